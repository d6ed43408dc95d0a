// Horizontal semi-coarsening multigrid for the FP32 inner solve (no
// counterpart in the reference, whose preconditioner is Jacobi-Richardson,
// wind_iterative.F90:1268-1315).
//
// Why this shape: the terrain-following grid has dz of 10 m at the ground and
// hundreds of metres aloft against dx = dy, so the levels of one column are
// coupled far more strongly than neighbouring columns near the ground and far
// more weakly aloft, and both ghost-plane rows are derivative (Neumann-type)
// conditions: the slowest error of a Krylov solve is smooth in x and y and
// constant in z, pinned only through the lateral identity rows. A vertical
// line smoother removes everything that varies within a column; halving the
// horizontal resolution level by level (down to one interior column, all
// levels kept) removes the smooth horizontal error at its own scale.
//
// Every level keeps the layout of the fine grid: rows [k][j][i], n = k P + c,
// a ring of lateral identity columns around nx x ny interior columns. Coarse
// column (I, J) covers the fine columns 2I-1, 2I and 2J-1, 2J (one of each at
// an odd edge). A coarse operator is the analytic operator rediscretised at
// 2 dx from the fine level's metrics averaged over these 2 x 2 columns (cell
// values over the four cells, face values over the two fine faces that make
// up the coarse face), assembled by the same set-up kernels as the fine
// operator. (Galerkin products with piecewise-constant transfer were tried
// first: on steep terrain the antisymmetric cross terms of the rows inside an
// aggregate fold into its vertical couplings, and the coarse columns lost the
// dominance the line smoother needs.) Restriction averages the rows of an
// aggregate (a coarse row is a per-cell equation like a fine one);
// prolongation interpolates bilinearly in x and y.
#pragma once

namespace windhip {

// One level of the hierarchy.
struct MgGeo {
    int nx, ny;   // interior columns
    int mx, my;   // columns including the lateral ring
    int mz;       // rows per column
    long P, N;    // columns, rows
};

inline MgGeo mg_geo(int nx, int ny, int mz)
{
    MgGeo m;
    m.nx = nx; m.ny = ny; m.mx = nx + 2; m.my = ny + 2; m.mz = mz;
    m.P = (long)m.mx * m.my;
    m.N = m.P * mz;
    return m;
}

constexpr int kMgThreads = 256;
inline int mg_blocks(long cols) { return (int)((cols + kMgThreads - 1) / kMgThreads); }

// Index set of fine columns under coarse column I of a level with nf
// interior columns: the ring maps to the ring.
__device__ __forceinline__ void mg_fine_range(int I, int nf, int nc, int &lo, int &hi)
{
    if (I <= 0) { lo = hi = 0; return; }
    if (I > nc) { lo = hi = nf + 1; return; }
    lo = 2 * I - 1;
    hi = min(2 * I, nf);
}
// fine face under coarse face I (the face left of coarse column I)
__device__ __forceinline__ int mg_fine_face(int I, int nf) { return I <= 0 ? 0 : min(2 * I - 1, nf + 1); }

// Inputs of the set-up chain at level c, averaged from level f: the cell
// fields over the fine cells of a coarse cell, u-face fields over the fine
// j-range at the fine face under the coarse face (v-faces alike), the terrain
// plane over the fine columns. One coarse row per thread (plane values from
// the threads of level 0).
struct MgMetricsIn {
    const double *dzdx, *dzdy, *jaco, *jaco_w, *alpha, *adv_dz, *dzdx_u, *jaco_u, *dzdy_v, *jaco_v, *terrain;
};
struct MgMetricsOut {
    double *dzdx, *dzdy, *jaco, *jaco_w, *alpha, *adv_dz, *dzdx_u, *jaco_u, *dzdy_v, *jaco_v, *terrain;
};
__global__ __launch_bounds__(kMgThreads) void k_mg_coarsen(MgGeo f, MgGeo c, MgMetricsIn a, MgMetricsOut o,
                                                           int alpha_only)
{
    const long n = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (n >= c.N) return;
    const int k = (int)(n / c.P);
    const long cc = n - (long)k * c.P;
    const int I = (int)(cc % c.mx), J = (int)(cc / c.mx);
    int i0, i1, j0, j1;
    mg_fine_range(I, f.nx, c.nx, i0, i1);
    mg_fine_range(J, f.ny, c.ny, j0, j1);
    const long fk = (long)k * f.P;
    auto cell = [&](const double *p) {
        double s = 0.0;
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) s += p[fk + (long)j * f.mx + i];
        return s / (double)((i1 - i0 + 1) * (j1 - j0 + 1));
    };
    o.alpha[n] = cell(a.alpha);
    if (alpha_only) return;
    o.dzdx[n] = cell(a.dzdx);
    o.dzdy[n] = cell(a.dzdy);
    o.jaco[n] = cell(a.jaco);
    o.jaco_w[n] = cell(a.jaco_w);
    o.adv_dz[n] = cell(a.adv_dz);
    const int fi = mg_fine_face(I, f.nx), fj = mg_fine_face(J, f.ny);
    double su = 0.0, sj = 0.0;
    for (int j = j0; j <= j1; ++j) {
        const long q = fk + (long)j * f.mx + fi;
        su += a.dzdx_u[q];
        sj += a.jaco_u[q];
    }
    o.dzdx_u[n] = su / (double)(j1 - j0 + 1);
    o.jaco_u[n] = sj / (double)(j1 - j0 + 1);
    double sv = 0.0, sk = 0.0;
    for (int i = i0; i <= i1; ++i) {
        const long q = fk + (long)fj * f.mx + i;
        sv += a.dzdy_v[q];
        sk += a.jaco_v[q];
    }
    o.dzdy_v[n] = sv / (double)(i1 - i0 + 1);
    o.jaco_v[n] = sk / (double)(i1 - i0 + 1);
    if (k == 0) {
        double st = 0.0;
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) st += a.terrain[(long)j * f.mx + i];
        o.terrain[cc] = st / (double)((i1 - i0 + 1) * (j1 - j0 + 1));
    }
}

// out = y + w L^-1 in on a level whose columns are too tall for the
// register-resident solve (y may be null; out must alias neither in nor y):
// the forward elimination is kept in out.
__global__ __launch_bounds__(kMgThreads) void k_mg_line(MgGeo g, const KState *st, const float *__restrict__ lo,
                                                        const float *__restrict__ cp, const float *__restrict__ ip,
                                                        const float *__restrict__ in, const float *__restrict__ y,
                                                        float *__restrict__ out, float w)
{
    if (st->done) return;
    const long c = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (c >= g.P) return;
    float dprev = 0.f;
    for (int k = 0; k < g.mz; ++k) {
        const long n = (long)k * g.P + c;
        dprev = (in[n] - lo[n] * dprev) * ip[n];
        out[n] = dprev;
    }
    float delta = 0.f;
    for (int k = g.mz - 1; k >= 0; --k) {
        const long n = (long)k * g.P + c;
        delta = out[n] - cp[n] * delta;
        out[n] = (y ? y[n] : 0.f) + w * delta;
    }
}

// b_c = R r_f, one coarse row per thread: every interior row is the mean of
// the rows of its aggregate; lateral rows get zero.
__global__ __launch_bounds__(kMgThreads) void k_mg_restrict(MgGeo f, MgGeo c, const KState *st,
                                                            const float *__restrict__ rf, float *__restrict__ bc)
{
    if (st->done) return;
    const long n = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (n >= c.N) return;
    const int k = (int)(n / c.P);
    const long cc = n - (long)k * c.P;
    const int I = (int)(cc % c.mx), J = (int)(cc / c.mx);
    if (I == 0 || I == c.mx - 1 || J == 0 || J == c.my - 1) {
        bc[n] = 0.f;
        return;
    }
    const int i0 = 2 * I - 1, j0 = 2 * J - 1;
    const long m = (long)k * f.P + (long)j0 * f.mx + i0;
    const bool hi = i0 + 1 <= f.nx, hj = j0 + 1 <= f.ny;
    float s = rf[m];
    if (hi) s += rf[m + 1];
    if (hj) {
        float t = rf[m + f.mx];
        if (hi) t += rf[m + f.mx + 1];
        s += t;
    }
    bc[n] = s / (float)((hi ? 2 : 1) * (hj ? 2 : 1));
}

// x_f += P e_c, one fine row per thread: cell-centred bilinear interpolation
// in x and y (a fine cell is the left or right, lower or upper half of its
// coarse cell: weights 9/16 on its own coarse cell, 3/16 on the two nearer
// neighbours, 1/16 on the diagonal one; ring values are zero), the level kept.
__global__ __launch_bounds__(kMgThreads) void k_mg_prolong(MgGeo f, MgGeo c, const KState *st,
                                                           const float *__restrict__ ec, float *__restrict__ xf)
{
    if (st->done) return;
    const long n = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (n >= f.N) return;
    const int k = (int)(n / f.P);
    const long fc = n - (long)k * f.P;
    const int i = (int)(fc % f.mx), j = (int)(fc / f.mx);
    if (i == 0 || i == f.mx - 1 || j == 0 || j == f.my - 1) return;
    const int I = (i + 1) >> 1, J = (j + 1) >> 1;
    const int In = (i & 1) ? I - 1 : I + 1, Jn = (j & 1) ? J - 1 : J + 1;
    const float *e = ec + (long)k * c.P;
    const float v = 0.5625f * e[(long)J * c.mx + I] + 0.1875f * (e[(long)J * c.mx + In] + e[(long)Jn * c.mx + I])
                  + 0.0625f * e[(long)Jn * c.mx + In];
    xf[n] += v;
}

}  // namespace windhip
