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
// Every level keeps the columns of the fine grid (a ring of lateral identity
// columns around nx x ny interior columns) and all its rows per column. The
// levels smoothed a wavefront per column (at most mg_wave_cols columns of up
// to 64 rows) store their vectors and their copies of the operator rows and
// line factors column-major, n = c S + k with S = mz, so the 64 lanes of a
// wavefront read one column's rows from consecutive addresses; the other
// levels keep the fine layout, rows [k][j][i], n = k P + c (MgIdx). Coarse
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

// Row n of (level k, column c): plane-major (S = 0, n = k P + c) or
// column-major (S = mz, n = c S + k); col(d) is the offset of the column d
// columns away at the same level.
struct MgIdx {
    long P, S;
    __device__ __forceinline__ long n(int k, long c) const { return S ? c * S + k : (long)k * P + c; }
    __device__ __forceinline__ long col(long d) const { return S ? d * S : d; }
};

// dst (column-major) = src (plane-major) over all N rows of a level
__global__ void k_mg_to_cm(long P, int mz, const float *__restrict__ src, float *__restrict__ dst)
{
    const long N = P * mz;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const long c = e / mz;
        const int k = (int)(e - c * mz);
        dst[e] = src[(long)k * P + c];
    }
}

constexpr int kMgThreads = 256;
inline int mg_blocks(long cols) { return (int)((cols + kMgThreads - 1) / kMgThreads); }
// the plane-by-level grid of the transfer kernels: kPlaneX x kPlaneY columns
// per block, kPlaneZ levels per thread (their loads issued together: one
// level per thread left the fine-level transfers waiting one memory round
// trip per 256 rows), one grid layer per kPlaneZ levels
constexpr int kPlaneX = 64, kPlaneY = kMgThreads / kPlaneX, kPlaneZ = 4;
inline dim3 mg_plane_grid(const MgGeo &g)
{
    return dim3((unsigned)((g.mx + kPlaneX - 1) / kPlaneX), (unsigned)((g.my + kPlaneY - 1) / kPlaneY),
                (unsigned)((g.mz + kPlaneZ - 1) / kPlaneZ));
}

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

// One damped line-relaxation step of a coarse level in one launch,
// out = y + w L^-1 (b - A y) (y null: out = w L^-1 b), for columns of at most
// 64 rows: one wavefront per column, lane k holding row k. On the small levels
// the two-kernel step (a residual march, then a Thomas march per column) is
// bound by the latency of its serial per-thread marches, not by bandwidth.
// Here every load is independent (one memory round trip): the residual row is
// evaluated from the lane's own five columns, with the k-1 and k+1 values
// taken from the neighbouring lanes, and both Thomas recurrences,
// d(k) = ip(k) (r(k) - lo(k) d(k-1)) and e(k) = d(k) - cp(k) e(k+1), are
// inclusive scans of affine maps z -> a z + c over the lanes (6 shuffle steps
// each way) on the factors of k_line_factor. y must be a level vector with
// mx + 1 guard elements on both sides (the neighbour loads of the ring
// columns fall on them); out must alias neither b nor y.
constexpr int kWaveCols = 4;  // columns (wavefronts) per block

// The damped line solve of one column held one row per lane: returns
// (L^-1 r)(k) for the lane's row k, with both Thomas recurrences as inclusive
// lane scans of affine maps on the factors lo, cp, ip of k_line_factor (lanes
// at or above mz hold zero maps).
__device__ __forceinline__ float wave_line_solve(int lane, bool act, float r, float lo, float cp, float ip)
{
    // forward: d(k) = a(k) d(k-1) + c(k), a = -ip lo, c = ip r, d(-1) = 0
    float a = act ? -(ip * lo) : 0.f, s = act ? ip * r : 0.f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const float pa = __shfl_up(a, off, 64), ps = __shfl_up(s, off, 64);
        if (lane >= off) {
            s = a * ps + s;
            a = a * pa;
        }
    }
    // backward: e(k) = -cp(k) e(k+1) + d(k), e(mz) = 0
    float a2 = act ? -cp : 0.f, s2 = act ? s : 0.f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const float pa = __shfl_down(a2, off, 64), ps = __shfl_down(s2, off, 64);
        if (lane + off < 64) {
            s2 = a2 * ps + s2;
            a2 = a2 * pa;
        }
    }
    return s2;
}

// Vector element access of the wavefront steps. COH: at the device-coherent
// level (agent-scope relaxed atomics), for vectors that one launch both
// writes and reads across blocks (k_mg_tail): no L2 write-back or invalidation
// is then needed between its steps; plain accesses otherwise.
template <bool COH>
__device__ __forceinline__ float vld(const float *p, long i)
{
    if constexpr (COH)
        return __hip_atomic_load(const_cast<float *>(p) + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return p[i];
}
template <bool COH>
__device__ __forceinline__ void vst(float *p, long i, float v)
{
    if constexpr (COH)
        __hip_atomic_store(p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        p[i] = v;
}

// A coarse level whose operator is the Galerkin product R A P of the next
// finer level's operator with the cycle's transfers (R the mean over an
// aggregate's columns, P the bilinear prolongation of k_mg_prolong): 27
// coefficients per row, for the 3 x 3 columns around the row's column at
// levels k-1, k, k+1. Stored column-major (row n = c mz + k, as the level's
// vectors) and planar: coefficient o = 9 (dk + 1) + 3 (dj + 1) + (di + 1) of
// row n at q[o N + n]. Lateral ring rows are identity rows; the ghost-plane
// rows hold only the offsets they have (zero elsewhere).
struct OpG {
    const float *q;
    long N;
};
__host__ __device__ constexpr int g_off(int di, int dk, int dj) { return 9 * (dk + 1) + 3 * (dj + 1) + (di + 1); }

// (A v)(row n) of a column held one row per lane: val(di, dj) is the lane's
// value at the column di, dj away (level k); the values at k-1 and k+1 come
// from the neighbouring lanes. The analytic operator reads the five columns
// C, W, E, S, N.
template <class Op, class Val>
__device__ __forceinline__ float wave_apply(const Op &op, int k, int mz, bool lat, long n, long c, Val val)
{
    Win<float> x;
    const float yc = val(0, 0), yw = val(-1, 0), ye = val(1, 0), ys = val(0, -1), yn = val(0, 1);
    x.c1 = yc; x.w1 = yw; x.e1 = ye; x.s1 = ys; x.n1 = yn;
    x.c0 = __shfl_up(yc, 1, 64); x.w0 = __shfl_up(yw, 1, 64); x.e0 = __shfl_up(ye, 1, 64);
    x.s0 = __shfl_up(ys, 1, 64); x.n0 = __shfl_up(yn, 1, 64);
    x.c2 = __shfl_down(yc, 1, 64); x.w2 = __shfl_down(yw, 1, 64); x.e2 = __shfl_down(ye, 1, 64);
    x.s2 = __shfl_down(ys, 1, 64); x.n2 = __shfl_down(yn, 1, 64);
    return op.row(k, mz, lat, x, op.load(n), op.bc(c));
}
// The Galerkin operator reads the nine columns around the row's column; its
// terms are summed in coefficient order.
template <class Val>
__device__ __forceinline__ float wave_apply(const OpG &op, int, int, bool lat, long n, long, Val val)
{
    float v[9], u[9], d[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) v[q] = val(q % 3 - 1, q / 3 - 1);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        u[q] = __shfl_up(v[q], 1, 64);
        d[q] = __shfl_down(v[q], 1, 64);
    }
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        const float xv = o < 9 ? u[o % 9] : (o < 18 ? v[o % 9] : d[o % 9]);
        s = s + op.q[o * op.N + n] * xv;
    }
    return lat ? v[4] : s;
}

// The residual row (b - A y)(row n) of a column held one row per lane.
template <class Op, bool COH = false>
__device__ __forceinline__ float wave_resid(const Op &op, const float *__restrict__ y, float bv, long n, long c,
                                            int k, int mz, bool lat, long dx, long dy)
{
    return bv - wave_apply(op, k, mz, lat, n, c, [&](int di, int dj) { return vld<COH>(y, n + di * dx + dj * dy); });
}

// Rows of a finer level as 27-point rows for the Galerkin product: a[o]
// (g_off numbering) of row (i, k, j), zero where the row has no such entry.
// A rediscretised level: its FP32 analytic operator (interior 15-point rows
// J = -H, K = -I, N = -L, O = -M, the bottom BC row, the top row; the
// wind_iterative.F90:912-1025 stencil), plane-major.
struct FineRedisc {
    OpA<float> op;
    long P;
    int mx, my, mz;
    __device__ void row(int i, int k, int j, double (&a)[27]) const
    {
#pragma unroll
        for (int o = 0; o < 27; ++o) a[o] = 0.0;
        const long c = (long)j * mx + i;
        if (i == 0 || i == mx - 1 || j == 0 || j == my - 1) { a[g_off(0, 0, 0)] = 1.0; return; }
        const typename OpA<float>::Bc b = op.bc(c);
        if (k == 0) {
            a[g_off(0, 0, -1)] = b.by; a[g_off(0, 1, -1)] = b.by;
            a[g_off(-1, 0, 0)] = b.bx; a[g_off(-1, 1, 0)] = b.bx;
            a[g_off(0, 0, 0)] = -b.bz; a[g_off(0, 1, 0)] = b.bz;
            a[g_off(1, 0, 0)] = -b.bx; a[g_off(1, 1, 0)] = -b.bx;
            a[g_off(0, 0, 1)] = -b.by; a[g_off(0, 1, 1)] = -b.by;
            return;
        }
        if (k == mz - 1) {
            a[g_off(0, 0, 0)] = 1.0 / (double)b.dzt;
            a[g_off(0, -1, 0)] = -1.0 / (double)b.dzt;
            return;
        }
        const typename OpA<float>::Cf f = op.load((long)k * P + c);
        const double inv = op.inv;
        a[g_off(0, 0, 0)] = ((double)op.neg4 - f.B) - f.C;
        a[g_off(0, 1, 0)] = f.B; a[g_off(0, -1, 0)] = f.C;
        a[g_off(1, 0, 0)] = inv; a[g_off(-1, 0, 0)] = inv; a[g_off(0, 0, 1)] = inv; a[g_off(0, 0, -1)] = inv;
        a[g_off(1, 1, 0)] = f.H; a[g_off(-1, 1, 0)] = f.I; a[g_off(1, -1, 0)] = -f.H; a[g_off(-1, -1, 0)] = -f.I;
        a[g_off(0, 1, 1)] = f.L; a[g_off(0, 1, -1)] = f.M; a[g_off(0, -1, 1)] = -f.L; a[g_off(0, -1, -1)] = -f.M;
    }
};
// A Galerkin level (column-major 27-point rows).
struct FineGal {
    OpG op;
    int mx, my, mz;
    __device__ void row(int i, int k, int j, double (&a)[27]) const
    {
        const long n = ((long)j * mx + i) * mz + k;
#pragma unroll
        for (int o = 0; o < 27; ++o) a[o] = op.q[o * op.N + n];
    }
};

// The Galerkin product A_c = R A_f P, one coarse row per thread, column-major
// output (OpG layout). R: the mean over the (up to) 2 x 2 fine columns 2I-1,
// 2I x 2J-1, 2J of coarse column (I, J) at the same level; P: the bilinear
// prolongation of k_mg_prolong (a fine column takes 9/16 of its own coarse
// cell, 3/16 of the two nearer neighbours, 1/16 of the diagonal one; ring
// columns, fine or coarse, carry nothing). Every index offset is a compile-time
// constant after unrolling: fine column 2I-1+t (t = aggregate position +
// fine offset, -1..2) has its own coarse cell at I + (t == -1 ? -1 : t == 2)
// and its neighbouring cell at I + (t == -1 || t == 2 ? 0 : t == 0 ? -1 : 1).
// The sums run in FP64; the coefficients are stored in FP32. Ring rows of the
// coarse level are identity rows.
template <class Fine>
__global__ __launch_bounds__(kMgThreads) void k_mg_rap(Fine fine, MgGeo f, MgGeo c, float *__restrict__ q)
{
    const long r = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (r >= c.N) return;
    const int mz = c.mz;
    const long cc = r / mz;
    const int k = (int)(r - cc * mz);
    const int I = (int)(cc % c.mx), J = (int)(cc / c.mx);
    double acc[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) acc[o] = 0.0;
    if (I == 0 || I == c.mx - 1 || J == 0 || J == c.my - 1) {
        acc[g_off(0, 0, 0)] = 1.0;
    } else {
        const int ni = 2 * I <= f.nx ? 2 : 1, nj = 2 * J <= f.ny ? 2 : 1;
        const double rw = 1.0 / (double)(ni * nj);
#pragma unroll
        for (int aj = 0; aj < 2; ++aj)
#pragma unroll
            for (int ai = 0; ai < 2; ++ai) {
                if (ai >= ni || aj >= nj) continue;
                const int i = 2 * I - 1 + ai, j = 2 * J - 1 + aj;
                double a[27];
                fine.row(i, k, j, a);
#pragma unroll
                for (int o = 0; o < 27; ++o) {
                    const int di = o % 3 - 1, dj = (o / 3) % 3 - 1, dk = o / 9 - 1;
                    const double v = a[o];
                    if (v == 0.0) continue;
                    const int it = i + di, jt = j + dj;
                    if (it <= 0 || it >= f.mx - 1 || jt <= 0 || jt >= f.my - 1) continue;  // fine ring: P row zero
                    const int tx = ai + di, ty = aj + dj;  // -1..2
                    const int ox = tx == -1 ? -1 : (tx == 2 ? 1 : 0), nxo = tx == -1 || tx == 2 ? 0 : (tx == 0 ? -1 : 1);
                    const int oy = ty == -1 ? -1 : (ty == 2 ? 1 : 0), nyo = ty == -1 || ty == 2 ? 0 : (ty == 0 ? -1 : 1);
                    const double rv = rw * v;
#pragma unroll
                    for (int pq = 0; pq < 4; ++pq) {
                        const int cx = (pq & 1) ? nxo : ox, cy = (pq & 2) ? nyo : oy;
                        const int It = I + cx, Jt = J + cy;
                        if (It <= 0 || It >= c.mx - 1 || Jt <= 0 || Jt >= c.my - 1) continue;  // coarse ring: zero
                        const double w = ((pq & 1) ? 0.25 : 0.75) * ((pq & 2) ? 0.25 : 0.75);
                        acc[g_off(cx, dk, cy)] += rv * w;
                    }
                }
            }
    }
#pragma unroll
    for (int o = 0; o < 27; ++o) q[o * c.N + r] = (float)acc[o];
}

// Line factors of a Galerkin level (column-major): the Thomas factors of each
// column's tridiagonal part (coefficients of k-1, k, k+1 in the own column),
// in FP64 as k_line_factor forms them, stored in FP32: lo = lower coefficient,
// ip = 1 / pivot, cp = upper / pivot.
__global__ __launch_bounds__(kMgThreads) void k_mg_factor_g(MgGeo g, OpG op, float *__restrict__ lo,
                                                            float *__restrict__ cp, float *__restrict__ ip)
{
    const long c = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (c >= g.P) return;
    double cprev = 0.0;
    for (int k = 0; k < g.mz; ++k) {
        const long n = c * g.mz + k;
        const double l = op.q[g_off(0, -1, 0) * op.N + n], dd = op.q[g_off(0, 0, 0) * op.N + n],
                     u = op.q[g_off(0, 1, 0) * op.N + n];
        const double m = dd - l * cprev;
        cprev = u / m;
        lo[n] = (float)l;
        ip[n] = (float)(1.0 / m);
        cp[n] = (float)cprev;
    }
}

// One damped line-smoothing step of column c of a level, lane k holding row
// k: out = y + w L^-1 (b - A y) (y null: out = w L^-1 b).
template <class Op, bool CM, bool COH = false>
__device__ __forceinline__ void wave_smooth_col(const Op &op, const Geo &g, const float *__restrict__ b,
                                                const float *__restrict__ y, const float *__restrict__ lo,
                                                const float *__restrict__ cp, const float *__restrict__ ip,
                                                float *__restrict__ out, float w, long c, int lane)
{
    const int mz = g.mz;
    const bool act = lane < mz;
    const int k = act ? lane : mz - 1;
    const MgIdx ix{g.P, CM ? mz : 0};
    const long n = ix.n(k, c), dx = ix.col(1), dy = ix.col(g.mx);
    const int i = (int)(c % g.mx), j = (int)(c / g.mx);
    const bool lat = i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1;
    float r = vld<COH>(b, n);
    float yc = 0.f;
    if (y) {
        yc = vld<COH>(y, n);
        r = r - wave_apply(op, k, mz, lat, n, c, [&](int di, int dj) {
                return di == 0 && dj == 0 ? yc : vld<COH>(y, n + di * dx + dj * dy);
            });
    }
    const float s2 = wave_line_solve(lane, act, r, lo[n], cp[n], ip[n]);
    if (act) vst<COH>(out, n, yc + w * s2);
}

template <class Op, bool CM = false>
__global__ __launch_bounds__(64 * kWaveCols) void k_mg_smooth_w(Op op, Geo g, const KState *st,
                                                               const float *__restrict__ b,
                                                               const float *__restrict__ y,
                                                               const float *__restrict__ lo,
                                                               const float *__restrict__ cp,
                                                               const float *__restrict__ ip,
                                                               float *__restrict__ out, float w)
{
    if (st->done) return;
    const long c = (long)blockIdx.x * kWaveCols + (threadIdx.x >> 6);
    if (c >= g.P) return;  // uniform per wavefront
    wave_smooth_col<Op, CM>(op, g, b, y, lo, cp, ip, out, w, c, threadIdx.x & 63);
}

// The descent into a coarse level in one launch, for coarse levels of at most
// 64 rows per column: the residual r_f = b_f - A_f y_f of the finer level at
// the (up to) 2 x 2 fine columns of each coarse column, their mean as the
// coarse rhs b_c (the restriction of k_mg_restrict, in its order), and the
// first damped line-smoothing step of the coarse level from zero,
// x_c = w L_c^-1 b_c. One wavefront per coarse column, lane k holding row k.
// y_f needs mx_f + 1 guard elements on both sides (the neighbour loads of
// the finer level's ring columns fall on them). Ring columns of the coarse
// level get b_c = x_c = 0. wave_down_col: coarse column cc.
template <class Op, bool CM, bool COH = false>
__device__ __forceinline__ void wave_down_col(const Op &opf, const Geo &gf, const MgGeo &c,
                                              const float *__restrict__ yf, const float *__restrict__ bf,
                                              const float *__restrict__ lo, const float *__restrict__ cp,
                                              const float *__restrict__ ip, float *__restrict__ bc,
                                              float *__restrict__ xc, float w, long cc, int lane)
{
    const int mz = c.mz;
    const bool act = lane < mz;
    const int k = act ? lane : mz - 1;
    const MgIdx ic{c.P, CM ? mz : 0}, iff{gf.P, CM ? mz : 0};
    const long n = ic.n(k, cc);
    const int I = (int)(cc % c.mx), J = (int)(cc / c.mx);
    if (I == 0 || I == c.mx - 1 || J == 0 || J == c.my - 1) {
        if (act) { vst<COH>(bc, n, 0.f); vst<COH>(xc, n, 0.f); }
        return;
    }
    const long mxf = gf.mx;
    const int i0 = 2 * I - 1, j0 = 2 * J - 1;
    const bool hi = i0 + 1 <= (int)mxf - 2, hj = j0 + 1 <= gf.my - 2;
    const long cf0 = (long)j0 * mxf + i0, m = iff.n(k, cf0), dx = iff.col(1), dy = iff.col(mxf);
    // the fine columns are interior columns of the finer level
    float s = wave_resid<Op, COH>(opf, yf, vld<COH>(bf, m), m, cf0, k, mz, false, dx, dy);
    if (hi) s += wave_resid<Op, COH>(opf, yf, vld<COH>(bf, m + dx), m + dx, cf0 + 1, k, mz, false, dx, dy);
    if (hj) {
        float t = wave_resid<Op, COH>(opf, yf, vld<COH>(bf, m + dy), m + dy, cf0 + mxf, k, mz, false, dx, dy);
        if (hi)
            t += wave_resid<Op, COH>(opf, yf, vld<COH>(bf, m + dy + dx), m + dy + dx, cf0 + mxf + 1, k, mz, false,
                                     dx, dy);
        s += t;
    }
    const float b = s / (float)((hi ? 2 : 1) * (hj ? 2 : 1));
    const float s2 = wave_line_solve(lane, act, b, lo[n], cp[n], ip[n]);
    if (act) {
        vst<COH>(bc, n, b);
        vst<COH>(xc, n, w * s2);
    }
}

template <class Op, bool CM = false>
__global__ __launch_bounds__(64 * kWaveCols) void k_mg_down_w(Op opf, Geo gf, MgGeo c, const KState *st,
                                                             const float *__restrict__ yf,
                                                             const float *__restrict__ bf,
                                                             const float *__restrict__ lo,
                                                             const float *__restrict__ cp,
                                                             const float *__restrict__ ip,
                                                             float *__restrict__ bc, float *__restrict__ xc, float w)
{
    if (st->done) return;
    const long cc = (long)blockIdx.x * kWaveCols + (threadIdx.x >> 6);
    if (cc >= c.P) return;  // uniform per wavefront
    wave_down_col<Op, CM>(opf, gf, c, yf, bf, lo, cp, ip, bc, xc, w, cc, threadIdx.x & 63);
}

// The return to a finer level in one launch, for levels of at most 64 rows per
// column: z = y + P e (the bilinear prolongation of k_mg_prolong of the coarse
// correction e) at the column and its four neighbours, then one damped
// line-smoothing step out = z + w L^-1 (b - A z). One wavefront per column of
// the finer level, lane k holding row k. y needs mx + 1 guard elements on
// both sides; out must alias neither y nor b. wave_up_col: column col.
template <class Op, bool CM, bool COH = false>
__device__ __forceinline__ void wave_up_col(const Op &op, const Geo &g, const MgGeo &c, const float *__restrict__ y,
                                            const float *__restrict__ ec, const float *__restrict__ b,
                                            const float *__restrict__ lo, const float *__restrict__ cp,
                                            const float *__restrict__ ip, float *__restrict__ out, float w, long col,
                                            int lane)
{
    const int mz = g.mz;
    const bool act = lane < mz;
    const int k = act ? lane : mz - 1;
    const long mx = g.mx;
    const MgIdx ix{g.P, CM ? mz : 0}, ic{c.P, CM ? mz : 0};
    const long n = ix.n(k, col);
    const int i = (int)(col % mx), j = (int)(col / mx);
    const bool lat = i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1;
    // y + P e at fine column (ii, jj) of row k; ring columns are not prolonged
    auto z = [&](int ii, int jj) -> float {
        const float yv = vld<COH>(y, ix.n(k, (long)jj * mx + ii));
        if (ii <= 0 || ii >= g.mx - 1 || jj <= 0 || jj >= g.my - 1) return yv;
        const int I = (ii + 1) >> 1, J = (jj + 1) >> 1;
        const int In = (ii & 1) ? I - 1 : I + 1, Jn = (jj & 1) ? J - 1 : J + 1;
        const float v = 0.5625f * vld<COH>(ec, ic.n(k, (long)J * c.mx + I)) +
                        0.1875f * (vld<COH>(ec, ic.n(k, (long)J * c.mx + In)) +
                                   vld<COH>(ec, ic.n(k, (long)Jn * c.mx + I))) +
                        0.0625f * vld<COH>(ec, ic.n(k, (long)Jn * c.mx + In));
        return yv + v;
    };
    // the ring's neighbours outside the level are never read by its rows
    const float zc = z(i, j);
    const float r = vld<COH>(b, n) - wave_apply(op, k, mz, lat, n, col, [&](int di, int dj) {
                        const int ii = i + di, jj = j + dj;
                        if (di == 0 && dj == 0) return zc;
                        return ii >= 0 && ii < g.mx && jj >= 0 && jj < g.my ? z(ii, jj) : zc;
                    });
    const float s2 = wave_line_solve(lane, act, r, lo[n], cp[n], ip[n]);
    if (act) vst<COH>(out, n, zc + w * s2);
}

template <class Op, bool CM = false>
__global__ __launch_bounds__(64 * kWaveCols) void k_mg_up_w(Op op, Geo g, MgGeo c, const KState *st,
                                                           const float *__restrict__ y,
                                                           const float *__restrict__ ec,
                                                           const float *__restrict__ b,
                                                           const float *__restrict__ lo,
                                                           const float *__restrict__ cp,
                                                           const float *__restrict__ ip,
                                                           float *__restrict__ out, float w)
{
    if (st->done) return;
    const long col = (long)blockIdx.x * kWaveCols + (threadIdx.x >> 6);
    if (col >= g.P) return;  // uniform per wavefront
    wave_up_col<Op, CM>(op, g, c, y, ec, b, lo, cp, ip, out, w, col, threadIdx.x & 63);
}

// The cycle below a column-major level as one persistent launch: the same
// steps as the k_mg_smooth_w / k_mg_down_w / k_mg_up_w launches of that part
// of the V-cycle, each a loop of the grid's wavefronts over the level's
// columns, separated by grid-wide barriers instead of kernel boundaries (the
// levels of at most 66 x 66 columns hold about a microsecond of work per
// step). Per column the arithmetic is that of the separate kernels, so the
// result is bitwise the same. The vectors the steps pass to each other are
// read and written at the device-coherent level (vld/vst<true>), so a barrier
// needs no L2 write-back or invalidation; the rows and factors, constant
// during the launch, stay cached. Measured at 512 x 512 x 40 (20 steps, 545
// blocks): 130-145 us per launch against 115 us for the 20 separate launches
// (a barrier costs what a kernel boundary does, 5-7 us), so the solve runs
// the separate launches unless asked for this form.
struct MgTailLevel {
    OpA<float> op;  // column-major FP32 rows
    Geo g;
    MgGeo m;
    const float *lo, *cp, *ip;  // column-major line factors
    float *b;                   // rhs (column-major)
};
enum { kTailSmooth = 0, kTailDown = 1, kTailUp = 2 };
struct MgTailStep {
    int kind;      // kTailSmooth: step on `level`; kTailDown / kTailUp: transfer between level and level + 1
    int level;
    float w;       // damping of the step's line solve
    const float *y;  // iterate on `level` (kTailSmooth: null = from zero)
    const float *e;  // kTailUp: the correction on level + 1
    float *out;      // kTailSmooth / kTailUp: new iterate on `level`; kTailDown: first iterate on level + 1
};
constexpr int kTailMaxSteps = 64, kTailMaxLevels = 12, kTailMaxBlocks = 4096;
struct MgTailPlan {
    int nsteps, nlevels;
    MgTailStep step[kTailMaxSteps];
    MgTailLevel lv[kTailMaxLevels];
};
// Grid barrier state (zeroed once). Barrier q of a launch uses the epoch
// E = G + q, G the value of gen at the launch's start: every block but block
// 0 writes E to its flag and waits for gen to reach E; block 0 waits for all
// flags to read E and then writes E to the eight copies of gen (one cache
// line each, polled by an eighth of the blocks). err is sticky: a wait that
// ran out sets it, every later launch returns at once, and the host reports
// it (and clears the state).
struct MgGridBar {
    unsigned gen[8][32];
    int err;
    unsigned pad[31];
    unsigned flag[kTailMaxBlocks];
};
constexpr int kTailWaves = 8;                   // wavefronts (columns in flight) per block
constexpr unsigned kTailSpinLimit = 1u << 20;   // polls of one wait (about a second) before it gives up

__device__ __forceinline__ unsigned ld_coh_u(const unsigned *p)
{
    return __hip_atomic_load(const_cast<unsigned *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh_u(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// All blocks of the grid (resident together: the host sizes the grid within
// the device's occupancy) meet here; the coherent stores before it are seen
// by the coherent loads after it in every block. Returns false when a wait
// ran out (here or in another block).
__device__ __forceinline__ bool grid_barrier(MgGridBar *bar, unsigned nblocks, unsigned epoch)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // this wavefront's stores have landed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (blockIdx.x == 0) {
        __syncthreads();
        for (unsigned spins = 0;; ++spins) {
            bool mine = true;
            for (unsigned q = threadIdx.x + 1; q < nblocks; q += blockDim.x) mine &= ld_coh_u(&bar->flag[q]) == epoch;
            if (__syncthreads_and(mine)) break;
            if ((spins & 255u) == 255u) {
                const bool bad = spins > kTailSpinLimit || __hip_atomic_load(&bar->err, __ATOMIC_RELAXED,
                                                                             __HIP_MEMORY_SCOPE_AGENT) != 0;
                if (__syncthreads_or(bad)) {
                    if (threadIdx.x == 0) __hip_atomic_store(&bar->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return false;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (threadIdx.x < 8) st_coh_u(&bar->gen[threadIdx.x][0], epoch);
        return true;
    }
    __shared__ int ok_s;
    __syncthreads();  // every wavefront's stores have landed
    if (threadIdx.x == 0) {
        int ok = 1;
        st_coh_u(&bar->flag[blockIdx.x], epoch);
        const unsigned *gp = &bar->gen[blockIdx.x & 7][0];
        unsigned spins = 0;
        while (ld_coh_u(gp) != epoch) {
            if (++spins > kTailSpinLimit ||
                __hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                __hip_atomic_store(&bar->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        ok_s = ok;
    }
    __syncthreads();
    return ok_s != 0;
}

__global__ __launch_bounds__(64 * kTailWaves) void k_mg_tail(const MgTailPlan *__restrict__ plan, const KState *st,
                                                             MgGridBar *bar)
{
    if (st->done) return;  // set before this launch: every block returns
    if (__hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const unsigned base = ld_coh_u(&bar->gen[blockIdx.x & 7][0]);  // the same in every copy between launches
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * kTailWaves + (threadIdx.x >> 6);
    const long nwaves = (long)gridDim.x * kTailWaves;
    const int ns = plan->nsteps;
    for (int q = 0; q < ns; ++q) {
        if (q > 0 && !grid_barrier(bar, gridDim.x, base + (unsigned)q)) return;
        const MgTailStep &t = plan->step[q];
        const MgTailLevel &L = plan->lv[t.level];
        if (t.kind == kTailSmooth) {
            for (long c = wave; c < L.g.P; c += nwaves)
                wave_smooth_col<OpA<float>, true, true>(L.op, L.g, L.b, t.y, L.lo, L.cp, L.ip, t.out, t.w, c, lane);
        } else if (t.kind == kTailDown) {
            const MgTailLevel &C = plan->lv[t.level + 1];
            for (long c = wave; c < C.m.P; c += nwaves)
                wave_down_col<OpA<float>, true, true>(L.op, L.g, C.m, t.y, L.b, C.lo, C.cp, C.ip, C.b, t.out, t.w, c,
                                                      lane);
        } else {
            const MgTailLevel &C = plan->lv[t.level + 1];
            for (long c = wave; c < L.g.P; c += nwaves)
                wave_up_col<OpA<float>, true, true>(L.op, L.g, C.m, t.y, t.e, L.b, L.lo, L.cp, L.ip, t.out, t.w, c,
                                                    lane);
        }
    }
}

// b_c = R r_f, one coarse row per thread: every interior row is the mean of
// the rows of its aggregate; lateral rows get zero.
__global__ __launch_bounds__(kMgThreads) void k_mg_restrict(MgGeo f, MgGeo c, const KState *st,
                                                            const float *__restrict__ rf, float *__restrict__ bc,
                                                            long fS, long cS)
{
    const MgIdx fi{f.P, fS}, ci{c.P, cS};
    if (st->done) return;
    // a 64 x 4 block of (I, J) per grid cell (x, y), the level as grid z: no
    // index divisions
    const int I = (int)(blockIdx.x * kPlaneX + (threadIdx.x & (kPlaneX - 1)));
    const int J = (int)(blockIdx.y * kPlaneY + (threadIdx.x / kPlaneX));
    const int k0 = (int)blockIdx.z * kPlaneZ;
    if (I >= c.mx || J >= c.my) return;
    const long n0 = (long)J * c.mx + I;
    if (I == 0 || I == c.mx - 1 || J == 0 || J == c.my - 1) {
#pragma unroll
        for (int q = 0; q < kPlaneZ; ++q)
            if (k0 + q < c.mz) bc[ci.n(k0 + q, n0)] = 0.f;
        return;
    }
    const int i0 = 2 * I - 1, j0 = 2 * J - 1;
    const bool hi = i0 + 1 <= f.nx, hj = j0 + 1 <= f.ny;
    const long cf0 = (long)j0 * f.mx + i0, dx = fi.col(1), dy = fi.col(f.mx);
    float sum[kPlaneZ];
#pragma unroll
    for (int q = 0; q < kPlaneZ; ++q) {
        const int k = k0 + q < c.mz ? k0 + q : c.mz - 1;  // levels past the top re-read the top one
        const long m = fi.n(k, cf0);
        float s = rf[m];
        if (hi) s += rf[m + dx];
        if (hj) {
            float t = rf[m + dy];
            if (hi) t += rf[m + dy + dx];
            s += t;
        }
        sum[q] = s;
    }
#pragma unroll
    for (int q = 0; q < kPlaneZ; ++q)
        if (k0 + q < c.mz) bc[ci.n(k0 + q, n0)] = sum[q] / (float)((hi ? 2 : 1) * (hj ? 2 : 1));
}

// x_f += P e_c, one fine row per thread: cell-centred bilinear interpolation
// in x and y (a fine cell is the left or right, lower or upper half of its
// coarse cell: weights 9/16 on its own coarse cell, 3/16 on the two nearer
// neighbours, 1/16 on the diagonal one; ring values are zero), the level kept.
__global__ __launch_bounds__(kMgThreads) void k_mg_prolong(MgGeo f, MgGeo c, const KState *st,
                                                           const float *__restrict__ ec, float *__restrict__ xf,
                                                           long fS, long cS)
{
    const MgIdx fi{f.P, fS}, ci{c.P, cS};
    if (st->done) return;
    const int i = (int)(blockIdx.x * kPlaneX + (threadIdx.x & (kPlaneX - 1)));
    const int j = (int)(blockIdx.y * kPlaneY + (threadIdx.x / kPlaneX));
    const int k0 = (int)blockIdx.z * kPlaneZ;
    if (i <= 0 || i >= f.mx - 1 || j <= 0 || j >= f.my - 1) return;
    const int I = (i + 1) >> 1, J = (j + 1) >> 1;
    const int In = (i & 1) ? I - 1 : I + 1, Jn = (j & 1) ? J - 1 : J + 1;
    const long cc = (long)J * c.mx + I, cx = (long)J * c.mx + In, cy = (long)Jn * c.mx + I, cd = (long)Jn * c.mx + In;
    const long nf = (long)j * f.mx + i;
    float v[kPlaneZ], xo[kPlaneZ];
#pragma unroll
    for (int q = 0; q < kPlaneZ; ++q) {
        const int k = k0 + q < f.mz ? k0 + q : f.mz - 1;  // levels past the top re-read the top one
        v[q] = 0.5625f * ec[ci.n(k, cc)] + 0.1875f * (ec[ci.n(k, cx)] + ec[ci.n(k, cy)]) + 0.0625f * ec[ci.n(k, cd)];
        xo[q] = xf[fi.n(k, nf)];
    }
#pragma unroll
    for (int q = 0; q < kPlaneZ; ++q)
        if (k0 + q < f.mz) xf[fi.n(k0 + q, nf)] = xo[q] + v[q];
}

// The residual r = b - A y of a level of one wavefront per column (the
// separate-kernel form of the descent on those levels), lane k holding row k.
template <class Op, bool CM = false>
__global__ __launch_bounds__(64 * kWaveCols) void k_mg_resid_w(Op op, Geo g, const KState *st,
                                                              const float *__restrict__ y,
                                                              const float *__restrict__ b, float *__restrict__ r)
{
    if (st->done) return;
    const int lane = threadIdx.x & 63;
    const long c = (long)blockIdx.x * kWaveCols + (threadIdx.x >> 6);
    if (c >= g.P) return;  // uniform per wavefront
    const int mz = g.mz;
    const bool act = lane < mz;
    const int k = act ? lane : mz - 1;
    const MgIdx ix{g.P, CM ? mz : 0};
    const long n = ix.n(k, c);
    const int i = (int)(c % g.mx), j = (int)(c / g.mx);
    const bool lat = i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1;
    const float rv = wave_resid(op, y, b[n], n, c, k, mz, lat, ix.col(1), ix.col(g.mx));
    if (act) r[n] = rv;
}


// The fine level's descent in one pass: the residual r = b - A y of the 2 x 2
// fine columns under each coarse column (I, J) and their mean as the coarse
// rhs (k_resid_c followed by k_mg_restrict, the same arithmetic element by
// element, without the residual's round trip through memory). One thread per
// coarse column and level chunk (kRRChunks chunks per column, 256 coarse
// columns per block), marching up with a three-level window of the twelve
// fine columns the four rows read:
//   row j0 - 1:        S0 (i0), S1 (i0 + 1)
//   row j0:      W0 (i0 - 1), C00, C10, E0 (i0 + 2)
//   row j0 + 1:  W1 (i0 - 1), C01, C11, E1 (i0 + 2)
//   row j0 + 2:        N0 (i0), N1 (i0 + 1)
// with i0 = 2I - 1, j0 = 2J - 1. y needs mx + 1 guard elements on both sides
// (the neighbour loads of the ring's fine columns fall on them); the fine
// level is plane-major, the coarse one of stride cS (MgIdx). Ring columns of
// the coarse level get zero.
constexpr int kRRChunks = 6;  // level chunks per coarse column of k_resid_restrict_c
template <class Op>
__global__ __launch_bounds__(kColThreads) void k_resid_restrict_c(Op op, Geo g, MgGeo c, const KState *st,
                                                                  const float *__restrict__ y,
                                                                  const float *__restrict__ in,
                                                                  float *__restrict__ bc, long cS)
{
    if (st->done) return;
    const int nb = gridDim.x, L = logical_block(nb);
    const int cbk = L / kRRChunks, q = L - cbk * kRRChunks;
    const long cc = (long)cbk * kColThreads + threadIdx.x;
    if (cc >= c.P) return;
    const int mz = __builtin_amdgcn_readfirstlane(g.mz);
    const int k0 = (int)((long)q * mz / kRRChunks), k1 = (int)((long)(q + 1) * mz / kRRChunks);
    const MgIdx ci{c.P, cS};
    const int I = (int)(cc % c.mx), J = (int)(cc / c.mx);
    if (I == 0 || I == c.mx - 1 || J == 0 || J == c.my - 1) {
        for (int k = k0; k < k1; ++k) bc[ci.n(k, cc)] = 0.f;
        return;
    }
    const int mx = __builtin_amdgcn_readfirstlane(g.mx);
    const long P = (long)mx * __builtin_amdgcn_readfirstlane(g.my);
    const long N = P * mz;
    const int i0 = 2 * I - 1, j0 = 2 * J - 1;
    const bool hi = i0 + 1 <= mx - 2, hj = j0 + 1 <= g.my - 2;
    const long c00 = (long)j0 * mx + i0;
    const unsigned oc = (unsigned)c00, voc = oc * 4u, pb = (unsigned)(P * 4);
    const long gd = mx + 1;
    // byte offsets (uniform) of the twelve columns from (y - gd) + c00, and of
    // the four fine columns from c00
    auto yo = [&](int di, int dj) { return (unsigned)((gd + (long)dj * mx + di) * 4); };
    const unsigned oW0 = yo(-1, 0), oC00 = yo(0, 0), oC10 = yo(1, 0), oE0 = yo(2, 0), oS0 = yo(0, -1),
                   oS1 = yo(1, -1), oW1 = yo(-1, 1), oC01 = yo(0, 1), oC11 = yo(1, 1), oE1 = yo(2, 1),
                   oN0 = yo(0, 2), oN1 = yo(1, 2);
    const long r10 = 1, r01 = mx, r11 = mx + 1;  // fine column offsets of the other three from c00
    const rsrc_t ry = make_rsrc(y - gd, (N + 2 * gd) * 4), rb = make_rsrc(in, N * 4);
    const typename Op::Src rcf = op.src(N);
    using Cf = typename Op::Cf;
    const typename Op::Bc b00 = op.bc(c00), b10 = op.bc(c00 + r10), b01 = op.bc(c00 + r01),
                          b11 = op.bc(c00 + r11);
    // the twelve columns at one level
    struct Lv { float W0, C00, C10, E0, S0, S1, W1, C01, C11, E1, N0, N1; };
    auto ld = [&](Lv &v, int k) {
        const unsigned so = (unsigned)k * pb;
        v.W0 = bload<float>(ry, voc, so + oW0); v.C00 = bload<float>(ry, voc, so + oC00);
        v.C10 = bload<float>(ry, voc, so + oC10); v.E0 = bload<float>(ry, voc, so + oE0);
        v.S0 = bload<float>(ry, voc, so + oS0); v.S1 = bload<float>(ry, voc, so + oS1);
        v.W1 = bload<float>(ry, voc, so + oW1); v.C01 = bload<float>(ry, voc, so + oC01);
        v.C11 = bload<float>(ry, voc, so + oC11); v.E1 = bload<float>(ry, voc, so + oE1);
        v.N0 = bload<float>(ry, voc, so + oN0); v.N1 = bload<float>(ry, voc, so + oN1);
    };
    // per level: b and the coefficient row of the four fine columns
    struct Rw { float b00, b10, b01, b11; Cf f00, f10, f01, f11; };
    auto ldr = [&](Rw &w, int k) {
        const unsigned so = (unsigned)k * pb;
        const long lo = (long)k * P;
        w.b00 = bload<float>(rb, voc, so); w.b10 = bload<float>(rb, voc, so + (unsigned)(r10 * 4));
        w.b01 = bload<float>(rb, voc, so + (unsigned)(r01 * 4));
        w.b11 = bload<float>(rb, voc, so + (unsigned)(r11 * 4));
        w.f00 = op.load_src(rcf, lo, so, oc, voc);
        w.f10 = op.load_src(rcf, lo + r10, so + (unsigned)(r10 * 4), oc, voc);
        w.f01 = op.load_src(rcf, lo + r01, so + (unsigned)(r01 * 4), oc, voc);
        w.f11 = op.load_src(rcf, lo + r11, so + (unsigned)(r11 * 4), oc, voc);
    };
    // the 5-column window of one fine column from the level windows a (k-1), b (k), c (k+1)
    auto win = [](float c0, float c1, float c2, float w0, float w1, float w2, float e0, float e1, float e2, float s0,
                  float s1, float s2, float n0, float n1, float n2) {
        Win<float> x;
        x.c0 = c0; x.c1 = c1; x.c2 = c2; x.w0 = w0; x.w1 = w1; x.w2 = w2; x.e0 = e0; x.e1 = e1; x.e2 = e2;
        x.s0 = s0; x.s1 = s1; x.s2 = s2; x.n0 = n0; x.n1 = n1; x.n2 = n2;
        return x;
    };
    Lv va, vb, vc;  // levels k-1, k, k+1
    Rw rw, rn;
    ld(vb, k0);
    if (k0 > 0) ld(va, k0 - 1);
    else va = vb;  // unused by the bottom row
    ld(vc, k0 + 1 < mz ? k0 + 1 : mz - 1);
    ldr(rw, k0);
    for (int k = k0; k < k1; ++k) {
        // operands of the next level, one step ahead
        Lv vn;
        const bool more = k + 1 < k1;
        if (more) {
            ld(vn, k + 2 < mz ? k + 2 : mz - 1);
            ldr(rn, k + 1);
        }
        auto row = [&](const Win<float> &x, const Cf &f, const typename Op::Bc &bcv) {
            if (k == 0) return op.template row_c<kRowBottom>(false, x, f, bcv);
            if (k == mz - 1) return op.template row_c<kRowTop>(false, x, f, bcv);
            return op.template row_c<kRowInterior>(false, x, f, bcv);
        };
        const float r00 = rw.b00 - row(win(va.C00, vb.C00, vc.C00, va.W0, vb.W0, vc.W0, va.C10, vb.C10, vc.C10,
                                           va.S0, vb.S0, vc.S0, va.C01, vb.C01, vc.C01), rw.f00, b00);
        float s = r00;
        if (hi) {
            const float r = rw.b10 - row(win(va.C10, vb.C10, vc.C10, va.C00, vb.C00, vc.C00, va.E0, vb.E0, vc.E0,
                                             va.S1, vb.S1, vc.S1, va.C11, vb.C11, vc.C11), rw.f10, b10);
            s += r;
        }
        if (hj) {
            float t = rw.b01 - row(win(va.C01, vb.C01, vc.C01, va.W1, vb.W1, vc.W1, va.C11, vb.C11, vc.C11,
                                       va.C00, vb.C00, vc.C00, va.N0, vb.N0, vc.N0), rw.f01, b01);
            if (hi)
                t += rw.b11 - row(win(va.C11, vb.C11, vc.C11, va.C01, vb.C01, vc.C01, va.E1, vb.E1, vc.E1,
                                      va.C10, vb.C10, vc.C10, va.N1, vb.N1, vc.N1), rw.f11, b11);
            s += t;
        }
        bc[ci.n(k, cc)] = s / (float)((hi ? 2 : 1) * (hj ? 2 : 1));
        if (more) {
            va = vb;
            vb = vc;
            vc = vn;
            rw = rn;
        }
    }
}

}  // namespace windhip
