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
// column (I, J) aggregates the fine columns 2I-1, 2I and 2J-1, 2J (one of
// each at an odd edge). Coarse operators are Galerkin products R A P with
// piecewise-constant P and R = P^T. The stencil has no (i+-1, j+-1) entries,
// so the product keeps the 15-point pattern of the probed operator
// (spmv, wind_iterative.F90:888-904). Aggregation reproduces first-derivative
// and cross terms exactly but doubles second differences along a coarsened
// direction; the same-level horizontal couplings (D, E, F, G) of interior
// rows are therefore scaled by hs (1/2) with the diagonal moved so that every
// row still sums to zero. Coefficients are accumulated in FP64 and stored in
// FP32; the column tridiagonals are factored in FP64 in the same pass.
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

// Offsets (di, dk, dj) of the stencil entries A..O (spmv, wind_iterative.F90:888-904).
__device__ __forceinline__ int off_i(int q) { return ((q == 3 || q == 7 || q == 9) ? 1 : 0) - ((q == 4 || q == 8 || q == 10) ? 1 : 0); }
__device__ __forceinline__ int off_k(int q) { return ((q == 1 || q == 7 || q == 8 || q == 11 || q == 12) ? 1 : 0) - ((q == 2 || q == 9 || q == 10 || q == 13 || q == 14) ? 1 : 0); }
__device__ __forceinline__ int off_j(int q) { return ((q == 5 || q == 11 || q == 13) ? 1 : 0) - ((q == 6 || q == 12 || q == 14) ? 1 : 0); }
// entry index of offset (di, dk, dj); -1 outside the pattern
__device__ __forceinline__ int stencil_index(int di, int dk, int dj)
{
    if (di == 0 && dj == 0) return dk == 0 ? 0 : (dk > 0 ? 1 : 2);
    if (dj == 0) {
        if (dk == 0) return di > 0 ? 3 : 4;
        if (dk > 0) return di > 0 ? 7 : 8;
        return di > 0 ? 9 : 10;
    }
    if (di == 0) {
        if (dk == 0) return dj > 0 ? 5 : 6;
        if (dk > 0) return dj > 0 ? 11 : 12;
        return dj > 0 ? 13 : 14;
    }
    return -1;
}

// Window with 1 at stencil entry q, 0 elsewhere: (A x)(row) of it is the
// row's coefficient of that entry.
template <typename T>
__device__ __forceinline__ Win<T> onehot(int q)
{
    Win<T> w = {};
    switch (q) {
    case 0: w.c1 = T(1); break;
    case 1: w.c2 = T(1); break;
    case 2: w.c0 = T(1); break;
    case 3: w.e1 = T(1); break;
    case 4: w.w1 = T(1); break;
    case 5: w.n1 = T(1); break;
    case 6: w.s1 = T(1); break;
    case 7: w.e2 = T(1); break;
    case 8: w.w2 = T(1); break;
    case 9: w.e0 = T(1); break;
    case 10: w.w0 = T(1); break;
    case 11: w.n2 = T(1); break;
    case 12: w.s2 = T(1); break;
    case 13: w.n0 = T(1); break;
    default: w.s0 = T(1); break;
    }
    return w;
}

// Coefficients of an interior-column row of the fine operator (analytic or
// probed, FP64), read off its own row evaluation.
template <class Op>
struct MgExtractOp {
    Op op;
    long P;
    int mz;
    __device__ void coefs(int k, long c, double a[15]) const
    {
        const long n = (long)k * P + c;
        const typename Op::Cf f = op.load(n);
        const typename Op::Bc b = op.bc(c);
#pragma unroll
        for (int q = 0; q < 15; ++q) a[q] = (double)op.row(k, mz, false, onehot<typename Op::type>(q), f, b);
    }
};

// Coefficients of a coarse level: entry 0 holds the row sum, 1..14 the
// off-diagonal couplings; the diagonal is the sum minus the couplings.
struct MgExtractC {
    const float *A;
    long P, N;
    __device__ void coefs(int k, long c, double a[15]) const
    {
        const long n = (long)k * P + c;
        double off = 0.0;
#pragma unroll
        for (int q = 1; q < 15; ++q) {
            a[q] = (double)A[q * N + n];
            off += a[q];
        }
        a[0] = (double)A[n] - off;
    }
};

// fine column index -> coarse column index (lateral ring to lateral ring)
__device__ __forceinline__ int mg_cmap(int i, int nf, int nc)
{
    if (i <= 0) return 0;
    if (i > nf) return nc + 1;
    return (i + 1) >> 1;
}

// Galerkin coarse operator of level c from level f, one coarse row per thread.
template <class Ext>
__global__ __launch_bounds__(kMgThreads) void k_mg_galerkin(Ext ext, MgGeo f, MgGeo c, float *__restrict__ Ac,
                                                            double hs)
{
    const long n = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (n >= c.N) return;
    const int k = (int)(n / c.P);
    const long cc = n - (long)k * c.P;
    const int I = (int)(cc % c.mx), J = (int)(cc / c.mx);
    double a[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) a[q] = 0.0;
    if (I == 0 || I == c.mx - 1 || J == 0 || J == c.my - 1) {
        a[0] = 1.0;
    } else {
        for (int dj = 0; dj < 2; ++dj) {
            const int j = 2 * J - 1 + dj;
            if (j > f.ny) continue;
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I - 1 + di;
                if (i > f.nx) continue;
                double fa[15];
                ext.coefs(k, (long)j * f.mx + i, fa);
                for (int q = 0; q < 15; ++q) {
                    if (fa[q] == 0.0) continue;
                    const int NI = mg_cmap(i + off_i(q), f.nx, c.nx), NJ = mg_cmap(j + off_j(q), f.ny, c.ny);
                    const int qc = stencil_index(NI - I, off_k(q), NJ - J);
                    if (qc >= 0) a[qc] += fa[q];
                }
            }
        }
        if (k >= 1 && k <= c.mz - 2) {
            const double hsum = ((a[3] + a[4]) + a[5]) + a[6];
            a[3] *= hs; a[4] *= hs; a[5] *= hs; a[6] *= hs;
            a[0] += (1.0 - hs) * hsum;
        }
    }
    double off = 0.0;
    for (int q = 1; q < 15; ++q) {
        Ac[q * c.N + n] = (float)a[q];
        off += a[q];
    }
    Ac[n] = (float)(a[0] + off);  // row sum (zero up to rounding on every non-lateral row)
}

// Thomas factors (lo, cp, 1/pivot) in FP64 of every column's tridiagonal
// (C, A, B) of level c as stored, i.e. of the operator the residual kernel
// applies: the diagonal is s - sum of the couplings, written as h - B - C with
// h = s - (couplings outside the column), so that the pivots of nearly
// singular columns (derivative conditions at both ends, weak horizontal
// coupling on the coarse levels) carry no cancellation of B and C.
__global__ __launch_bounds__(kMgThreads) void k_mg_factor(MgGeo c, const float *__restrict__ Ac,
                                                          float *__restrict__ lo, float *__restrict__ cp,
                                                          float *__restrict__ ip)
{
    const long cc = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (cc >= c.P) return;
    double cprev = 0.0;
    for (int k = 0; k < c.mz; ++k) {
        const long n = (long)k * c.P + cc;
        double h = (double)Ac[n];
        for (int q = 3; q < 15; ++q) h -= (double)Ac[q * c.N + n];
        const double l = (double)Ac[2 * c.N + n], u = (double)Ac[c.N + n];
        double piv = (h - u) - l * (1.0 + cprev);  // d - l cp_{k-1} with d = h - l - u
        if (fabs(piv) < 1e-300) piv = copysign(1e-300, piv);
        const double inv = 1.0 / piv;
        cprev = u * inv;
        lo[n] = (float)l;
        cp[n] = (float)cprev;
        ip[n] = (float)inv;
    }
}

// r = b - A x on a coarse level, one row per thread: lateral rows are
// identity; every other row sums to zero, so (A x)(row) = s x_c + sum_q a_q
// (x_q - x_c) with s the stored row sum (differences keep the smooth
// components accurate in FP32). (The V-cycle kernels do nothing once the
// inner solve whose preconditioner they apply has converged.)
__global__ __launch_bounds__(kMgThreads) void k_mg_resid(MgGeo g, const KState *st, const float *__restrict__ A,
                                                         const float *__restrict__ x, const float *__restrict__ b,
                                                         float *__restrict__ r)
{
    if (st->done) return;
    const long n = (long)blockIdx.x * kMgThreads + threadIdx.x;
    if (n >= g.N) return;
    const int k = (int)(n / g.P);
    const long c = n - (long)k * g.P;
    const int i = (int)(c % g.mx), j = (int)(c / g.mx);
    const float xc = x[n];
    if (i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1) {
        r[n] = b[n] - xc;
        return;
    }
    const long N = g.N, P = g.P, mx = g.mx;
    float ax = A[n] * xc;
    ax += A[3 * N + n] * (x[n + 1] - xc);
    ax += A[4 * N + n] * (x[n - 1] - xc);
    ax += A[5 * N + n] * (x[n + mx] - xc);
    ax += A[6 * N + n] * (x[n - mx] - xc);
    if (k + 1 < g.mz) {
        const long u = n + P;
        ax += A[1 * N + n] * (x[u] - xc);
        ax += A[7 * N + n] * (x[u + 1] - xc);
        ax += A[8 * N + n] * (x[u - 1] - xc);
        ax += A[11 * N + n] * (x[u + mx] - xc);
        ax += A[12 * N + n] * (x[u - mx] - xc);
    }
    if (k > 0) {
        const long d = n - P;
        ax += A[2 * N + n] * (x[d] - xc);
        ax += A[9 * N + n] * (x[d + 1] - xc);
        ax += A[10 * N + n] * (x[d - 1] - xc);
        ax += A[13 * N + n] * (x[d + mx] - xc);
        ax += A[14 * N + n] * (x[d - mx] - xc);
    }
    r[n] = b[n] - ax;
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

// b_c = R r_f, one coarse row per thread: every interior row sums the rows of
// its aggregate; lateral rows get zero.
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
    bc[n] = s;
}

// x_f += P e_c, one fine row per thread: every interior row takes its
// aggregate's value.
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
    xf[n] += ec[(long)k * c.P + (long)((j + 1) >> 1) * c.mx + ((i + 1) >> 1)];
}

}  // namespace windhip
