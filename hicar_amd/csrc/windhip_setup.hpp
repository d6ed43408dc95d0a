// windhip_setup.hpp -- per-nest and per-call kernels around the Krylov loop:
// field gathers from the caller's Fortran arrays, module-array staging
// (init_module_vars), analytic assembly, diagonal / BC planes / rhs, the u/v/w
// correction G, the divergence D and operator probing.
#pragma once

#include "windhip_kernels.hpp"

namespace windhip {

// A caller field array: Fortran (i,k,j) with lower bounds (i0,k0,j0) and
// extents ni, nk (the j extent is implied). Element (i,k,j) lives at
// (i-i0) + ni*((k-k0) + nk*(j-j0)).
template <typename F>
struct Fld {
    F *p;
    int i0, k0, j0, ni, nk;
    __device__ __forceinline__ F &at(int i, int k, int j) const
    {
        return p[(long)(i - i0) + (long)ni * ((long)(k - k0) + (long)nk * (long)(j - j0))];
    }
    __device__ __forceinline__ F &at2(int i, int j) const { return p[(long)(i - i0) + (long)ni * (long)(j - j0)]; }
};

struct Dims {
    int nx, ny, nz, mx, my, mz;
    long P;
    double dx;
    __device__ __forceinline__ long r(int i, int k, int j) const { return (long)k * P + (long)j * mx + i; }
    __device__ __forceinline__ long q(int i, int j) const { return (long)j * mx + i; }
};

// dst(rows, k*P + j*mx + i) = src(i,k,j) * scale for i in [ia,ib], k in [ka,kb], j in [ja,jb];
// changed (may be null) is set to 1 when any gathered value differs bitwise
// from what dst held before (the solver's test for a new alpha).
template <typename F>
__global__ void k_gather3(Dims d, Fld<F> src, double *__restrict__ dst, int ia, int ib, int ka, int kb,
                          int ja, int jb, double scale, double *changed)
{
    const int ni = ib - ia + 1, nk = kb - ka + 1, nj = jb - ja + 1;
    const long tot = (long)ni * nk * nj;
    bool diff = false;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = ia + (int)(e % ni);
        const long t = e / ni;
        const int k = ka + (int)(t % nk);
        const int j = ja + (int)(t / nk);
        const long n = d.r(i, k, j);
        const double v = scale * (double)src.at(i, k, j);
        if (changed) diff |= __double_as_longlong(dst[n]) != __double_as_longlong(v);
        dst[n] = v;
    }
    if (changed && diff) *changed = 1.0;
}

template <typename F>
__global__ void k_gather2(Dims d, Fld<F> src, double *__restrict__ dst, int ia, int ib, int ja, int jb)
{
    const int ni = ib - ia + 1, nj = jb - ja + 1;
    const long tot = (long)ni * nj;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = ia + (int)(e % ni);
        const int j = ja + (int)(e / ni);
        dst[d.q(i, j)] = (double)src.at2(i, j);
    }
}

// Caller field -> natural (i,k,j) layout over (nx, nz, ny), FP64.
template <typename F>
__global__ void k_to_natural(Dims d, Fld<F> src, double *__restrict__ dst)
{
    const long tot = (long)d.nx * d.nz * d.ny;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const long t = e / d.nx;
        const int k = 1 + (int)(t % d.nz);
        const int j = 1 + (int)(t / d.nz);
        dst[e] = (double)src.at(i, k, j);
    }
}

// min/max of alpha and NaN count of div over the cells (calc_iter_winds
// :322-329 and :412-420). out[0] = min, out[1] = max, out[2] = NaN count.
__global__ __launch_bounds__(kBlock) void k_scan_inputs(Dims d, const double *__restrict__ alpha,
                                                        const double *__restrict__ b, double *part,
                                                        unsigned *counter, double *out, int nblocks)
{
    __shared__ double smin[kBlock], smax[kBlock], snan[kBlock];
    __shared__ int is_last;
    double mn = HUGE_VAL, mxv = -HUGE_VAL, nn = 0.0;
    // a wavefront per (j, k) row of cells, its lanes along i: contiguous
    // loads and one 32-bit index division per row
    constexpr int kW = kBlock / 64;
    const int lane = threadIdx.x & 63, rows = d.ny * d.nz;
    for (int row = blockIdx.x * kW + (threadIdx.x >> 6); row < rows; row += gridDim.x * kW) {
        const long base = d.r(0, 1 + row / d.ny, 1 + row % d.ny);
        for (int i = 1 + lane; i <= d.nx; i += 64) {
            const double av = alpha[base + i], bv = b[base + i];
            mn = fmin(mn, av);
            mxv = fmax(mxv, av);
            if (bv != bv) nn += 1.0;
        }
    }
    smin[threadIdx.x] = mn; smax[threadIdx.x] = mxv; snan[threadIdx.x] = nn;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            snan[threadIdx.x] += snan[threadIdx.x + s];
        }
        __syncthreads();
    }
    // block partials through agent-scope stores (no fence: a release here
    // would write back the whole L2), the arrival through an agent-scope
    // counter; the last block folds the partials with all its threads
    if (threadIdx.x == 0) {
        __hip_atomic_store(&part[3 * blockIdx.x], smin[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&part[3 * blockIdx.x + 1], smax[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&part[3 * blockIdx.x + 2], snan[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);  // the partials have landed before the counter moves
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        is_last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                  (unsigned)(nblocks - 1);
    }
    __syncthreads();
    if (!is_last) return;
    double a = HUGE_VAL, bmax = -HUGE_VAL, c = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += kBlock) {
        a = fmin(a, ld_coherent(&part[3 * q]));
        bmax = fmax(bmax, ld_coherent(&part[3 * q + 1]));
        c += ld_coherent(&part[3 * q + 2]);
    }
    __syncthreads();
    smin[threadIdx.x] = a; smax[threadIdx.x] = bmax; snan[threadIdx.x] = c;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            snan[threadIdx.x] += snan[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smin[0]; out[1] = smax[0]; out[2] = snan[0];
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// init_module_vars (wind_iterative.F90:2350-2409): dz_if, surface slopes.
// adv_dz, dzdx, dzdy on the row grid; terrain on the plane grid.
__global__ void k_module_vars(Dims d, const double *__restrict__ adv_dz, const double *__restrict__ dzdx,
                              const double *__restrict__ dzdy, const double *__restrict__ terrain,
                              double *__restrict__ dz_if, double *__restrict__ sx, double *__restrict__ sy)
{
    const long tot = (long)d.nx * d.ny;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const int j = 1 + (int)(e / d.nx);
        for (int k = 2; k <= d.nz; ++k)
            dz_if[d.r(i, k, j)] = (adv_dz[d.r(i, k, j)] + adv_dz[d.r(i, k - 1, j)]) / 2;
        dz_if[d.r(i, 1, j)] = adv_dz[d.r(i, 1, j)];
        dz_if[d.r(i, d.nz + 1, j)] = adv_dz[d.r(i, d.nz, j)];
        double vx = dzdx[d.r(i, 1, j)], vy = dzdy[d.r(i, 1, j)];
        if (i >= 2 && i <= d.nx - 1) vx = (terrain[d.q(i + 1, j)] - terrain[d.q(i - 1, j)]) / (2 * d.dx);
        if (j >= 2 && j <= d.ny - 1) vy = (terrain[d.q(i, j + 1)] - terrain[d.q(i, j - 1)]) / (2 * d.dx);
        sx[d.q(i, j)] = vx;
        sy[d.q(i, j)] = vy;
    }
}

struct Metrics {
    const double *dzdx, *dzdy, *jaco, *jaco_w, *alpha, *adv_dz, *dz_if;  // rows
    const double *dzdx_u, *jaco_u, *dzdy_v, *jaco_v;                     // rows (u: i to nx+1, v: j to ny+1)
    const double *sx, *sy;                                               // planes
    const double *mf_mx_u, *mf_my_u, *mf_my_v, *mf_mx_v, *mf_mxy;        // planes or null
};

// Analytic assembly. B, C, A per cell as update_coefs_gpu
// (wind_iterative.F90:2555-2603); with full != 0 also H, I, L, M as
// initialize_coefs (:2455-2465) (J = -H, K = -I, N = -L, O = -M are implied).
// Writes FP64 coefficients and their FP32 roundings.
__global__ void k_assemble(Dims d, Metrics m, int full, double *B, double *C, double *H, double *I,
                           double *L, double *M, float *B32, float *C32, float *H32, float *I32,
                           float *L32, float *M32)
{
    const long tot = (long)d.nx * d.ny * d.nz;
    const double dx = d.dx;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const long t = e / d.nx;
        const int j = 1 + (int)(t % d.ny);
        const int k = 1 + (int)(t / d.ny);
        const long n = d.r(i, k, j);
        const double adv_dz_k = m.adv_dz[d.r(1, k, 1)];
        const double dzdx = m.dzdx[n], dzdy = m.dzdy[n], al = m.alpha[n];
        double M_up, M_dwn;
        if (k == d.nz) {
            M_up = (dzdx * dzdx + dzdy * dzdy) + al * al;
        } else {
            const long n1 = d.r(i, k + 1, j);
            const double a1 = m.adv_dz[d.r(1, k + 1, 1)];
            const double sum_up = adv_dz_k + a1;
            const double jw = m.jaco_w[n];
            const double dzdx1 = m.dzdx[n1], dzdy1 = m.dzdy[n1], al1 = m.alpha[n1];
            M_up = ((dzdx * (dzdx * a1 + dzdx1 * adv_dz_k)) / sum_up) / jw
                 + ((dzdy * (dzdy * a1 + dzdy1 * adv_dz_k)) / sum_up) / jw;
            M_up = M_up + ((al * al * a1 + al1 * al1 * adv_dz_k) / sum_up) / jw;
        }
        if (k == 1) {
            M_dwn = (dzdx * m.sx[d.q(i, j)] + dzdy * m.sy[d.q(i, j)]) + al * al;
        } else {
            const long nm = d.r(i, k - 1, j);
            const double am1 = m.adv_dz[d.r(1, k - 1, 1)];
            const double sum_dwn = adv_dz_k + am1;
            const double jw = m.jaco_w[nm];
            const double dzdxm = m.dzdx[nm], dzdym = m.dzdy[nm], alm = m.alpha[nm];
            M_dwn = ((dzdx * (dzdx * am1 + dzdxm * adv_dz_k)) / sum_dwn) / jw
                  + ((dzdy * (dzdy * am1 + dzdym * adv_dz_k)) / sum_dwn) / jw;
            M_dwn = M_dwn + ((al * al * am1 + alm * alm * adv_dz_k) / sum_dwn) / jw;
        }
        const double jc = m.jaco[n];
        double b = M_up / ((jc * adv_dz_k) * m.dz_if[d.r(1, k + 1, 1)]);
        double c = M_dwn / ((jc * adv_dz_k) * m.dz_if[d.r(1, k, 1)]);
        const double md = ((m.dz_if[d.r(i, k + 1, j)] + m.dz_if[n]) * 2.0) * dx;
        const long ue = d.r(i + 1, k, j), vn = d.r(i, k, j + 1);
        const double X = -(((m.dzdx_u[ue] / m.jaco_u[ue] - m.dzdx_u[n] / m.jaco_u[n])
                            + (m.dzdy_v[vn] / m.jaco_v[vn] - m.dzdy_v[n] / m.jaco_v[n])) / md);
        b = b + X;
        c = c - X;
        B[n] = b; C[n] = c;
        B32[n] = (float)b; C32[n] = (float)c;
        if (full) {
            const double mdh = (2 * dx) * (m.dz_if[d.r(1, k + 1, 1)] + m.dz_if[d.r(1, k, 1)]);
            const double jac = m.jaco[n];
            const double ax = dzdx / jac, ay = dzdy / jac;
            const double hx = ax + m.dzdx_u[ue] / m.jaco_u[ue];
            const double ix = ax + m.dzdx_u[n] / m.jaco_u[n];
            const double ly = ay + m.dzdy_v[vn] / m.jaco_v[vn];
            const double my = ay + m.dzdy_v[n] / m.jaco_v[n];
            const double h = -(hx / mdh), iv = ix / mdh, l = -(ly / mdh), mv = my / mdh;
            H[n] = h; I[n] = iv; L[n] = l; M[n] = mv;
            H32[n] = (float)h; I32[n] = (float)iv; L32[n] = (float)l; M32[n] = (float)mv;
        }
    }
}

// Bottom/top BC planes for the current alpha (spmv :962-988 reads alpha live).
__global__ void k_bc_planes(Dims d, Metrics m, double *bcx, double *bcy, double *bcz, double *dztop,
                            float *bcx32, float *bcy32, float *bcz32, float *dztop32)
{
    const long tot = (long)d.nx * d.ny;
    const double two_dx = 2.0 * d.dx;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const int j = 1 + (int)(e / d.nx);
        const long q = d.q(i, j);
        const double sx = m.sx[q], sy = m.sy[q];
        const double al = m.alpha[d.r(i, 1, j)];
        const double denom = (2.0 * ((sx * sx + sy * sy) + al * al)) / m.jaco[d.r(i, 1, j)];
        const double cy = sy / (denom * two_dx);
        const double cx = sx / (denom * two_dx);
        const double cz = 1.0 / m.dz_if[d.r(i, 1, j)];
        const double dt = m.dz_if[d.r(i, d.nz + 1, j)];
        bcx[q] = cx; bcy[q] = cy; bcz[q] = cz; dztop[q] = dt;
        bcx32[q] = (float)cx; bcy32[q] = (float)cy; bcz32[q] = (float)cz; dztop32[q] = (float)dt;
    }
}

// D^-1 over all rows (extract_diagonal, wind_iterative.F90:1041-1075), FP64 and
// the FP32 operator's own diagonal.
__global__ void k_diagonal(Dims d, int probed, const double *A15, const double *B, const double *C,
                           const float *B32, const float *C32, const float *A15_32, const double *bcz,
                           const double *dztop, const float *bcz32, const float *dztop32,
                           double *dinv, float *dinv32)
{
    const double inv = 1.0 / (d.dx * d.dx);
    const double neg4 = -4.0 * inv;
    const float invf = (float)inv, neg4f = (float)neg4;
    const long N = d.P * d.mz;
    for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (long)gridDim.x * blockDim.x) {
        const int k = (int)(n / d.P);
        const long c = n - (long)k * d.P;
        const int j = (int)(c / d.mx), i = (int)(c % d.mx);
        double dd;
        float df;
        if (i <= 0 || j <= 0 || i >= d.mx - 1 || j >= d.my - 1) { dd = 1.0; df = 1.0f; }
        else if (k >= d.mz - 1) {
            if (probed) { dd = 1.0; df = 1.0f; }
            else { dd = clamp_d<double>(1.0 / dztop[c]); df = clamp_d<float>(1.0f / dztop32[c]); }
        } else if (k <= 0) {
            if (probed) { dd = 1.0; df = 1.0f; }
            else { dd = clamp_d<double>(-bcz[c]); df = clamp_d<float>(-bcz32[c]); }
        } else if (probed) {
            dd = clamp_d<double>(A15[n]); df = clamp_d<float>(A15_32[n]);
        } else {
            dd = clamp_d<double>((neg4 - B[n]) - C[n]);
            df = clamp_d<float>((neg4f - B32[n]) - C32[n]);
        }
        dinv[n] = 1.0 / dd;
        dinv32[n] = 1.0f / df;
    }
    (void)invf;
}

// ------------------------------------------------------------- correction G
// calc_updated_winds (wind_iterative.F90:1742-2071) with an FP64 lambda.
// density (mass-grid Fld) is used only when adv_den; map factors may be null.
template <typename F>
__global__ void k_correct_u(Dims d, Metrics m, const double *__restrict__ lam, Fld<F> u, Fld<const F> rho,
                            int use_rho)
{
    const long tot = (long)(d.nx + 1) * d.ny;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % (d.nx + 1));
        const int j = 1 + (int)(e / (d.nx + 1));
        const double mf = m.mf_mx_u ? m.mf_mx_u[d.q(i, j)] : 1.0;
        double ut_m = 0.0, ut_0 = (lam[d.r(i, 0, j)] + lam[d.r(i - 1, 0, j)]) / 2;
        double ut_1 = (lam[d.r(i, 1, j)] + lam[d.r(i - 1, 1, j)]) / 2;
        for (int k = 1; k <= d.nz; ++k) {
            // window: ut_m = ut(k-1), ut_0 = ut(k), ut_1 = ut(k+1)
            ut_m = ut_0; ut_0 = ut_1;
            ut_1 = (lam[d.r(i, k + 1, j)] + lam[d.r(i - 1, k + 1, j)]) / 2;
            double dl;
            if (k == 1) dl = (ut_1 - ut_0) / m.dz_if[d.r(1, 2, 1)];
            else if (k == d.nz) dl = (ut_0 - ut_m) / m.dz_if[d.r(1, d.nz, 1)];
            else dl = (ut_1 - ut_m) / (m.dz_if[d.r(1, k + 1, 1)] + m.dz_if[d.r(1, k, 1)]);
            double ru;
            if (use_rho) {
                if (i == 1) ru = (double)rho.at(1, k, j);
                else if (i == d.nx + 1) ru = (double)rho.at(d.nx, k, j);
                else ru = 0.5 * ((double)rho.at(i, k, j) + (double)rho.at(i - 1, k, j));
            } else {
                ru = (i == 1 || i == d.nx + 1) ? 1.0 : 0.5 * (1.0 + 1.0);
            }
            const long n = d.r(i, k, j);
            const double grad = (lam[n] - lam[d.r(i - 1, k, j)]) / d.dx - (m.dzdx_u[n] * dl) / m.jaco_u[n];
            F &uv = u.at(i, k, j);
            uv = (F)((double)uv + ((0.5 * mf) * grad) / ru);
        }
    }
}

template <typename F>
__global__ void k_correct_v(Dims d, Metrics m, const double *__restrict__ lam, Fld<F> v, Fld<const F> rho,
                            int use_rho)
{
    const long tot = (long)d.nx * (d.ny + 1);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const int j = 1 + (int)(e / d.nx);
        const double mf = m.mf_my_v ? m.mf_my_v[d.q(i, j)] : 1.0;
        double vt_m = 0.0, vt_0 = (lam[d.r(i, 0, j)] + lam[d.r(i, 0, j - 1)]) / 2;
        double vt_1 = (lam[d.r(i, 1, j)] + lam[d.r(i, 1, j - 1)]) / 2;
        for (int k = 1; k <= d.nz; ++k) {
            vt_m = vt_0; vt_0 = vt_1;
            vt_1 = (lam[d.r(i, k + 1, j)] + lam[d.r(i, k + 1, j - 1)]) / 2;
            double dl;
            if (k == 1) dl = (vt_1 - vt_0) / m.dz_if[d.r(1, 2, 1)];
            else if (k == d.nz) dl = (vt_0 - vt_m) / m.dz_if[d.r(1, d.nz, 1)];
            else dl = (vt_1 - vt_m) / (m.dz_if[d.r(1, k + 1, 1)] + m.dz_if[d.r(1, k, 1)]);
            double rv;
            if (use_rho) {
                if (j == 1) rv = (double)rho.at(i, k, 1);
                else if (j == d.ny + 1) rv = (double)rho.at(i, k, d.ny);
                else rv = 0.5 * ((double)rho.at(i, k, j) + (double)rho.at(i, k, j - 1));
            } else {
                rv = (j == 1 || j == d.ny + 1) ? 1.0 : 0.5 * (1.0 + 1.0);
            }
            const long n = d.r(i, k, j);
            const double grad = (lam[n] - lam[d.r(i, k, j - 1)]) / d.dx - (m.dzdy_v[n] * dl) / m.jaco_v[n];
            F &vv = v.at(i, k, j);
            vv = (F)((double)vv + ((0.5 * mf) * grad) / rv);
        }
    }
}

template <typename F>
__global__ void k_correct_w(Dims d, Metrics m, const double *__restrict__ lam, Fld<F> w, Fld<const F> rho,
                            int use_rho)
{
    const long tot = (long)d.nx * d.ny;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const int j = 1 + (int)(e / d.nx);
        for (int k = 1; k <= d.nz - 1; ++k) {
            const long n = d.r(i, k, j), n1 = d.r(i, k + 1, j);
            double rw;
            if (use_rho) {
                const double dzk = m.adv_dz[n], dzk1 = m.adv_dz[n1];
                rw = ((double)rho.at(i, k, j) * dzk1 + (double)rho.at(i, k + 1, j) * dzk) / (dzk + dzk1);
            } else {
                const double dzk = m.adv_dz[n], dzk1 = m.adv_dz[n1];
                rw = (1.0 * dzk1 + 1.0 * dzk) / (dzk + dzk1);
            }
            const double al = m.alpha[n];
            F &wv = w.at(i, k, j);
            wv = (F)((double)wv + ((((0.5 * (al * al)) * (lam[n1] - lam[n])) / m.dz_if[n1]) / m.jaco_w[n]) / rw);
        }
    }
}

// ------------------------------------------------------------- divergence D
// calc_divergence (wind.F90:250-523) on one tile (output type O, by default
// the fields' type).
template <typename F, typename O = F>
__global__ void k_divergence(Dims d, Metrics m, Fld<const F> u, Fld<const F> v, Fld<const F> w,
                             Fld<const F> rho, int use_rho, int adv_den, int horz_only, int use_dqdt,
                             Fld<O> out)
{
    const long tot = (long)d.nx * d.ny;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const int j = 1 + (int)(e / d.nx);
        auto R = [&](int ii, int kk, int jj) -> double { return use_rho ? (double)rho.at(ii, kk, jj) : 1.0; };
        const double m0 = m.mf_my_u ? m.mf_my_u[d.q(i, j)] : 1.0;
        const double m1 = m.mf_my_u ? m.mf_my_u[d.q(i + 1, j)] : 1.0;
        const double m2 = m.mf_mx_v ? m.mf_mx_v[d.q(i, j)] : 1.0;
        const double m3 = m.mf_mx_v ? m.mf_mx_v[d.q(i, j + 1)] : 1.0;
        const double mxy = m.mf_mxy ? m.mf_mxy[d.q(i, j)] : 1.0;
        double wm_prev = 0.0;
        for (int k = 1; k <= d.nz; ++k) {
            auto umet = [&](int f) -> double {
                const double uu = (double)u.at(f, k, j);
                const long n = d.r(f, k, j);
                if (adv_den) {
                    if (f == 1) return (uu * m.jaco_u[n]) * (1.5 * R(1, k, j) - 0.5 * R(2, k, j));
                    if (f == d.nx + 1) return (uu * m.jaco_u[n]) * (1.5 * R(d.nx, k, j) - 0.5 * R(d.nx - 1, k, j));
                    return ((uu * m.jaco_u[n]) * (R(f - 1, k, j) + R(f, k, j))) / 2;
                }
                return use_dqdt ? uu : uu * m.jaco_u[n];
            };
            auto vmet = [&](int f) -> double {
                const double vv = (double)v.at(i, k, f);
                const long n = d.r(i, k, f);
                if (adv_den) {
                    if (f == 1) return (vv * m.jaco_v[n]) * (1.5 * R(i, k, 1) - 0.5 * R(i, k, 2));
                    if (f == d.ny + 1) return (vv * m.jaco_v[n]) * (1.5 * R(i, k, d.ny) - 0.5 * R(i, k, d.ny - 1));
                    return ((vv * m.jaco_v[n]) * (R(i, k, f - 1) + R(i, k, f))) / 2;
                }
                return use_dqdt ? vv : vv * m.jaco_v[n];
            };
            double dv = (mxy * (((umet(i + 1) / m1 - umet(i) / m0) + vmet(j + 1) / m3) - vmet(j) / m2)) / d.dx;
            if (!horz_only) {
                const long n = d.r(i, k, j);
                double wm;
                const double ww = (double)w.at(i, k, j);
                if (adv_den) {
                    if (k < d.nz) {
                        const double dzk = m.adv_dz[n], dzk1 = m.adv_dz[d.r(i, k + 1, j)];
                        const double ri = (R(i, k, j) * dzk1 + R(i, k + 1, j) * dzk) / (dzk + dzk1);
                        wm = (ww * m.jaco_w[n]) * ri;
                    } else {
                        wm = (ww * m.jaco_w[n]) * R(i, k, j);
                    }
                } else {
                    wm = ww * m.jaco_w[n];
                }
                if (k == 1) dv = dv + wm / m.adv_dz[n];
                else dv = dv + (wm - wm_prev) / m.adv_dz[n];
                dv = dv / m.jaco[n];
                wm_prev = wm;
            }
            out.at(i, k, j) = (O)dv;
        }
    }
}

// ------------------------------------------------------- the caller's steps
// These run in FP64; values the reference keeps in arrays of the caller's type
// (the divergence, rho_i, w, the intermediate w of the restaggering) are
// rounded to that type where it stores them.
//
// calc_w (wind.F90:193-234): w from the horizontal divergence by upward
// integration of each column, with the interface density
// rho_i(k) = (rho(k) dz(k+1) + rho(k+1) dz(k)) / (dz(k) + dz(k+1)), rho_i(top)
// = rho(top), dz = advection_dz, when adv_den.
template <typename F>
__global__ void k_calc_w(Dims d, Metrics m, Fld<const double> div, Fld<const F> rho, int use_rho, int adv_den,
                         Fld<F> w)
{
    const long tot = (long)d.nx * d.ny;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const int j = 1 + (int)(e / d.nx);
        auto R = [&](int kk) -> double { return use_rho ? (double)rho.at(i, kk, j) : 1.0; };
        auto rho_i = [&](int kk) -> double {
            if (kk == d.nz) return R(kk);
            const double dz0 = m.adv_dz[d.r(i, kk, j)], dz1 = m.adv_dz[d.r(i, kk + 1, j)];
            return (double)(F)((R(kk) * dz1 + R(kk + 1) * dz0) / (dz0 + dz1));
        };
        double wprev = 0.0, riprev = 0.0;
        for (int k = 1; k <= d.nz; ++k) {
            const long n = d.r(i, k, j);
            const double dv = (double)(F)div.at(i, k, j), dz = m.adv_dz[n], jw = m.jaco_w[n];
            double wk;
            if (adv_den) {
                const double ri = rho_i(k);
                if (k == 1) wk = 0.0 - (dv * dz) / (jw * ri);
                else wk = ((wprev * riprev) * m.jaco_w[d.r(i, k - 1, j)] - dv * dz) / (jw * ri);
                riprev = ri;
            } else {
                if (k == 1) wk = (0.0 - dv * dz) / jw;
                else wk = (wprev * m.jaco_w[d.r(i, k - 1, j)] - dv * dz) / jw;
            }
            w.at(i, k, j) = (F)wk;
            wprev = (double)(F)wk;  // the reference reads w(k-1) back from the array
        }
    }
}

// calc_w_real (wind.F90:1175-1227): the real vertical velocity on the mass
// grid, (u dz/dx)_i,i+1 / 2 + (v dz/dy)_j,j+1 / 2 + (w_ext(k) + w_ext(k+1)) / 2
// with w_ext(1) = 0 and w_ext(k+1) = w_grid(k) jaco_w(k); a null w_grid is 0.
template <typename F>
__device__ __forceinline__ double w_real_at(const Dims &d, const Metrics &m, const Fld<const F> &u,
                                            const Fld<const F> &v, const F *wg_p, const Fld<const F> &wg, int i,
                                            int k, int j)
{
    const long n = d.r(i, k, j);
    const double uw = ((double)u.at(i, k, j) * m.dzdx_u[n] + (double)u.at(i + 1, k, j) * m.dzdx_u[d.r(i + 1, k, j)]) * 0.5;
    const double vw = ((double)v.at(i, k, j) * m.dzdy_v[n] + (double)v.at(i, k, j + 1) * m.dzdy_v[d.r(i, k, j + 1)]) * 0.5;
    double lo = 0.0, hi = 0.0;
    if (wg_p) {
        if (k > 1) lo = (double)wg.at(i, k - 1, j) * m.jaco_w[d.r(i, k - 1, j)];
        hi = (double)wg.at(i, k, j) * m.jaco_w[n];
    }
    return (uw + vw) + (lo + hi) * 0.5;
}

template <typename F>
__global__ void k_w_real(Dims d, Metrics m, Fld<const F> u, Fld<const F> v, Fld<const F> wg, Fld<F> out)
{
    const long tot = (long)d.nx * d.ny * d.nz;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const long t = e / d.nx;
        const int j = 1 + (int)(t % d.ny);
        const int k = 1 + (int)(t / d.ny);
        out.at(i, k, j) = (F)w_real_at<F>(d, m, u, v, wg.p, wg, i, k, j);
    }
}

// calc_idealized_wgrid (wind.F90:940-998): the grid-relative w that leaves a
// prescribed w_real, t = w_real - calc_w_real(u, v, 0) on the mass grid,
// restaggered to the level interfaces with advection_dz weights and divided
// by jaco_w; the top level is 0.
template <typename F>
__global__ void k_idealized_wgrid(Dims d, Metrics m, Fld<const F> u, Fld<const F> v, Fld<const F> w_real,
                                  Fld<F> w)
{
    const long tot = (long)d.nx * d.ny;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const int j = 1 + (int)(e / d.nx);
        auto tm = [&](int k) -> double {
            const double h = (double)(F)w_real_at<F>(d, m, u, v, nullptr, u, i, k, j);
            return (double)(F)((double)w_real.at(i, k, j) - h);
        };
        double tk = tm(1);
        for (int k = 1; k < d.nz; ++k) {
            const double t1 = tm(k + 1);
            const double a0 = m.adv_dz[d.r(i, k, j)], a1 = m.adv_dz[d.r(i, k + 1, j)];
            const double st = (double)(F)((tk * a1 + t1 * a0) / (a0 + a1));
            w.at(i, k, j) = (F)(st / m.jaco_w[d.r(i, k, j)]);
            tk = t1;
        }
        w.at(i, d.nz, j) = (F)0.0;
    }
}

// b = -2 div on interior rows (compute_rhs_3d, wind_iterative.F90:1513-1530);
// the BC rows of b are zero from allocation and never written.
template <typename F>
__global__ void k_rhs(Dims d, Fld<const F> div, double *__restrict__ b)
{
    const long tot = (long)d.nx * d.ny * d.nz;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const long t = e / d.nx;
        const int k = 1 + (int)(t % d.nz);
        const int j = 1 + (int)(t / d.nz);
        b[d.r(i, k, j)] = -2.0 * (double)div.at(i, k, j);
    }
}

// ------------------------------------------------------------- probing
// probe_lambda_pattern (wind_iterative.F90:1096-1115)
__global__ void k_probe_pattern(Dims d, int ca, int cb, int cc, double *__restrict__ x)
{
    const long N = d.P * d.mz;
    for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (long)gridDim.x * blockDim.x) {
        const int k = (int)(n / d.P);
        const long c = n - (long)k * d.P;
        const int j = (int)(c / d.mx), i = (int)(c % d.mx);
        const bool on = i >= 1 && i <= d.mx - 2 && k >= 1 && k <= d.mz - 2 && j >= 1 && j <= d.my - 2
                        && i % 3 == ca && k % 3 == cb && j % 3 == cc;
        x[n] = on ? 1.0 : 0.0;
    }
}

// probe_record (wind_iterative.F90:1169-1231): T is the divergence response on
// the natural (nx,nz,ny) layout; leak holds max |2T| off the stencil as the bit
// pattern of a non-negative double.
__global__ void k_probe_record(Dims d, int ca, int cb, int cc, const double *__restrict__ T,
                               double *const *cf, unsigned long long *leak)
{
    const long tot = (long)d.nx * d.ny * d.nz;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(e % d.nx);
        const long t = e / d.nx;
        const int k = 1 + (int)(t % d.nz);
        const int j = 1 + (int)(t / d.nz);
        int di = (ca - i % 3 + 3) % 3; if (di == 2) di = -1;
        int dk = (cb - k % 3 + 3) % 3; if (dk == 2) dk = -1;
        int dj = (cc - j % 3 + 3) % 3; if (dj == 2) dj = -1;
        const double tv = 2.0 * T[(long)(i - 1) + (long)d.nx * ((long)(k - 1) + (long)d.nz * (j - 1))];
        int q = -1;
        if (dk == 0) {
            if (di == 0 && dj == 0) q = 0; else if (di == 1 && dj == 0) q = 3;
            else if (di == -1 && dj == 0) q = 4; else if (di == 0 && dj == 1) q = 5;
            else if (di == 0 && dj == -1) q = 6;
        } else if (dk == 1) {
            if (di == 0 && dj == 0) q = 1; else if (di == 1 && dj == 0) q = 7;
            else if (di == -1 && dj == 0) q = 8; else if (di == 0 && dj == 1) q = 11;
            else if (di == 0 && dj == -1) q = 12;
        } else {
            if (di == 0 && dj == 0) q = 2; else if (di == 1 && dj == 0) q = 9;
            else if (di == -1 && dj == 0) q = 10; else if (di == 0 && dj == 1) q = 13;
            else if (di == 0 && dj == -1) q = 14;
        }
        if (q >= 0) cf[q][d.r(i, k, j)] = tv;
        else atomicMax(leak, (unsigned long long)__double_as_longlong(fabs(tv)));
    }
}

__global__ void k_round_f32(long n, const double *__restrict__ a, float *__restrict__ b)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        b[e] = (float)a[e];
}

__global__ void k_widen_f64(long n, const float *__restrict__ a, double *__restrict__ b)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        b[e] = (double)a[e];
}

// ------------------------------------------------------------------ learned guess
// Network input from the solver's fields, one thread per (x, y) column of the
// padded volume (Dp, Hp, Wp multiples of 8, channels-last, 8 channels):
// 0 rhs b * bscale, 1 dz/dx, 2 dz/dy, 3 Jacobian, 4 alpha, 5 height of the
// cell centre above the surface / column height, 6..7 zero; zero outside the
// grid.
// OT: float, or __bf16 (the network's own input type: the rounding of its
// FP32 input conversion, one pass less)
template <typename OT>
__global__ void k_net_input(Dims d, int Dp, int Hp, int Wp, const double *__restrict__ b, double bscale,
                            const double *__restrict__ dzdx, const double *__restrict__ dzdy,
                            const double *__restrict__ jaco, const double *__restrict__ alpha,
                            const double *__restrict__ adv_dz, OT *__restrict__ in)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)Hp * Wp) return;
    const int x = (int)(q % Wp), y = (int)(q / Wp);
    const bool col = x < d.nx && y < d.ny;
    double top = 0.0;
    if (col)
        for (int k = 1; k <= d.nz; ++k) top += adv_dz[d.r(x + 1, k, y + 1)];
    double h = 0.0;
    for (int z = 0; z < Dp; ++z) {
        OT *o = in + (((long)z * Hp + y) * Wp + x) * 8;
        float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (col && z < d.nz) {
            const long n = d.r(x + 1, z + 1, y + 1);
            const double dz = adv_dz[n];
            v[0] = (float)(b[n] * bscale);
            v[1] = (float)dzdx[n];
            v[2] = (float)dzdy[n];
            v[3] = (float)jaco[n];
            v[4] = (float)alpha[n];
            v[5] = (float)((h + 0.5 * dz) / top);
            h += dz;
        }
        for (int c = 0; c < 8; ++c) o[c] = (OT)v[c];
    }
}

// lambda rows from the network output: interior = scale * out, lateral rows
// 0 (their rhs), ghost planes copied from the adjacent interior level.
__global__ void k_net_output(Dims d, int Dp, int Hp, int Wp, const float *__restrict__ out, double scale,
                             double *__restrict__ x)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.P) return;
    const int i = (int)(q % d.mx), j = (int)(q / d.mx);
    const bool lat = i == 0 || i == d.mx - 1 || j == 0 || j == d.my - 1;
    for (int k = 1; k <= d.nz; ++k)
        x[d.r(i, k, j)] = lat ? 0.0 : scale * (double)out[((long)(k - 1) * Hp + (j - 1)) * Wp + (i - 1)];
    x[d.r(i, 0, j)] = x[d.r(i, 1, j)];
    x[d.r(i, d.mz - 1, j)] = x[d.r(i, d.nz, j)];
}

}  // namespace windhip
