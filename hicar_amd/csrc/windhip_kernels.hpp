// windhip_kernels.hpp -- gfx950 HIP kernels of the wind solve.
//
// Layouts in HBM (one tile):
//   rows  [k][j][i] with k slowest: n = k*P + c, c = j*mx + i, P = mx*my,
//         mx = nx+2, my = ny+2, mz = nz+2. The domain cell (i,k,j) of the
//         reference (1-based) is row (i,k,j); i = 0, mx-1 and j = 0, my-1 are the
//         lateral identity rows, k = 0 and mz-1 the ghost planes
//         (wind_iterative.F90:2279-2302). Cell, u-face, v-face and dz_if fields
//         are stored on the same row grid, so every array shares one index.
//   planes [j][i] (P values) for 2-D fields (surface slopes, BC coefficients,
//         map factors).
//
// Work decomposition of the stencil kernels: one thread per (i,j) column,
// marching k from the bottom ghost plane to the top one with a three-plane
// register window of the five columns C, W, E, S, N; 256 consecutive columns
// per block. Each column's loads are unit-stride across a wavefront.
// Blocks are remapped so that each XCD processes a contiguous range of
// columns (the j +- 1 rows of a block then sit in the same XCD's L2).
//
// Arithmetic contract: every expression rounds exactly as the reference writes
// it (left-to-right sums, no FMA: the file is compiled with -ffp-contract=off),
// and every dot product is summed in the canonical order
//   per column sum over k (from 0) -> 64-lane xor butterfly per wave ->
//   ((w0 + w1) + w2) + w3 per block -> lane t folds partials t, t+256, ...
//   -> the same block reduction.
// The CPU oracle sums in this same order, so the FP64 solve is bit-identical to
// it iteration by iteration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace windhip {

constexpr int kBlock = 256;
constexpr double kBreakdownEps = 1.0e-30;  // wind_iterative.F90:61

struct Geo {
    int mx, my, mz;
    long P;   // columns
    int nb;   // column blocks = ceil(P / 256)
};

// Device-resident scalars of one BiCGStab (bicgstab_solve, wind_iterative.F90:563-569).
struct KState {
    double rho, alpha, omega, beta;
    double rnorm, res0, target, bnorm;
    double sigma, ts, tt, rs, rt, ss;
    double tol_rel, tol_abs;
    double outer_rnorm, outer_bnorm, outer_target, inner_tol;
    int iter, status, done, fmode;  // fmode: 0 none, 1 x, 2 x+r, 3 x+r+p
    int max_iters, outer_done;
    unsigned cnt[4];
};

// Analytic operator (spmv analytic branch, wind_iterative.F90:912-1025) in
// compressed form: J=-H, K=-I, N=-L, O=-M, D=E=F=G=1/dx^2, A=-4/dx^2-B-C.
template <typename T>
struct OpA {
    const T *B, *C, *H, *I, *L, *M;   // rows
    const T *bcx, *bcy, *bcz, *dztop; // planes: bottom-row dzdx/(denom 2dx), dzdy/(...), 1/dz_if(1); top dz_if(nz+1)
    T inv, neg4;                      // 1/dx^2, -4/dx^2
};

// Probed 15-array operator (spmv probed branch, wind_iterative.F90:837-910).
template <typename T>
struct OpP {
    const T *cf[15];  // A..O, rows
};

__device__ __forceinline__ int logical_block(int nb)
{
    // Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
    // contiguous logical range [start(x), start(x) + count(x)).
    int b = blockIdx.x;
    int x = b & 7, r = b >> 3;
    int q = nb >> 3, rem = nb & 7;
    return x * q + (x < rem ? x : rem) + r;
}

template <typename T>
__device__ __forceinline__ T clamp_d(T d)
{
    if (fabs((double)d) < kBreakdownEps) d = d < T(0) ? T(-kBreakdownEps) : T(kBreakdownEps);
    return d;
}

__device__ __forceinline__ double wave_butterfly(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double ld_coherent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Canonical block reduction of NV per-thread values; the last block to finish
// folds all partials and thread 0 of it calls fin(total). Returns nothing.
template <int NV, class Fin>
__device__ __forceinline__ void reduce_finalize(double (&acc)[NV], double *partials,
                                                unsigned *counter, int L, int nb, Fin fin)
{
    __shared__ double red[4][NV];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double s = wave_butterfly(acc[v]);
        if (lane == 0) red[wv][v] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            partials[(long)L * NV + v] = ((red[0][v] + red[1][v]) + red[2][v]) + red[3][v];
        __threadfence();
        unsigned old = atomicAdd(counter, 1u);
        is_last = (old == (unsigned)(nb - 1));
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double fold[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) fold[v] = 0.0;
    for (int q = threadIdx.x; q < nb; q += kBlock) {
#pragma unroll
        for (int v = 0; v < NV; ++v) fold[v] = fold[v] + ld_coherent(&partials[(long)q * NV + v]);
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double s = wave_butterfly(fold[v]);
        if (lane == 0) red[wv][v] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) tot[v] = ((red[0][v] + red[1][v]) + red[2][v]) + red[3][v];
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fin(tot);
    }
}

// ------------------------------------------------------------------ the march
// Visits the rows of column c bottom to top and calls epi(k, n, y, xc, dinv)
// with y = (A x)(row), xc = x(row) and dinv = 1/diag(A)(row) (extract_diagonal,
// wind_iterative.F90:1041-1075) when NEED_DINV.
template <typename T, bool NEED_DINV, class Epi>
__device__ __forceinline__ void march_analytic(const OpA<T> &op, const Geo &g, const T *__restrict__ x,
                                               long c, int i, int j, Epi &epi)
{
    const long P = g.P;
    const bool lat = (i == 0) | (i == g.mx - 1) | (j == 0) | (j == g.my - 1);
    const long cw = lat ? c : c - 1, ce = lat ? c : c + 1;
    const long cs = lat ? c : c - g.mx, cn = lat ? c : c + g.mx;
    T c0 = x[c], w0 = x[cw], e0 = x[ce], s0 = x[cs], n0 = x[cn];
    T c1 = x[P + c], w1 = x[P + cw], e1 = x[P + ce], s1 = x[P + cs], n1 = x[P + cn];
    {   // bottom BC row, wind_iterative.F90:962-988
        const T bx = op.bcx[c], by = op.bcy[c], bz = op.bcz[c];
        T y = by * s0 + by * s1 + bx * w0 + (-bz) * c0 + (-bx) * e0 + bx * w1 + bz * c1
              + (-bx) * e1 + (-by) * n0 + (-by) * n1;
        T dv = T(1);
        if (lat) y = c0;
        if (NEED_DINV) {
            T d = lat ? T(1) : clamp_d<T>(-bz);
            dv = T(1) / d;
        }
        epi(0, c, y, c0, dv);
    }
    const T inv = op.inv, neg4 = op.neg4;
    for (int k = 1; k <= g.mz - 2; ++k) {
        const long o2 = (long)(k + 1) * P;
        const T c2 = x[o2 + c], w2 = x[o2 + cw], e2 = x[o2 + ce], s2 = x[o2 + cs], n2 = x[o2 + cn];
        const long n = (long)k * P + c;
        const T B = op.B[n], Cc = op.C[n], H = op.H[n], I = op.I[n], L = op.L[n], M = op.M[n];
        const T a = (neg4 - B) - Cc;
        // interior 15-point row, wind_iterative.F90:1005-1020
        T y = (-M) * s0 + inv * s1 + M * s2 + (-I) * w0 + Cc * c0 + (-H) * e0 + inv * w1 + a * c1
              + inv * e1 + I * w2 + B * c2 + H * e2 + (-L) * n0 + inv * n1 + L * n2;
        T dv = T(1);
        if (lat) y = c1;
        if (NEED_DINV) {
            T d = lat ? T(1) : clamp_d<T>(a);
            dv = T(1) / d;
        }
        epi(k, n, y, c1, dv);
        c0 = c1; w0 = w1; e0 = e1; s0 = s1; n0 = n1;
        c1 = c2; w1 = w2; e1 = e2; s1 = s2; n1 = n2;
    }
    {   // top BC row, wind_iterative.F90:950-960
        const long n = (long)(g.mz - 1) * P + c;
        const T dzt = op.dztop[c];
        T y = (c1 - c0) / dzt;
        T dv = T(1);
        if (lat) y = c1;
        if (NEED_DINV) {
            T d = lat ? T(1) : clamp_d<T>(T(1) / dzt);
            dv = T(1) / d;
        }
        epi(g.mz - 1, n, y, c1, dv);
    }
}

// Probed operator: lateral and ghost-plane identity rows, 15 stored arrays.
template <typename T, bool NEED_DINV, class Epi>
__device__ __forceinline__ void march_probed(const OpP<T> &op, const Geo &g, const T *__restrict__ x,
                                             long c, int i, int j, Epi &epi)
{
    const long P = g.P;
    const bool lat = (i == 0) | (i == g.mx - 1) | (j == 0) | (j == g.my - 1);
    const long cw = lat ? c : c - 1, ce = lat ? c : c + 1;
    const long cs = lat ? c : c - g.mx, cn = lat ? c : c + g.mx;
    T c0 = x[c], w0 = x[cw], e0 = x[ce], s0 = x[cs], n0 = x[cn];
    T c1 = x[P + c], w1 = x[P + cw], e1 = x[P + ce], s1 = x[P + cs], n1 = x[P + cn];
    epi(0, c, c0, c0, T(1));
    for (int k = 1; k <= g.mz - 2; ++k) {
        const long o2 = (long)(k + 1) * P;
        const T c2 = x[o2 + c], w2 = x[o2 + cw], e2 = x[o2 + ce], s2 = x[o2 + cs], n2 = x[o2 + cn];
        const long n = (long)k * P + c;
        // wind_iterative.F90:888-904
        T y = op.cf[14][n] * s0 + op.cf[6][n] * s1 + op.cf[12][n] * s2 + op.cf[10][n] * w0
              + op.cf[2][n] * c0 + op.cf[9][n] * e0 + op.cf[4][n] * w1 + op.cf[0][n] * c1
              + op.cf[3][n] * e1 + op.cf[8][n] * w2 + op.cf[1][n] * c2 + op.cf[7][n] * e2
              + op.cf[13][n] * n0 + op.cf[5][n] * n1 + op.cf[11][n] * n2;
        T dv = T(1);
        if (lat) y = c1;
        if (NEED_DINV) {
            T d = lat ? T(1) : clamp_d<T>(op.cf[0][n]);
            dv = T(1) / d;
        }
        epi(k, n, y, c1, dv);
        c0 = c1; w0 = w1; e0 = e1; s0 = s1; n0 = n1;
        c1 = c2; w1 = w2; e1 = e2; s1 = s2; n1 = n2;
    }
    epi(g.mz - 1, (long)(g.mz - 1) * P + c, c1, c1, T(1));
}

template <typename T, bool NEED_DINV, class Epi>
__device__ __forceinline__ void march(const OpA<T> &opa, const OpP<T> &opp, bool probed, const Geo &g,
                                      const T *__restrict__ x, long c, int i, int j, Epi &epi)
{
    if (probed) march_probed<T, NEED_DINV>(opp, g, x, c, i, j, epi);
    else march_analytic<T, NEED_DINV>(opa, g, x, c, i, j, epi);
}

#define WINDHIP_COLUMN_PROLOGUE                                   \
    const int L = logical_block(g.nb);                            \
    const long c = (long)L * kBlock + threadIdx.x;                \
    const bool active = c < g.P;                                  \
    const int i = active ? (int)(c % g.mx) : 0;                   \
    const int j = active ? (int)(c / g.mx) : 0;

// ------------------------------------------------------------------ kernels
// y = A x (spmv)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv(OpA<T> opa, OpP<T> opp, int probed, Geo g,
                                                 const T *__restrict__ x, T *__restrict__ y)
{
    WINDHIP_COLUMN_PROLOGUE
    if (!active) return;
    auto epi = [&](int, long n, T v, T, T) { y[n] = v; };
    march<T, false>(opa, opp, probed, g, x, c, i, j, epi);
}

// One Richardson sweep of apply_precond (wind_iterative.F90:1300-1314):
// out = y + D^-1 (in - A y), y the current iterate.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sweep(OpA<T> opa, OpP<T> opp, int probed, Geo g,
                                                  const KState *st, const T *__restrict__ y,
                                                  const T *__restrict__ in, T *__restrict__ out)
{
    if (st->done) return;
    WINDHIP_COLUMN_PROLOGUE
    if (!active) return;
    auto epi = [&](int, long n, T ay, T yc, T dv) { out[n] = yc + dv * (in[n] - ay); };
    march<T, true>(opa, opp, probed, g, y, c, i, j, epi);
}

// Set-up of bicgstab_solve (wind_iterative.F90:586-640): t = A x0,
// r = b - t, r_hat = r, p = r, p_hat(sweep 1) = D^-1 r, rho0 = <r_hat, r>,
// ||r0||^2, ||b||^2.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_init(OpA<T> opa, OpP<T> opp, int probed, Geo g, KState *st,
                                                 double *partials, const T *__restrict__ x,
                                                 const T *__restrict__ b, const T *__restrict__ dinv,
                                                 T *__restrict__ r, T *__restrict__ rh,
                                                 T *__restrict__ p, T *__restrict__ ph)
{
    WINDHIP_COLUMN_PROLOGUE
    double acc[3] = {0.0, 0.0, 0.0};
    if (active) {
        auto epi = [&](int, long n, T t, T, T) {
            const T bv = b[n];
            const T rv = T(1) * bv + T(-1) * t;
            r[n] = rv; rh[n] = rv; p[n] = rv;
            ph[n] = dinv[n] * rv;
            acc[0] = acc[0] + (double)rv * (double)rv;
            acc[1] = acc[1] + (double)rv * (double)rv;
            acc[2] = acc[2] + (double)bv * (double)bv;
        };
        march<T, false>(opa, opp, probed, g, x, c, i, j, epi);
    }
    reduce_finalize<3>(acc, partials, &st->cnt[0], L, g.nb, [&](const double *tot) {
        st->rho = tot[0];
        st->rnorm = sqrt(tot[1]);
        st->res0 = st->rnorm;
        st->bnorm = sqrt(tot[2]);
        st->target = fmax(st->tol_abs, st->tol_rel * sqrt(tot[2]));
        st->iter = 0;
        st->omega = 1.0;
        st->alpha = 1.0;
        st->fmode = 0;
        if (st->rnorm <= st->target) { st->status = 0; st->done = 1; }
        else { st->status = 1; st->done = (st->max_iters <= 0); }
    });
}

// v = A p_hat and sigma = <r_hat, v> (wind_iterative.F90:657-688).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sigma(OpA<T> opa, OpP<T> opp, int probed, Geo g, KState *st,
                                                  double *partials, const T *__restrict__ ph,
                                                  const T *__restrict__ rh, T *__restrict__ v)
{
    if (st->done) {
        // clear the update mode left by the iteration that finished the solve
        if (blockIdx.x == 0 && threadIdx.x == 0) st->fmode = 0;
        return;
    }
    WINDHIP_COLUMN_PROLOGUE
    double acc[1] = {0.0};
    if (active) {
        auto epi = [&](int, long n, T y, T, T) {
            v[n] = y;
            acc[0] = acc[0] + (double)rh[n] * (double)y;
        };
        march<T, false>(opa, opp, probed, g, ph, c, i, j, epi);
    }
    reduce_finalize<1>(acc, partials, &st->cnt[1], L, g.nb, [&](const double *tot) {
        const int it = st->iter + 1;
        st->iter = it;
        st->sigma = tot[0];
        st->fmode = 0;
        if (fabs(tot[0]) < kBreakdownEps) { st->status = 2; st->done = 1; }
        else st->alpha = st->rho / tot[0];
    });
}

// s = r - alpha v (vec_axpby_into, :690) and the first sweep s_hat = D^-1 s.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_s(long N, const KState *st, const T *__restrict__ r,
                                                     const T *__restrict__ v, const T *__restrict__ dinv,
                                                     T *__restrict__ s, T *__restrict__ sh)
{
    if (st->done) return;
    const T na = T(-st->alpha);
    for (long n = (long)blockIdx.x * kBlock + threadIdx.x; n < N; n += (long)gridDim.x * kBlock) {
        const T sv = T(1) * r[n] + na * v[n];
        s[n] = sv;
        sh[n] = dinv[n] * sv;
    }
}

// t = A s_hat and the five fused dots <t,s>, <t,t>, <r_hat,s>, <r_hat,t>, <s,s>
// (wind_iterative.F90:703-785 up to the update of x).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_omega(OpA<T> opa, OpP<T> opp, int probed, Geo g, KState *st,
                                                  double *partials, const T *__restrict__ sh,
                                                  const T *__restrict__ s, const T *__restrict__ rh,
                                                  T *__restrict__ t)
{
    if (st->done) return;
    WINDHIP_COLUMN_PROLOGUE
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (active) {
        auto epi = [&](int, long n, T tv, T, T) {
            t[n] = tv;
            const double td = tv, sd = s[n], rd = rh[n];
            acc[0] = acc[0] + td * sd;
            acc[1] = acc[1] + td * td;
            acc[2] = acc[2] + rd * sd;
            acc[3] = acc[3] + rd * td;
            acc[4] = acc[4] + sd * sd;
        };
        march<T, false>(opa, opp, probed, g, sh, c, i, j, epi);
    }
    reduce_finalize<5>(acc, partials, &st->cnt[2], L, g.nb, [&](const double *tot) {
        const double ts = tot[0], tt = tot[1], rs = tot[2], rt = tot[3], ss = tot[4];
        st->ts = ts; st->tt = tt; st->rs = rs; st->rt = rt; st->ss = ss;
        if (tt < kBreakdownEps) { st->status = 2; st->done = 1; st->fmode = 0; return; }
        const double omega = ts / tt;
        const double rho_new = rs - omega * rt;
        double rn2 = (ss - (2.0 * omega) * ts) + (omega * omega) * tt;
        if (rn2 < 0.0) rn2 = 0.0;
        st->omega = omega;
        st->rnorm = sqrt(rn2);
        if (st->rnorm <= st->target) { st->status = 0; st->done = 1; st->fmode = 1; return; }
        if (fabs(omega) < kBreakdownEps) { st->status = 2; st->done = 1; st->fmode = 1; return; }
        if (fabs(st->rho) < kBreakdownEps) { st->status = 2; st->done = 1; st->fmode = 2; return; }
        st->beta = (rho_new / st->rho) * (st->alpha / omega);
        st->rho = rho_new;
        st->fmode = 3;
        if (st->iter >= st->max_iters) { st->status = 1; st->done = 1; }
        if (!(st->rnorm == st->rnorm) || st->rnorm > 1.0e300) { st->status = 2; st->done = 1; }
    });
}

// x += alpha p_hat + omega s_hat (vec_axpy2, :755); r = s - omega t (:771);
// p = r + beta (p - omega v) (vec_p_update, :783); next p_hat(sweep 1) = D^-1 p.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_xrp(long N, const KState *st, T *__restrict__ x,
                                                       const T *__restrict__ ph, const T *__restrict__ sh,
                                                       const T *__restrict__ s, const T *__restrict__ t,
                                                       T *__restrict__ r, T *__restrict__ p,
                                                       const T *__restrict__ v, const T *__restrict__ dinv,
                                                       T *__restrict__ ph_next)
{
    const int mode = st->fmode;
    if (mode == 0) return;
    const T a = T(st->alpha), om = T(st->omega), be = T(st->beta), nom = T(-st->omega);
    for (long n = (long)blockIdx.x * kBlock + threadIdx.x; n < N; n += (long)gridDim.x * kBlock) {
        x[n] = (x[n] + a * ph[n]) + om * sh[n];
        if (mode >= 2) {
            const T rv = T(1) * s[n] + nom * t[n];
            r[n] = rv;
            if (mode >= 3) {
                const T pv = rv + be * (p[n] - om * v[n]);
                p[n] = pv;
                ph_next[n] = dinv[n] * pv;
            }
        }
    }
}

// ----------------------------------------------- FP64 refinement around FP32
// The FP32 BiCGStab runs on the residual equation of the FP64 iterate x64:
// its own iterate y32 is flushed into x64 (k_ir_accumulate) and its recursive
// residual r32 is replaced by the true residual b - A x64, computed here in
// FP64, whenever the recursive norm has fallen by the refinement factor
// (a reliable update). `restart` also resets the shadow residual and the search
// direction (r_hat = p = r), which a breakdown or a finished inner solve needs;
// otherwise p and r_hat are kept and rho = <r_hat, r> is recomputed.
// Sums: 0 ||r64||^2, 1 ||b||^2, 2 <r32, r32>, 3 <r_hat32, r32>.
__global__ __launch_bounds__(kBlock) void k_ir_resid(OpA<double> opa, OpP<double> opp, int probed, Geo g,
                                                     KState *st_out, KState *st_in, double *partials, int restart,
                                                     const double *__restrict__ x, const double *__restrict__ b,
                                                     const float *__restrict__ dinv32, float *__restrict__ r32,
                                                     float *__restrict__ rh32, float *__restrict__ p32,
                                                     float *__restrict__ ph32)
{
    WINDHIP_COLUMN_PROLOGUE
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
        auto epi = [&](int, long n, double t, double, double) {
            const double bv = b[n];
            const double rv = bv - t;
            const float rf = (float)rv;
            r32[n] = rf;
            float rhv;
            if (restart) {
                rh32[n] = rf; p32[n] = rf;
                ph32[n] = dinv32[n] * rf;
                rhv = rf;
            } else {
                rhv = rh32[n];
            }
            acc[0] = acc[0] + rv * rv;
            acc[1] = acc[1] + bv * bv;
            acc[2] = acc[2] + (double)rf * (double)rf;
            acc[3] = acc[3] + (double)rhv * (double)rf;
        };
        march<double, false>(opa, opp, probed, g, x, c, i, j, epi);
    }
    reduce_finalize<4>(acc, partials, &st_out->cnt[3], L, g.nb, [&](const double *tot) {
        st_out->outer_rnorm = sqrt(tot[0]);
        st_out->outer_bnorm = sqrt(tot[1]);
        st_out->outer_target = fmax(st_out->tol_abs, st_out->tol_rel * st_out->outer_bnorm);
        st_out->outer_done = st_out->outer_rnorm <= st_out->outer_target;
        st_in->rho = tot[3];
        st_in->rnorm = sqrt(tot[2]);
        st_in->res0 = st_in->rnorm;
        st_in->bnorm = st_out->outer_bnorm;
        // the recursion aims a little below the FP64 target so that the true
        // residual passes at the check that follows
        st_in->target = 0.5 * st_out->outer_target;
        if (restart) { st_in->omega = 1.0; st_in->alpha = 1.0; }
        st_in->fmode = 0;
        if (st_out->outer_done) { st_in->status = 0; st_in->done = 1; }
        else { st_in->status = 1; st_in->done = (st_in->iter >= st_in->max_iters); }
    });
}

// x += d (FP64 accumulation of the FP32 correction), d = 0 for the next inner solve.
__global__ __launch_bounds__(kBlock) void k_ir_accumulate(long N, double *__restrict__ x, float *__restrict__ d)
{
    for (long n = (long)blockIdx.x * kBlock + threadIdx.x; n < N; n += (long)gridDim.x * kBlock) {
        x[n] = x[n] + (double)d[n];
        d[n] = 0.0f;
    }
}

// ||b - A x||^2 in FP64 (true residual), stored as st->outer_rnorm.
__global__ __launch_bounds__(kBlock) void k_true_resid(OpA<double> opa, OpP<double> opp, int probed, Geo g,
                                                       KState *st, double *partials,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ b)
{
    WINDHIP_COLUMN_PROLOGUE
    double acc[2] = {0.0, 0.0};
    if (active) {
        auto epi = [&](int, long n, double t, double, double) {
            const double bv = b[n];
            const double rv = bv - t;
            acc[0] = acc[0] + rv * rv;
            acc[1] = acc[1] + bv * bv;
        };
        march<double, false>(opa, opp, probed, g, x, c, i, j, epi);
    }
    reduce_finalize<2>(acc, partials, &st->cnt[3], L, g.nb, [&](const double *tot) {
        st->outer_rnorm = sqrt(tot[0]);
        st->outer_bnorm = sqrt(tot[1]);
    });
}

__global__ void k_state_reset(KState *st, int max_iters, double tol_rel, double tol_abs, double inner_tol)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        KState z = {};
        z.max_iters = max_iters;
        z.tol_rel = tol_rel;
        z.tol_abs = tol_abs;
        z.inner_tol = inner_tol;
        z.status = 1;
        *st = z;
    }
}

}  // namespace windhip
