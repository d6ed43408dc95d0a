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
// Work decomposition of the stencil kernels: a block is a tile of 8
// wavefronts x 64 lanes, one (i,j) column per thread; each thread marches its
// column bottom to top through the levels k. Lanes 1..62 of waves 1..6 are the
// tile's inner columns, the rest its halo (columns of the neighbouring tiles,
// recomputed where a fused kernel needs them). Unit-stride loads across a
// wavefront; blocks are remapped so each XCD processes a contiguous range of
// tiles (the j +- 1 rows of a tile then sit in the same XCD's L2). The fused
// sweep + matvec kernels stage the swept plane in an LDS ring and software-
// pipeline their global loads one level ahead through buffer instructions.
//
// Arithmetic contract: every expression rounds exactly as the reference writes
// it (left-to-right sums, no FMA: the file is compiled with -ffp-contract=off),
// and every dot product is summed in the canonical order
//   per column sum over k (from 0) -> 64-lane xor butterfly per wave ->
//   waves summed in order per tile -> lane t of the last block folds tile
//   partials t, t+512, ... -> the same block reduction.
// The CPU oracle sums in this same order, so the FP64 solve is bit-identical to
// it iteration by iteration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace windhip {

constexpr int kBlock = 256;          // elementwise kernels
constexpr int kWaves = 8;            // stencil kernels: wavefronts per block = tile rows incl. halo
constexpr int kThreads = 64 * kWaves;
constexpr int kTileI = 62;           // inner columns of a tile row (64 lanes less the halo pair)
constexpr int kTileJ = kWaves - 2;   // inner rows of a tile
constexpr double kBreakdownEps = 1.0e-30;  // wind_iterative.F90:61

// Tiling of the (i,j) columns: tile (ti,tj) covers lanes l = 0..63 of wave w =
// 0..7 at column i = 62 ti + l - 1, j = 6 tj + w - 1; lanes 1..62 of waves 1..6
// are its inner columns (every column is inner to exactly one tile), the rest
// its halo.
struct Geo {
    int mx, my, mz;
    long P;        // columns
    int nti, ntj;  // tiles in i and j
    int nt;        // tiles
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

// Profiling: whether the next launch of a solve does work, as it sees the
// state: bit 0 the solve is done, bit 1 an x/r/p update mode is set.
__global__ void k_prof_gate(const KState *st, int *out)
{
    if (threadIdx.x == 0) *out = (st->done ? 1 : 0) | (st->fmode ? 2 : 0);
}

// Three-level window of the five columns C, W, E, S, N: levels k-1, k, k+1.
template <typename T>
struct Win {
    T c0, c1, c2, w0, w1, w2, e0, e1, e2, s0, s1, s2, n0, n1, n2;
    __device__ __forceinline__ void shift()
    {
        c0 = c1; w0 = w1; e0 = e1; s0 = s1; n0 = n1;
        c1 = c2; w1 = w2; e1 = e2; s1 = s2; n1 = n2;
    }
};

template <typename T>
__device__ __forceinline__ T clamp_d(T d)
{
    if (fabs((double)d) < kBreakdownEps) d = copysign(T(kBreakdownEps), d);
    return d;
}

// Element e of the plane that starts at base, addressed as a uniform 64-bit
// base plus a 32-bit byte offset, which maps onto the saddr + voffset form of
// the global load/store instructions (no per-lane 64-bit address arithmetic).
template <typename T>
__device__ __forceinline__ const T &at(const T *base, unsigned e)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + e * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T &at(T *base, unsigned e)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + e * (unsigned)sizeof(T));
}

// Buffer-resource access for the streaming kernels: one SGPR resource per
// array, a 32-bit per-lane byte offset (voffset, shared by every array of the
// same layout) and a uniform 32-bit byte offset of the level (soffset). The
// host guarantees that every row array spans less than 4 GiB.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 family: 32-bit data format, raw (stride 0) buffer

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, long bytes)
{
    // base and size are uniform by construction; readfirstlane states it, so a
    // value the compiler happened to keep in VGPRs cannot turn every buffer
    // access into a per-lane waterfall loop
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    const int nr = __builtin_amdgcn_readfirstlane((int)(unsigned)bytes);
    void *base = (void *)(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, nr, kRsrcWord3);
}
// the level offset is uniform by construction; state it (free when the value
// already sits in an SGPR)
__device__ __forceinline__ int uni(unsigned so) { return __builtin_amdgcn_readfirstlane((int)so); }
template <typename T>
__device__ __forceinline__ T bload(rsrc_t r, unsigned vo, unsigned so);
template <>
__device__ __forceinline__ float bload<float>(rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, uni(so), 0));
}
template <>
__device__ __forceinline__ double bload<double>(rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, uni(so), 0));
}
__device__ __forceinline__ void bstore(float v, rsrc_t r, unsigned vo, unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo, uni(so), 0);
}
__device__ __forceinline__ void bstore(double v, rsrc_t r, unsigned vo, unsigned so)
{
    using u2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)vo, uni(so), 0);
}

enum RowClass { kRowBottom, kRowInterior, kRowTop };

// Analytic operator (spmv analytic branch, wind_iterative.F90:912-1025) in
// compressed form: J=-H, K=-I, N=-L, O=-M, D=E=F=G=1/dx^2, A=-4/dx^2-B-C.
template <typename T>
struct OpA {
    using type = T;
    const T *B, *C, *H, *I, *L, *M;   // rows
    const T *bcx, *bcy, *bcz, *dztop; // planes: bottom-row dzdx/(denom 2dx), dzdy/(...), 1/dz_if(1); top dz_if(nz+1)
    T inv, neg4;                      // 1/dx^2, -4/dx^2
    struct Cf { T B, C, H, I, L, M; };
    struct Bc { T bx, by, bz, dzt; };
    __device__ __forceinline__ Cf load(long n) const { return Cf{B[n], C[n], H[n], I[n], L[n], M[n]}; }
    // buffer resources of the six arrays; row = level offset lo (so bytes) + column e (vo bytes)
    struct Src { rsrc_t B, C, H, I, L, M; };
    __device__ __forceinline__ Src src(long N) const
    {
        const long nb = N * (long)sizeof(T);
        return Src{make_rsrc(B, nb), make_rsrc(C, nb), make_rsrc(H, nb),
                   make_rsrc(I, nb), make_rsrc(L, nb), make_rsrc(M, nb)};
    }
    __device__ __forceinline__ Cf load_src(const Src &r, long, unsigned so, unsigned, unsigned vo) const
    {
        return Cf{bload<T>(r.B, vo, so), bload<T>(r.C, vo, so), bload<T>(r.H, vo, so),
                  bload<T>(r.I, vo, so), bload<T>(r.L, vo, so), bload<T>(r.M, vo, so)};
    }
    __device__ __forceinline__ Bc bc(long c) const { return Bc{bcx[c], bcy[c], bcz[c], dztop[c]}; }
    // (A x)(row) of a row of class RC (kRowBottom: ghost plane k = 0,
    // kRowInterior, kRowTop: ghost plane k = mz-1). FP64 evaluates the
    // reference's sum term by term in its order (bit-identical to the oracle).
    // FP32 evaluates the same operator in difference form: every row of A sums
    // to zero (A = -4/dx^2 - B - C, J = -H, ...), so A x is a combination of
    // neighbour differences, whose rounding is relative to the differences
    // rather than to |x|; the smooth, slowly converging modes keep their
    // accuracy in FP32.
    template <int RC>
    __device__ __forceinline__ T row_c(bool lat, const Win<T> &x, const Cf &f, const Bc &b) const
    {
        if (lat) return x.c1;
        if constexpr (sizeof(T) == 8) {
            if constexpr (RC == kRowBottom)  // bottom BC row, wind_iterative.F90:962-988
                return b.by * x.s1 + b.by * x.s2 + b.bx * x.w1 + (-b.bz) * x.c1 + (-b.bx) * x.e1 + b.bx * x.w2
                       + b.bz * x.c2 + (-b.bx) * x.e2 + (-b.by) * x.n1 + (-b.by) * x.n2;
            else if constexpr (RC == kRowTop)
                return (x.c1 - x.c0) / b.dzt;  // top BC row, :950-960
            else {
                const T a = (neg4 - f.B) - f.C;
                // interior 15-point row, :1005-1020
                return (-f.M) * x.s0 + inv * x.s1 + f.M * x.s2 + (-f.I) * x.w0 + f.C * x.c0 + (-f.H) * x.e0
                       + inv * x.w1 + a * x.c1 + inv * x.e1 + f.I * x.w2 + f.B * x.c2 + f.H * x.e2
                       + (-f.L) * x.n0 + inv * x.n1 + f.L * x.n2;
            }
        } else {
            if constexpr (RC == kRowBottom)
                return b.bz * (x.c2 - x.c1) + b.bx * ((x.w1 + x.w2) - (x.e1 + x.e2))
                       + b.by * ((x.s1 + x.s2) - (x.n1 + x.n2));
            else if constexpr (RC == kRowTop)
                return (x.c1 - x.c0) / b.dzt;
            else
                return inv * (((x.w1 - x.c1) + (x.e1 - x.c1)) + ((x.s1 - x.c1) + (x.n1 - x.c1)))
                       + f.B * (x.c2 - x.c1) + f.C * (x.c0 - x.c1) + f.H * (x.e2 - x.e0)
                       + f.I * (x.w2 - x.w0) + f.L * (x.n2 - x.n0) + f.M * (x.s2 - x.s0);
        }
    }
    __device__ __forceinline__ T row(int k, int mz, bool lat, const Win<T> &x, const Cf &f, const Bc &b) const
    {
        if (k == 0) return row_c<kRowBottom>(lat, x, f, b);
        if (k == mz - 1) return row_c<kRowTop>(lat, x, f, b);
        return row_c<kRowInterior>(lat, x, f, b);
    }
    // 1/diag(A)(row) as extract_diagonal (wind_iterative.F90:1041-1075) computes it
    template <int RC>
    __device__ __forceinline__ T dinv_c(bool lat, const Cf &f, const Bc &b) const
    {
        T d;
        if (lat) d = T(1);
        else if constexpr (RC == kRowTop) d = clamp_d<T>(T(1) / b.dzt);
        else if constexpr (RC == kRowBottom) d = clamp_d<T>(-b.bz);
        else d = clamp_d<T>((neg4 - f.B) - f.C);
        return T(1) / d;
    }
    __device__ __forceinline__ T dinv(int k, int mz, bool lat, const Cf &f, const Bc &b) const
    {
        if (k == 0) return dinv_c<kRowBottom>(lat, f, b);
        if (k == mz - 1) return dinv_c<kRowTop>(lat, f, b);
        return dinv_c<kRowInterior>(lat, f, b);
    }
    // The vertical line of row k (coefficients of x(k-1), x(k), x(k+1)): the
    // tridiagonal part of A used by the line relaxation.
    __device__ __forceinline__ void tri(int k, int mz, bool lat, const Cf &f, const Bc &b, T &l, T &dd, T &u) const
    {
        if (lat) { l = T(0); dd = T(1); u = T(0); return; }
        if (k == 0) { l = T(0); dd = -b.bz; u = b.bz; return; }
        if (k == mz - 1) { l = T(-1) / b.dzt; dd = T(1) / b.dzt; u = T(0); return; }
        l = f.C; dd = (neg4 - f.B) - f.C; u = f.B;
    }
    // one Richardson update y + D^-1 (in - A y) (apply_precond, :1300-1314)
    __device__ __forceinline__ T relax(T yc, T dv, T inv_, T ay) const
    {
        return yc + dv * (inv_ - ay);
    }
};

// The analytic FP32 operator with its six coefficient rows stored in bf16 (the
// FP32 rows rounded to nearest even; FP32's range, 8 significant bits) and the
// arithmetic of OpA<float>: the operator of the multigrid-preconditioned FP32
// inner solve, whose 1e-1 inner tolerance is far above the 4e-3 relative
// rounding, at 12 instead of 24 bytes of coefficients per cell. The FP64 outer
// solve and its residual keep the FP64 operator. Byte offsets passed in are
// those of the FP32 rows (4 bytes per element) and are scaled here (x3 for the
// 12-byte interleaved rows, /2 for a plain bf16 row).
struct OpAB : OpA<float> {
    // the six coefficient rows of a level row interleaved as bf16 pairs, 12
    // bytes per row: {B | C << 16, H | I << 16, L | M << 16}
    const unsigned *Q;
    static __device__ __forceinline__ float bf(unsigned short h) { return __builtin_bit_cast(float, (unsigned)h << 16); }
    static __device__ __forceinline__ float lo(unsigned w) { return __builtin_bit_cast(float, w << 16); }
    static __device__ __forceinline__ float hi(unsigned w) { return __builtin_bit_cast(float, w & 0xFFFF0000u); }
    __device__ __forceinline__ Cf load(long n) const
    {
        const unsigned *p = Q + 3 * n;
        return Cf{lo(p[0]), hi(p[0]), lo(p[1]), hi(p[1]), lo(p[2]), hi(p[2])};
    }
    // one bf16 element of a plain bf16 row (the line solve's factor rows)
    static __device__ __forceinline__ float ld(rsrc_t r, unsigned vo, unsigned so)
    {
        return bf(__builtin_amdgcn_raw_buffer_load_b16(r, (int)(vo >> 1), uni(so >> 1), 0));
    }
    struct Src { rsrc_t q; };
    __device__ __forceinline__ Src src(long N) const { return Src{make_rsrc(Q, N * 12)}; }
    // one 12-byte load per lane: the row's offset is 3x its FP32 offset
    __device__ __forceinline__ Cf load_src(const Src &r, long, unsigned so, unsigned, unsigned vo) const
    {
        typedef unsigned u32x3 __attribute__((ext_vector_type(3)));
        const u32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r.q, (int)(vo * 3u), uni(so * 3u), 0);
        return Cf{lo(v.x), hi(v.x), lo(v.y), hi(v.y), lo(v.z), hi(v.z)};
    }
};

// FP32 -> bf16, round to nearest even (finite values)
__device__ __forceinline__ unsigned bf16_rne(float v)
{
    const unsigned u = __builtin_bit_cast(unsigned, v);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__global__ void k_to_bf16(const float *__restrict__ src, unsigned short *__restrict__ dst, long n)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        dst[e] = (unsigned short)bf16_rne(src[e]);
}
// two FP32 rows -> one row of bf16 pairs {a | b << 16}
__global__ void k_to_bf16_pair(const float *__restrict__ a, const float *__restrict__ b, unsigned *__restrict__ dst,
                               long n)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        dst[e] = bf16_rne(a[e]) | (bf16_rne(b[e]) << 16);
}
// the six FP32 coefficient rows -> the interleaved bf16 rows of OpAB
__global__ void k_to_bf16_rows(const float *__restrict__ B, const float *__restrict__ C, const float *__restrict__ H,
                               const float *__restrict__ I, const float *__restrict__ L, const float *__restrict__ M,
                               unsigned *__restrict__ Q, long n)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        Q[3 * e] = bf16_rne(B[e]) | (bf16_rne(C[e]) << 16);
        Q[3 * e + 1] = bf16_rne(H[e]) | (bf16_rne(I[e]) << 16);
        Q[3 * e + 2] = bf16_rne(L[e]) | (bf16_rne(M[e]) << 16);
    }
}

// Probed 15-array operator (spmv probed branch, wind_iterative.F90:837-910):
// lateral and ghost-plane identity rows.
template <typename T>
struct OpP {
    using type = T;
    const T *cf[15];  // A..O, rows
    struct Cf { T a[15]; };
    struct Bc {};
    __device__ __forceinline__ Cf load(long n) const
    {
        Cf f;
#pragma unroll
        for (int q = 0; q < 15; ++q) f.a[q] = cf[q][n];
        return f;
    }
    struct Src {};
    __device__ __forceinline__ Src src(long) const { return Src{}; }
    __device__ __forceinline__ Cf load_src(const Src &, long lo, unsigned, unsigned e, unsigned) const
    {
        Cf f;
#pragma unroll
        for (int q = 0; q < 15; ++q) f.a[q] = at(cf[q] + lo, e);
        return f;
    }
    __device__ __forceinline__ Bc bc(long) const { return Bc{}; }
    template <int RC>
    __device__ __forceinline__ T row_c(bool lat, const Win<T> &x, const Cf &f, const Bc &bc) const
    {
        if constexpr (RC != kRowInterior) return x.c1;
        else return row(1, 3, lat, x, f, bc);
    }
    template <int RC>
    __device__ __forceinline__ T dinv_c(bool lat, const Cf &f, const Bc &bc) const
    {
        if constexpr (RC != kRowInterior) return T(1);
        else return dinv(1, 3, lat, f, bc);
    }
    __device__ __forceinline__ T relax(T yc, T dv, T inv_, T ay) const { return yc + dv * (inv_ - ay); }
    __device__ __forceinline__ void tri(int k, int mz, bool lat, const Cf &f, const Bc &, T &l, T &dd, T &u) const
    {
        if (lat || k == 0 || k == mz - 1) { l = T(0); dd = T(1); u = T(0); return; }
        l = f.a[2]; dd = f.a[0]; u = f.a[1];  // C, A, B
    }

    __device__ __forceinline__ T row(int k, int mz, bool lat, const Win<T> &x, const Cf &f, const Bc &) const
    {
        if (lat || k == 0 || k == mz - 1) return x.c1;
        // wind_iterative.F90:888-904 (A..O = 0..14)
        return f.a[14] * x.s0 + f.a[6] * x.s1 + f.a[12] * x.s2 + f.a[10] * x.w0 + f.a[2] * x.c0 + f.a[9] * x.e0
               + f.a[4] * x.w1 + f.a[0] * x.c1 + f.a[3] * x.e1 + f.a[8] * x.w2 + f.a[1] * x.c2 + f.a[7] * x.e2
               + f.a[13] * x.n0 + f.a[5] * x.n1 + f.a[11] * x.n2;
    }
    __device__ __forceinline__ T dinv(int k, int mz, bool lat, const Cf &f, const Bc &) const
    {
        if (lat || k == 0 || k == mz - 1) return T(1);
        return T(1) / clamp_d<T>(f.a[0]);
    }
};

__device__ __forceinline__ int logical_block(int nb)
{
    // Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
    // contiguous logical range [start(x), start(x) + count(x)) so neighbouring
    // tiles share an L2.
    int b = blockIdx.x;
    int x = b & 7, r = b >> 3;
    int q = nb >> 3, rem = nb & 7;
    return x * q + (x < rem ? x : rem) + r;
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

// Canonical reduction of NV per-thread values over a block of NW waves: 64-lane
// xor butterfly per wave, waves summed in order; block partials are folded by
// the last block to finish (lane t sums partials t, t+64 NW, ... from 0),
// reduced the same way, and thread 0 of that block calls fin(total). The
// partials move through agent-scope atomic stores and loads, the arrival
// through an agent-scope counter.
template <int NV, int NW = kWaves, class Fin>
__device__ __forceinline__ void reduce_finalize(double (&acc)[NV], double *partials, unsigned *counter, int L,
                                                int nb, Fin fin)
{
    __shared__ double red[NW][NV];
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
        for (int v = 0; v < NV; ++v) {
            double s = red[0][v];
            for (int q = 1; q < NW; ++q) s = s + red[q][v];
            // written through to the device-coherent level (an agent-scope
            // store); a __threadfence() here would be a release of every
            // store of the kernel, i.e. a write-back of the XCD's whole L2
            __hip_atomic_store(&partials[(long)L * NV + v], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);  // the partials have landed before the counter moves
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        unsigned old = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (old == (unsigned)(nb - 1));
    }
    __syncthreads();
    if (!is_last) return;
    double fold[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) fold[v] = 0.0;
    // a batch of FB partials is loaded before any of them is added (each add
    // right after its load waited one memory round trip per partial); the
    // adds keep the order q = t, t + 64 NW, ...
    constexpr int FB = NV == 1 ? 8 : (NV <= 2 ? 4 : (NV < 8 ? 2 : 1));
    constexpr int S = NW * 64;
    for (int q0 = threadIdx.x; q0 < nb; q0 += FB * S) {
        double pv[FB][NV];
#pragma unroll
        for (int e = 0; e < FB; ++e)
#pragma unroll
            for (int v = 0; v < NV; ++v)
                pv[e][v] = q0 + e * S < nb ? ld_coherent(&partials[(long)(q0 + e * S) * NV + v]) : 0.0;
#pragma unroll
        for (int e = 0; e < FB; ++e)
            if (q0 + e * S < nb)
#pragma unroll
                for (int v = 0; v < NV; ++v) fold[v] = fold[v] + pv[e][v];
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
        for (int v = 0; v < NV; ++v) {
            double s = red[0][v];
            for (int q = 1; q < NW; ++q) s = s + red[q][v];
            tot[v] = s;
        }
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fin(tot);
    }
}

// Position of this thread in its tile.
struct TilePos {
    int L, lane, w, i, j;
    bool valid, inner, lat;
    long c, cw, ce, cs, cn;
    unsigned oc, ow, oe, os, on;  // the same column offsets in 32 bits (P < 2^31)
};

__device__ __forceinline__ TilePos tile_pos(const Geo &g)
{
    TilePos t;
    t.L = logical_block(g.nt);
    const int ti = t.L % g.nti, tj = t.L / g.nti;
    t.lane = threadIdx.x & 63;
    t.w = threadIdx.x >> 6;
    t.i = kTileI * ti + t.lane - 1;
    t.j = kTileJ * tj + t.w - 1;
    t.valid = t.i >= 0 && t.i < g.mx && t.j >= 0 && t.j < g.my;
    t.inner = t.valid && t.lane >= 1 && t.lane <= kTileI && t.w >= 1 && t.w <= kTileJ;
    t.lat = t.valid && (t.i == 0 || t.i == g.mx - 1 || t.j == 0 || t.j == g.my - 1);
    t.c = t.valid ? (long)t.j * g.mx + t.i : 0;
    const bool clampn = !t.valid || t.lat;
    t.cw = clampn ? t.c : t.c - 1;
    t.ce = clampn ? t.c : t.c + 1;
    t.cs = clampn ? t.c : t.c - g.mx;
    t.cn = clampn ? t.c : t.c + g.mx;
    t.oc = (unsigned)t.c; t.ow = (unsigned)t.cw; t.oe = (unsigned)t.ce;
    t.os = (unsigned)t.cs; t.on = (unsigned)t.cn;
    return t;
}


template <typename T>
__device__ __forceinline__ void load_level(const T *__restrict__ x, long o, const TilePos &t, T &c, T &w, T &e,
                                           T &s, T &n)
{
    c = x[o + t.c]; w = x[o + t.cw]; e = x[o + t.ce]; s = x[o + t.cs]; n = x[o + t.cn];
}

// ------------------------------------------------------------------ the march
// Visits the rows of the thread's column bottom to top and calls
// epi(k, n, y, xc, dinv) with y = (A x)(row), xc = x(row) and dinv = D^-1(row)
// when NEED_DINV. x is read from global memory; the loads of level k+2 are
// issued before level k is evaluated.
template <class Op, bool NEED_DINV, class Epi>
__device__ __forceinline__ void march(const Op &op, const Geo &g, const typename Op::type *__restrict__ x,
                                      const TilePos &t, Epi &epi)
{
    using T = typename Op::type;
    const long P = g.P;
    const int mz = g.mz;
    Win<T> win;
    load_level(x, 0, t, win.c1, win.w1, win.e1, win.s1, win.n1);
    load_level(x, P, t, win.c2, win.w2, win.e2, win.s2, win.n2);
    const typename Op::Bc bc = op.bc(t.c);
    typename Op::Cf cf = op.load(P + t.c);  // level 1; the ghost planes use the BC planes only
    for (int k = 0; k < mz; ++k) {
        T nc, nw, ne, ns, nn;
        if (k + 2 <= mz - 1) load_level(x, (long)(k + 2) * P, t, nc, nw, ne, ns, nn);
        typename Op::Cf cn;
        if (k >= 1 && k + 1 <= mz - 2) cn = op.load((long)(k + 1) * P + t.c);
        const T yv = op.row(k, mz, t.lat, win, cf, bc);
        const T dv = NEED_DINV ? op.dinv(k, mz, t.lat, cf, bc) : T(1);
        epi(k, (long)k * P + t.c, yv, win.c1, dv);
        win.shift();
        if (k + 2 <= mz - 1) { win.c2 = nc; win.w2 = nw; win.e2 = ne; win.s2 = ns; win.n2 = nn; }
        if (k >= 1 && k + 1 <= mz - 2) cf = cn;
    }
}

// Barrier for the LDS ring: waits for this wave's LDS writes only, so global
// loads issued for later levels stay in flight across it (a full
// __syncthreads() would drain them at every level).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Two stages in one pass: stage 1 is a Richardson sweep out1 = y + D^-1 (in - A y)
// over every column of the tile (halo included), staged plane by plane in an
// LDS ring; stage 2 evaluates (A out1)(row) one level behind on the inner
// columns from that ring and hands it to epi2. Each coefficient is read from
// HBM once for both stages.
//
// Step m evaluates stage 1 at level m and stage 2 at level m-1, m = 0..mz.
// Its operands (x at level m+1, coefficients and `in` at level m, epilogue
// operands of level m-1) form one Slot, loaded one step ahead into one of two
// register sets that alternate with the step, so no loaded register is copied
// while its load is in flight. The steps that touch a ghost plane (0, 1, mz-1,
// mz) are peeled out of the loop; the loop body is the interior stencil only.
// All global accesses are buffer loads/stores: one resource per array, the
// lane's column as the 32-bit voffset, the level (and for x the neighbour
// direction) as the uniform soffset.
enum StepPos { kFirst, kSecond, kMid, kTop, kLast };

// Register copy the register allocator cannot coalesce away. Used for values
// taken out of a pipeline slot that outlive the slot's refill: if such a copy
// were merged with the slot register, the refill would land in another
// register and be copied back at the loop back-edge while its load is still
// in flight, which waits for every outstanding load.
__device__ __forceinline__ float keep(float v)
{
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ double keep(double v)
{
    double r;
    asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
template <typename T, class Cf>
__device__ __forceinline__ Cf keep_cf(const Cf &c)
{
    Cf r;
    constexpr int n = sizeof(Cf) / sizeof(T);
    const T *src = reinterpret_cast<const T *>(&c);
    T *dst = reinterpret_cast<T *>(&r);
#pragma unroll
    for (int q = 0; q < n; ++q) dst[q] = keep(src[q]);
    return r;
}

template <typename T>
__device__ __forceinline__ void load_x5(rsrc_t r, unsigned vo, unsigned sS, unsigned sW, unsigned sC, unsigned sE,
                                        unsigned sN, T &c, T &w, T &e, T &s, T &n)
{
    s = bload<T>(r, vo, sS); w = bload<T>(r, vo, sW); c = bload<T>(r, vo, sC);
    e = bload<T>(r, vo, sE); n = bload<T>(r, vo, sN);
}

template <class Op, class Epi2>
__device__ __forceinline__ void fused_sweep_matvec(const Op &op, const Geo &g, const TilePos &t,
                                                   const typename Op::type *__restrict__ y,
                                                   const typename Op::type *__restrict__ in,
                                                   typename Op::type *__restrict__ out1, Epi2 &epi2)
{
    using T = typename Op::type;
    using Cf = typename Op::Cf;
    using Pre = typename Epi2::Pre;
    struct Slot {
        T c, w, e, s, n;
        Cf cf;
        T in;
        Pre ep;
    };
    __shared__ T ring[4][kWaves][64];
    // grid scalars pinned to SGPRs (they are uniform; see make_rsrc)
    const int mx = __builtin_amdgcn_readfirstlane(g.mx);
    const int mz = __builtin_amdgcn_readfirstlane(g.mz);
    const long P = (long)mx * __builtin_amdgcn_readfirstlane(g.my);
    const long N = P * mz;
    constexpr unsigned sz = sizeof(T);
    const unsigned pb = (unsigned)(P * (long)sz);  // bytes per level
    const unsigned voc = t.oc * sz;
    // x is addressed from y - (mx + 1): neighbour d of column c sits at byte
    // (c + d + mx + 1) * sz; the row arrays carry mx + 1 guard elements on
    // both sides, so every lane stays inside the allocation.
    const long gd = mx + 1;
    const unsigned oS = sz, oW = (unsigned)mx * sz, oC = (unsigned)(mx + 1) * sz,
                   oE = (unsigned)(mx + 2) * sz, oN = (unsigned)(2 * mx + 1) * sz;
    const rsrc_t ry = make_rsrc(y - gd, (N + 2 * gd) * (long)sz), rin = make_rsrc(in, N * (long)sz),
                 rout = make_rsrc(out1, N * (long)sz);
    const typename Op::Src rcf = op.src(N);
    epi2.setup(N);
    // operands of step m; levels are clamped into [0, mz-1]
    auto fill = [&](Slot &sl, int m, auto pos) {
        constexpr int ps = decltype(pos)::value;
        if (ps != kTop && ps != kLast) {
            const int ly = m + 1 < mz - 1 ? m + 1 : mz - 1;
            const unsigned so = (unsigned)ly * pb;
            load_x5(ry, voc, so + oS, so + oW, so + oC, so + oE, so + oN, sl.c, sl.w, sl.e, sl.s, sl.n);
        }
        if (ps == kSecond || ps == kMid) {
            const int lv = m < mz - 2 ? m : mz - 2;
            sl.cf = op.load_src(rcf, (long)lv * P, (unsigned)lv * pb, t.oc, voc);
        }
        if (ps != kLast) {
            const int lv = m < mz - 1 ? m : mz - 1;
            sl.in = bload<T>(rin, voc, (unsigned)lv * pb);
        }
        if (ps != kFirst) {
            const int ke = m - 1 < mz - 1 ? m - 1 : mz - 1;
            sl.ep = epi2.load((unsigned)ke * pb, voc);
        }
    };
    Win<T> win = {};
    Cf cf_prev = {};
    // boundary-condition planes of the column (ghost-plane steps only)
    typename Op::Bc bc = op.bc(t.c);
    // stage 1 at level m (ps: kFirst bottom row, kMid interior, kTop top row)
    // then stage 2 at level m-1 (kSecond bottom row, kMid/kTop interior, kLast top row)
    auto step = [&](const Slot &cur, int m, auto pos) {
        constexpr int ps = decltype(pos)::value;
        if (ps != kLast) {
            if (ps != kFirst) win.shift();
            win.c2 = keep(cur.c); win.w2 = keep(cur.w); win.e2 = keep(cur.e);
            win.s2 = keep(cur.s); win.n2 = keep(cur.n);
            T v1 = T(0);
            if (t.valid) {
                constexpr int rc = ps == kFirst ? kRowBottom : (ps == kTop ? kRowTop : kRowInterior);
                const T ay = op.template row_c<rc>(t.lat, win, cur.cf, bc);
                const T dv = op.template dinv_c<rc>(t.lat, cur.cf, bc);
                v1 = op.relax(win.c1, dv, cur.in, ay);
                if (t.inner) bstore(v1, rout, voc, (unsigned)m * pb);
            }
            ring[m & 3][t.w][t.lane] = v1;
        }
        lds_barrier();
        if (ps != kFirst && t.inner) {
            const int k = m - 1;
            Win<T> x2;
            const int p0 = (k + 3) & 3, p1 = k & 3, p2 = (k + 1) & 3;
            x2.c1 = ring[p1][t.w][t.lane];
            x2.w1 = ring[p1][t.w][t.lane - 1]; x2.e1 = ring[p1][t.w][t.lane + 1];
            x2.s1 = ring[p1][t.w - 1][t.lane]; x2.n1 = ring[p1][t.w + 1][t.lane];
            if (ps != kSecond) {
                x2.c0 = ring[p0][t.w][t.lane];
                x2.w0 = ring[p0][t.w][t.lane - 1]; x2.e0 = ring[p0][t.w][t.lane + 1];
                x2.s0 = ring[p0][t.w - 1][t.lane]; x2.n0 = ring[p0][t.w + 1][t.lane];
            } else {
                x2.c0 = x2.w0 = x2.e0 = x2.s0 = x2.n0 = T(0);
            }
            if (ps != kLast) {
                x2.c2 = ring[p2][t.w][t.lane];
                x2.w2 = ring[p2][t.w][t.lane - 1]; x2.e2 = ring[p2][t.w][t.lane + 1];
                x2.s2 = ring[p2][t.w - 1][t.lane]; x2.n2 = ring[p2][t.w + 1][t.lane];
            } else {
                x2.c2 = x2.w2 = x2.e2 = x2.s2 = x2.n2 = T(0);
            }
            constexpr int rc = ps == kSecond ? kRowBottom : (ps == kLast ? kRowTop : kRowInterior);
            const T y2 = op.template row_c<rc>(t.lat, x2, cf_prev, bc);
            epi2.apply(k, (unsigned)k * pb, voc, y2, x2.c1, cur.ep);
        }
        if (ps == kSecond || ps == kMid) cf_prev = keep_cf<T>(cur.cf);  // stage 2 of the next step evaluates this level
    };
    using PF = std::integral_constant<int, kFirst>;
    using PS = std::integral_constant<int, kSecond>;
    using PM = std::integral_constant<int, kMid>;
    using PT = std::integral_constant<int, kTop>;
    using PL = std::integral_constant<int, kLast>;
    Slot sa, sb;
    load_x5(ry, voc, oS, oW, oC, oE, oN, win.c1, win.w1, win.e1, win.s1, win.n1);
    fill(sa, 0, PF{});
    fill(sb, 1, PS{});
    step(sa, 0, PF{});
    // interior steps m = 2..mz-2 in pairs; the set holding step m alternates
    const int nmid = mz - 3;
    fill(sa, 2, PM{});
    step(sb, 1, PS{});
    int m = 2;
    // the scheduling barriers keep each set's refill after the last use of its
    // previous contents, so both live in the same registers
    for (; m + 1 <= mz - 2; m += 2) {
        __builtin_amdgcn_sched_barrier(0);
        fill(sb, m + 1, PM{});
        step(sa, m, PM{});
        __builtin_amdgcn_sched_barrier(0);
        fill(sa, m + 2, PM{});
        step(sb, m + 1, PM{});
    }
    __builtin_amdgcn_sched_barrier(0);
    if (nmid & 1) {  // one interior step left: step mz-1's operands go to sb
        fill(sb, m + 1, PT{});
        step(sa, m, PM{});
        fill(sa, mz, PL{});
        step(sb, mz - 1, PT{});
        step(sa, mz, PL{});
    } else {  // step mz-1's operands are in sa
        fill(sb, mz, PL{});
        step(sa, mz - 1, PT{});
        step(sb, mz, PL{});
    }
}

// ------------------------------------------------------------------ iteration scalars
// sigma = <r_hat, v>, alpha = rho / sigma (wind_iterative.F90:649-688)
__device__ __forceinline__ void finish_sigma(KState *st, const double *tot)
{
    const int it = st->iter + 1;
    st->iter = it;
    st->sigma = tot[0];
    st->fmode = 0;
    if (fabs(tot[0]) < kBreakdownEps) { st->status = 2; st->done = 1; }
    else st->alpha = st->rho / tot[0];
}

// omega, rho_new, ||r|| from <t,s>, <t,t>, <r_hat,s>, <r_hat,t>, <s,s>; the
// stop rule and breakdown guards (wind_iterative.F90:695-785)
__device__ __forceinline__ void finish_omega(KState *st, const double *tot)
{
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
}

// ------------------------------------------------------------------ kernels
// y = A x (spmv)
template <class Op>
__global__ __launch_bounds__(kThreads) void k_spmv(Op op, Geo g, const typename Op::type *__restrict__ x,
                                                   typename Op::type *__restrict__ y)
{
    using T = typename Op::type;
    const TilePos t = tile_pos(g);
    if (!t.inner) return;
    auto epi = [&](int, long n, T v, T, T) { y[n] = v; };
    march<Op, false>(op, g, x, t, epi);
}

// One Richardson sweep of apply_precond (wind_iterative.F90:1300-1314):
// out = y + D^-1 (in - A y), y the current iterate.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_sweep(Op op, Geo g, const KState *st,
                                                    const typename Op::type *__restrict__ y,
                                                    const typename Op::type *__restrict__ in,
                                                    typename Op::type *__restrict__ out)
{
    using T = typename Op::type;
    if (st->done) return;
    const TilePos t = tile_pos(g);
    if (!t.inner) return;
    auto epi = [&](int, long n, T ay, T yc, T dv) { out[n] = yc + dv * (in[n] - ay); };
    march<Op, true>(op, g, y, t, epi);
}

// Set-up of bicgstab_solve (wind_iterative.F90:586-640): t = A x0,
// r = b - t, r_hat = r, p = r, p_hat(sweep 1) = D^-1 r, rho0 = <r_hat, r>,
// ||r0||^2, ||b||^2.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_init(Op op, Geo g, KState *st, double *partials,
                                                   const typename Op::type *__restrict__ x,
                                                   const typename Op::type *__restrict__ b,
                                                   const typename Op::type *__restrict__ dinv,
                                                   typename Op::type *__restrict__ r,
                                                   typename Op::type *__restrict__ rh,
                                                   typename Op::type *__restrict__ p,
                                                   typename Op::type *__restrict__ ph)
{
    using T = typename Op::type;
    const TilePos t = tile_pos(g);
    double acc[3] = {0.0, 0.0, 0.0};
    if (t.inner) {
        auto epi = [&](int, long n, T tv, T, T) {
            const T bv = b[n];
            const T rv = T(1) * bv + T(-1) * tv;
            r[n] = rv; rh[n] = rv; p[n] = rv;
            ph[n] = dinv[n] * rv;
            acc[0] = acc[0] + (double)rv * (double)rv;
            acc[1] = acc[1] + (double)rv * (double)rv;
            acc[2] = acc[2] + (double)bv * (double)bv;
        };
        march<Op, false>(op, g, x, t, epi);
    }
    reduce_finalize<3>(acc, partials, &st->cnt[0], t.L, g.nt, [&](const double *tot) {
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

// Last preconditioner sweep on p fused with v = A p_hat and sigma = <r_hat, v>
// (wind_iterative.F90:649-688).
template <class Op>
__global__ __launch_bounds__(kThreads) void k_sweep_sigma(Op op, Geo g, KState *st, double *partials,
                                                          const typename Op::type *__restrict__ ph_in,
                                                          const typename Op::type *__restrict__ p,
                                                          typename Op::type *__restrict__ ph,
                                                          const typename Op::type *__restrict__ rh,
                                                          typename Op::type *__restrict__ v)
{
    using T = typename Op::type;
    if (st->done) {
        // clear the update mode left by the iteration that finished the solve
        if (blockIdx.x == 0 && threadIdx.x == 0) st->fmode = 0;
        return;
    }
    const TilePos t = tile_pos(g);
    struct Epi {
        const T *rh;
        T *v;
        double acc[1];
        rsrc_t rrh, rv;
        struct Pre { T rh; };
        __device__ void setup(long N) { rrh = make_rsrc(rh, N * (long)sizeof(T)); rv = make_rsrc(v, N * (long)sizeof(T)); }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<T>(rrh, vo, so)}; }
        __device__ void apply(int, unsigned so, unsigned vo, T yv, T, const Pre &p)
        {
            bstore(yv, rv, vo, so);
            acc[0] = acc[0] + (double)p.rh * (double)yv;
        }
    } epi{rh, v, {0.0}};
    fused_sweep_matvec(op, g, t, ph_in, p, ph, epi);
    double (&acc)[1] = epi.acc;
    reduce_finalize<1>(acc, partials, &st->cnt[1], t.L, g.nt, [&](const double *tot) { finish_sigma(st, tot); });
}

// s = r - alpha v (vec_axpby_into, :690) and the first sweep s_hat = D^-1 s.
// 16-byte vectors of the flat update kernels
template <typename T> struct Vec16;
template <> struct Vec16<double> { using type = double __attribute__((ext_vector_type(2))); static constexpr int n = 2; };
template <> struct Vec16<float> { using type = float __attribute__((ext_vector_type(4))); static constexpr int n = 4; };
template <typename... P>
__device__ __forceinline__ bool aligned16(const P *...p)
{
    return ((reinterpret_cast<unsigned long long>(p) | ...) & 15ull) == 0;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_s(long N, const KState *st, const T *__restrict__ r,
                                                     const T *__restrict__ v, const T *__restrict__ dinv,
                                                     T *__restrict__ s, T *__restrict__ sh)
{
    if (st->done) return;
    const T na = T(-st->alpha);
    const long tid = (long)blockIdx.x * kBlock + threadIdx.x, stride = (long)gridDim.x * kBlock;
    long n0 = 0;
    if (aligned16(r, v, dinv, s, sh)) {  // 16 bytes per lane and array
        using V = typename Vec16<T>::type;
        constexpr int W = Vec16<T>::n;
        n0 = N / W * W;
        for (long q = tid; q < N / W; q += stride) {
            const V rv = reinterpret_cast<const V *>(r)[q], vv = reinterpret_cast<const V *>(v)[q],
                    dv = reinterpret_cast<const V *>(dinv)[q];
            V sv, hv;
#pragma unroll
            for (int e = 0; e < W; ++e) {
                sv[e] = T(1) * rv[e] + na * vv[e];
                hv[e] = dv[e] * sv[e];
            }
            reinterpret_cast<V *>(s)[q] = sv;
            reinterpret_cast<V *>(sh)[q] = hv;
        }
    }
    for (long n = n0 + tid; n < N; n += stride) {
        const T sv = T(1) * r[n] + na * v[n];
        s[n] = sv;
        sh[n] = dinv[n] * sv;
    }
}

// Last sweep on s fused with t = A s_hat and the five dots <t,s>, <t,t>,
// <r_hat,s>, <r_hat,t>, <s,s> (wind_iterative.F90:695-785 up to the update of x).
template <class Op>
__global__ __launch_bounds__(kThreads) void k_sweep_omega(Op op, Geo g, KState *st, double *partials,
                                                          const typename Op::type *__restrict__ sh_in,
                                                          const typename Op::type *__restrict__ s,
                                                          typename Op::type *__restrict__ sh,
                                                          const typename Op::type *__restrict__ rh,
                                                          typename Op::type *__restrict__ tt_out)
{
    using T = typename Op::type;
    if (st->done) return;
    const TilePos t = tile_pos(g);
    struct Epi {
        const T *s, *rh;
        T *out;
        double acc[5];
        rsrc_t rs, rrh, rout;
        struct Pre { T s, rh; };
        __device__ void setup(long N)
        {
            const long nb = N * (long)sizeof(T);
            rs = make_rsrc(s, nb); rrh = make_rsrc(rh, nb); rout = make_rsrc(out, nb);
        }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<T>(rs, vo, so), bload<T>(rrh, vo, so)}; }
        __device__ void apply(int, unsigned so, unsigned vo, T tv, T, const Pre &p)
        {
            bstore(tv, rout, vo, so);
            const double td = tv, sd = p.s, rd = p.rh;
            acc[0] = acc[0] + td * sd;
            acc[1] = acc[1] + td * td;
            acc[2] = acc[2] + rd * sd;
            acc[3] = acc[3] + rd * td;
            acc[4] = acc[4] + sd * sd;
        }
    } epi{s, rh, tt_out, {0.0, 0.0, 0.0, 0.0, 0.0}};
    fused_sweep_matvec(op, g, t, sh_in, s, sh, epi);
    double (&acc)[5] = epi.acc;
    reduce_finalize<5>(acc, partials, &st->cnt[2], t.L, g.nt, [&](const double *tot) { finish_omega(st, tot); });
}

// x += alpha p_hat + omega s_hat (vec_axpy2, :755); r = s - omega t (:771);
// p = r + beta (p - omega v) (vec_p_update, :783); next p_hat(sweep 1) = D^-1 p
// (skipped when dinv is null: the line-preconditioned iteration has no use for it).
// FIRST: the first iteration of a solve from x = 0 whose p is still r_hat:
// x is not read (0 + ...), p is read from p_in (the same values, the same
// arithmetic) and written to p.
template <typename T, bool FIRST = false>
__global__ __launch_bounds__(kBlock) void k_update_xrp(long N, const KState *st, T *__restrict__ x,
                                                       const T *__restrict__ ph, const T *__restrict__ sh,
                                                       const T *__restrict__ s, const T *__restrict__ t,
                                                       T *__restrict__ r, T *__restrict__ p,
                                                       const T *__restrict__ v, const T *__restrict__ dinv,
                                                       T *__restrict__ ph_next, const T *__restrict__ p_in = nullptr)
{
    const int mode = st->fmode;
    if (mode == 0) return;
    const T a = T(st->alpha), om = T(st->omega), be = T(st->beta), nom = T(-st->omega);
    const long tid = (long)blockIdx.x * kBlock + threadIdx.x, stride = (long)gridDim.x * kBlock;
    const T *pr = FIRST ? p_in : p;
    long n0 = 0;
    if (mode >= 3 && aligned16(x, ph, sh, s, t, r, p, pr, v, dinv, ph_next)) {
        // the full update, 16 bytes per lane and array (the same per-element
        // arithmetic as the loop below)
        using V = typename Vec16<T>::type;
        constexpr int W = Vec16<T>::n;
        n0 = N / W * W;
        for (long q = tid; q < N / W; q += stride) {
            V xv;
            if constexpr (FIRST) {
#pragma unroll
                for (int e = 0; e < W; ++e) xv[e] = T(0);
            } else {
                xv = reinterpret_cast<const V *>(x)[q];
            }
            const V phv = reinterpret_cast<const V *>(ph)[q],
                    shv = reinterpret_cast<const V *>(sh)[q], sv = reinterpret_cast<const V *>(s)[q],
                    tv = reinterpret_cast<const V *>(t)[q], pv0 = reinterpret_cast<const V *>(pr)[q],
                    vv = reinterpret_cast<const V *>(v)[q];
            V xo, ro, po;
#pragma unroll
            for (int e = 0; e < W; ++e) {
                xo[e] = (xv[e] + a * phv[e]) + om * shv[e];
                ro[e] = T(1) * sv[e] + nom * tv[e];
                po[e] = ro[e] + be * (pv0[e] - om * vv[e]);
            }
            reinterpret_cast<V *>(x)[q] = xo;
            reinterpret_cast<V *>(r)[q] = ro;
            reinterpret_cast<V *>(p)[q] = po;
            if (dinv) {
                const V dv = reinterpret_cast<const V *>(dinv)[q];
                V ho;
#pragma unroll
                for (int e = 0; e < W; ++e) ho[e] = dv[e] * po[e];
                reinterpret_cast<V *>(ph_next)[q] = ho;
            }
        }
    }
    for (long n = n0 + tid; n < N; n += stride) {
        x[n] = ((FIRST ? T(0) : x[n]) + a * ph[n]) + om * sh[n];
        if (mode >= 2) {
            const T rv = T(1) * s[n] + nom * t[n];
            r[n] = rv;
            if (mode >= 3) {
                const T pv = rv + be * (pr[n] - om * v[n]);
                p[n] = pv;
                if (dinv) ph_next[n] = dinv[n] * pv;
            }
        }
    }
}

// ------------------------------------------------------------------ line relaxation
// One vertical line-relaxation step out = y + M^-1 (in - A y), M the
// tridiagonal (vertical) part of A in every column: the Thomas algorithm, the
// elimination on the way up (its modified coefficients stored in cq, dq) and
// the back substitution on the way down. The terrain-following grid couples
// the levels of a column far more strongly than neighbouring columns
// (dz << dx near the ground), so this removes the stiff vertical error
// components that point relaxation leaves behind. out must not alias y.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_line_relax(Op op, Geo g, const typename Op::type *__restrict__ y,
                                                         const typename Op::type *__restrict__ in,
                                                         typename Op::type *__restrict__ cq,
                                                         typename Op::type *__restrict__ dq,
                                                         typename Op::type *__restrict__ out)
{
    using T = typename Op::type;
    const TilePos t = tile_pos(g);
    if (!t.inner) return;
    const long P = g.P;
    const int mz = g.mz;
    Win<T> win;
    load_level(y, 0, t, win.c1, win.w1, win.e1, win.s1, win.n1);
    load_level(y, P, t, win.c2, win.w2, win.e2, win.s2, win.n2);
    win.c0 = win.w0 = win.e0 = win.s0 = win.n0 = T(0);
    const typename Op::Bc bc = op.bc(t.c);
    T cprev = T(0), dprev = T(0);
    for (int k = 0; k < mz; ++k) {
        const long n = (long)k * P + t.c;
        typename Op::Cf cf = {};
        if (k >= 1 && k <= mz - 2) cf = op.load(n);
        const T r = in[n] - op.row(k, mz, t.lat, win, cf, bc);
        T l, dd, u;
        op.tri(k, mz, t.lat, cf, bc, l, dd, u);
        const T m = dd - l * cprev;
        cprev = u / m;
        dprev = (r - l * dprev) / m;
        cq[n] = cprev;
        dq[n] = dprev;
        win.shift();
        if (k + 2 <= mz - 1) load_level(y, (long)(k + 2) * P, t, win.c2, win.w2, win.e2, win.s2, win.n2);
    }
    T delta = T(0);
    for (int k = mz - 1; k >= 0; --k) {
        const long n = (long)k * P + t.c;
        delta = dq[n] - cq[n] * delta;
        out[n] = y[n] + delta;
    }
}

// ---------------------------------------------- line-preconditioned FP32 BiCGStab
// Inner solver of WINDHIP_MIXED_LINE: the reference's BiCGStab iteration with
// the preconditioner M^-1 = two vertical line-relaxation steps
// (y1 = L^-1 in, y2 = y1 + L^-1 (in - A y1), L the tridiagonal column part of
// A) instead of the two Jacobi-Richardson sweeps.

// Thomas factors of L per column, computed in FP64 from the FP64 operator and
// stored in FP32: lo = coefficient of x(k-1), ip = 1 / pivot, cp = upper
// coefficient / pivot.
template <class Op>
__global__ __launch_bounds__(kThreads) void k_line_factor(Op op, Geo g, float *__restrict__ lo,
                                                          float *__restrict__ cp, float *__restrict__ ip)
{
    const TilePos t = tile_pos(g);
    if (!t.inner) return;
    const typename Op::Bc bc = op.bc(t.c);
    double cprev = 0.0;
    for (int k = 0; k < g.mz; ++k) {
        const long n = (long)k * g.P + t.c;
        typename Op::Cf cf = {};
        if (k >= 1 && k <= g.mz - 2) cf = op.load(n);
        double l, dd, u;
        op.tri(k, g.mz, t.lat, cf, bc, l, dd, u);
        const double m = dd - l * cprev;
        cprev = u / m;
        lo[n] = (float)l;
        ip[n] = (float)(1.0 / m);
        cp[n] = (float)cprev;
    }
}

// out = L^-1 in
__global__ __launch_bounds__(kThreads) void k_line_solve(Geo g, const KState *st, const float *__restrict__ in,
                                                         const float *__restrict__ lo, const float *__restrict__ cp,
                                                         const float *__restrict__ ip, float *__restrict__ dq,
                                                         float *__restrict__ out)
{
    if (st->done) return;
    const TilePos t = tile_pos(g);
    if (!t.inner) return;
    float dprev = 0.f;
    for (int k = 0; k < g.mz; ++k) {
        const long n = (long)k * g.P + t.c;
        dprev = (in[n] - lo[n] * dprev) * ip[n];
        dq[n] = dprev;
    }
    float delta = 0.f;
    for (int k = g.mz - 1; k >= 0; --k) {
        const long n = (long)k * g.P + t.c;
        delta = dq[n] - cp[n] * delta;
        out[n] = delta;
    }
}

// out = y + L^-1 (in - A y); out must not alias y
template <class Op>
__global__ __launch_bounds__(kThreads) void k_line_step(Op op, Geo g, const KState *st,
                                                        const float *__restrict__ y, const float *__restrict__ in,
                                                        const float *__restrict__ lo, const float *__restrict__ cp,
                                                        const float *__restrict__ ip, float *__restrict__ dq,
                                                        float *__restrict__ out)
{
    if (st->done) return;
    const TilePos t = tile_pos(g);
    if (!t.inner) return;
    float dprev = 0.f;
    auto epi = [&](int, long n, float ay, float, float) {
        dprev = ((in[n] - ay) - lo[n] * dprev) * ip[n];
        dq[n] = dprev;
    };
    march<Op, false>(op, g, y, t, epi);
    float delta = 0.f;
    for (int k = g.mz - 1; k >= 0; --k) {
        const long n = (long)k * g.P + t.c;
        delta = dq[n] - cp[n] * delta;
        out[n] = y[n] + delta;
    }
}

// v = A p_hat and sigma = <r_hat, v>
template <class Op>
__global__ __launch_bounds__(kThreads) void k_mv_sigma(Op op, Geo g, KState *st, double *partials,
                                                       const typename Op::type *__restrict__ ph,
                                                       const typename Op::type *__restrict__ rh,
                                                       typename Op::type *__restrict__ v)
{
    using T = typename Op::type;
    if (st->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->fmode = 0;
        return;
    }
    const TilePos t = tile_pos(g);
    double acc[1] = {0.0};
    if (t.inner) {
        auto epi = [&](int, long n, T yv, T, T) {
            v[n] = yv;
            acc[0] = acc[0] + (double)rh[n] * (double)yv;
        };
        march<Op, false>(op, g, ph, t, epi);
    }
    reduce_finalize<1>(acc, partials, &st->cnt[1], t.L, g.nt, [&](const double *tot) { finish_sigma(st, tot); });
}

// s = r - alpha v
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_s_only(long N, const KState *st, const T *__restrict__ r,
                                                          const T *__restrict__ v, T *__restrict__ s)
{
    if (st->done) return;
    const T na = T(-st->alpha);
    const long tid = (long)blockIdx.x * kBlock + threadIdx.x, stride = (long)gridDim.x * kBlock;
    long n0 = 0;
    if (aligned16(r, v, s)) {  // 16 bytes per lane and array, the same per-element arithmetic
        using V = typename Vec16<T>::type;
        constexpr int W = Vec16<T>::n;
        n0 = N / W * W;
        for (long q = tid; q < N / W; q += stride) {
            const V rv = reinterpret_cast<const V *>(r)[q], vv = reinterpret_cast<const V *>(v)[q];
            V sv;
#pragma unroll
            for (int e = 0; e < W; ++e) sv[e] = T(1) * rv[e] + na * vv[e];
            reinterpret_cast<V *>(s)[q] = sv;
        }
    }
    for (long n = n0 + tid; n < N; n += stride) s[n] = T(1) * r[n] + na * v[n];
}

// t = A s_hat and the five dots <t,s>, <t,t>, <r_hat,s>, <r_hat,t>, <s,s>
template <class Op>
__global__ __launch_bounds__(kThreads) void k_mv_omega(Op op, Geo g, KState *st, double *partials,
                                                       const typename Op::type *__restrict__ sh,
                                                       const typename Op::type *__restrict__ s,
                                                       const typename Op::type *__restrict__ rh,
                                                       typename Op::type *__restrict__ tt_out)
{
    using T = typename Op::type;
    if (st->done) return;
    const TilePos t = tile_pos(g);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (t.inner) {
        auto epi = [&](int, long n, T tv, T, T) {
            tt_out[n] = tv;
            const double td = tv, sd = s[n], rd = rh[n];
            acc[0] = acc[0] + td * sd;
            acc[1] = acc[1] + td * td;
            acc[2] = acc[2] + rd * sd;
            acc[3] = acc[3] + rd * td;
            acc[4] = acc[4] + sd * sd;
        };
        march<Op, false>(op, g, sh, t, epi);
    }
    reduce_finalize<5>(acc, partials, &st->cnt[2], t.L, g.nt, [&](const double *tot) { finish_omega(st, tot); });
}

// ---------------------------------------------- flat column march (FP32 line path)
// Single-stage FP32 kernels of the line-preconditioned iteration carry no
// bitwise contract, so they take one column per lane over all P columns
// (256 consecutive columns per block, XCD-aware block order) instead of the
// tiles, whose halo lanes would idle. Their reductions use the same
// block-partial / last-block fold with one partial per block.
constexpr int kColThreads = 256;  // flat column kernels: 4 wavefronts per block
constexpr int kColWaves = kColThreads / 64;

struct ColPos {
    int L;
    bool valid, lat;
    long c;
    unsigned oc;
};

__device__ __forceinline__ ColPos col_pos(const Geo &g, int nb)
{
    ColPos t;
    t.L = logical_block(nb);
    const long c = (long)t.L * kColThreads + threadIdx.x;
    t.valid = c < g.P;
    t.c = t.valid ? c : 0;
    const int i = (int)(t.c % g.mx), j = (int)(t.c / g.mx);
    t.lat = i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1;
    t.oc = (unsigned)t.c;
    return t;
}

// Pipelined march of (A x)(row) over levels [k0, k1) of one column, bottom to
// top: the operands of level m (x at m+1, coefficients, the epilogue's
// pre-loads) are loaded one level ahead into alternating register sets, all
// through buffer resources; the ghost planes are peeled where they fall in the
// range. epi.apply(k, so, vo, ax, xc, pre) per level.
// XT: the element type of x when it differs from the operator's (an FP32
// vector under the FP64 operator: widened on load).
template <class Op, class Epi, typename XT = typename Op::type>
__device__ __forceinline__ void march_pipe(const Op &op, const Geo &g, const ColPos &t, const XT *__restrict__ x,
                                           int k0, int k1, Epi &epi)
{
    using T = typename Op::type;
    using Cf = typename Op::Cf;
    using Pre = typename Epi::Pre;
    struct Slot {
        T c, w, e, s, n;
        Cf cf;
        Pre ep;
    };
    const int mx = __builtin_amdgcn_readfirstlane(g.mx);
    const int mz = __builtin_amdgcn_readfirstlane(g.mz);
    const long P = (long)mx * __builtin_amdgcn_readfirstlane(g.my);
    const long N = P * mz;
    constexpr unsigned sz = sizeof(T), xs = sizeof(XT);
    const unsigned pb = (unsigned)(P * (long)sz);
    const unsigned voc = t.oc * sz;
    // offsets of x (element size xs)
    const unsigned xpb = (unsigned)(P * (long)xs), xvoc = t.oc * xs;
    const long gd = mx + 1;
    const unsigned oS = xs, oW = (unsigned)mx * xs, oC = (unsigned)(mx + 1) * xs, oE = (unsigned)(mx + 2) * xs,
                   oN = (unsigned)(2 * mx + 1) * xs;
    const rsrc_t rx = make_rsrc(x - gd, (N + 2 * gd) * (long)xs);
    const typename Op::Src rcf = op.src(N);
    epi.setup(N);
    auto lx5 = [&](unsigned so, T &c, T &w, T &e, T &s, T &n) {
        if constexpr (sizeof(XT) == sizeof(T)) {
            load_x5(rx, xvoc, so + oS, so + oW, so + oC, so + oE, so + oN, c, w, e, s, n);
        } else {
            XT c_, w_, e_, s_, n_;
            load_x5(rx, xvoc, so + oS, so + oW, so + oC, so + oE, so + oN, c_, w_, e_, s_, n_);
            c = (T)c_; w = (T)w_; e = (T)e_; s = (T)s_; n = (T)n_;
        }
    };
    auto fill = [&](Slot &sl, int m, auto pos) {
        constexpr int ps = decltype(pos)::value;
        if (ps != kTop) {
            const unsigned so = (unsigned)(m + 1 < mz - 1 ? m + 1 : mz - 1) * xpb;
            lx5(so, sl.c, sl.w, sl.e, sl.s, sl.n);
        }
        if (ps == kMid) {
            const int lv = m < mz - 2 ? m : mz - 2;
            sl.cf = op.load_src(rcf, (long)lv * P, (unsigned)lv * pb, t.oc, voc);
        }
        sl.ep = epi.load((unsigned)(m < mz - 1 ? m : mz - 1) * pb, voc);
    };
    Win<T> win = {};
    const typename Op::Bc bc = op.bc(t.c);
    auto step = [&](const Slot &cur, int m, auto pos) {
        constexpr int ps = decltype(pos)::value;
        if (ps != kFirst) win.shift();
        win.c2 = keep(cur.c); win.w2 = keep(cur.w); win.e2 = keep(cur.e);
        win.s2 = keep(cur.s); win.n2 = keep(cur.n);
        constexpr int rc = ps == kFirst ? kRowBottom : (ps == kTop ? kRowTop : kRowInterior);
        const T ax = op.template row_c<rc>(t.lat, win, cur.cf, bc);
        epi.apply(m, (unsigned)m * pb, voc, ax, win.c1, cur.ep);
    };
    using PF = std::integral_constant<int, kFirst>;
    using PM = std::integral_constant<int, kMid>;
    using PT = std::integral_constant<int, kTop>;
    Slot sa, sb;
    int m = k0;
    if (k0 == 0) {
        lx5(0u, win.c1, win.w1, win.e1, win.s1, win.n1);
        fill(sa, 0, PF{});
        fill(sb, 1, PM{});
        step(sa, 0, PF{});
        m = 1;
    } else {
        // window before the first step: level k0-1 as c1, level k0 as c2
        const unsigned s1 = (unsigned)(k0 - 1) * xpb, s2 = (unsigned)k0 * xpb;
        lx5(s1, win.c1, win.w1, win.e1, win.s1, win.n1);
        lx5(s2, win.c2, win.w2, win.e2, win.s2, win.n2);
        fill(sb, k0, PM{});
    }
    // sb holds the operands of step m; interior steps m .. mend-1
    const int mend = k1 == mz ? mz - 1 : k1;
    for (; m + 1 < mend; m += 2) {
        fill(sa, m + 1, PM{});
        step(sb, m, PM{});
        fill(sb, m + 2, PM{});
        step(sa, m + 1, PM{});
    }
    if (m < mend) {  // one interior step left
        if (k1 == mz) {
            fill(sa, mz - 1, PT{});
            step(sb, m, PM{});
            step(sa, mz - 1, PT{});
        } else {
            step(sb, m, PM{});
        }
    } else if (k1 == mz) {
        step(sb, mz - 1, PT{});
    }
}

constexpr int kChunksK = 3;  // level chunks per column of the flat stencil kernels

// 1-D grid of col_blocks(g) * kChunksK blocks; consecutive logical blocks are
// the level chunks of the same columns (same XCD, shared halo levels in L2).
struct ChunkPos {
    ColPos t;
    int k0, k1, L, nb;
};

__device__ __forceinline__ ChunkPos chunk_pos(const Geo &g)
{
    ChunkPos c;
    c.nb = gridDim.x;
    c.L = logical_block(c.nb);
    const int cb = c.L / kChunksK, q = c.L - cb * kChunksK;
    const long col = (long)cb * kColThreads + threadIdx.x;
    c.t.L = c.L;
    c.t.valid = col < g.P;
    c.t.c = c.t.valid ? col : 0;
    const int i = (int)(c.t.c % g.mx), j = (int)(c.t.c / g.mx);
    c.t.lat = i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1;
    c.t.oc = (unsigned)c.t.c;
    c.k0 = (int)((long)q * g.mz / kChunksK);
    c.k1 = (int)((long)(q + 1) * g.mz / kChunksK);
    return c;
}

inline int col_blocks(const Geo &g) { return (int)((g.P + kColThreads - 1) / kColThreads); }
inline int chunk_blocks(const Geo &g) { return col_blocks(g) * kChunksK; }

// v = A p_hat and sigma = <r_hat, v> (FP32 line path)
template <class Op>
__global__ __launch_bounds__(kColThreads) void k_mv_sigma_c(Op op, Geo g, KState *st, double *partials,
                                                            const typename Op::type *__restrict__ ph,
                                                            const typename Op::type *__restrict__ rh,
                                                            typename Op::type *__restrict__ v)
{
    using T = typename Op::type;
    if (st->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->fmode = 0;
        return;
    }
    const ChunkPos c = chunk_pos(g);
    struct Epi {
        const T *rh;
        T *v;
        double acc[1];
        bool on;
        rsrc_t rr, rv;
        struct Pre { T rh; };
        __device__ void setup(long N) { rr = make_rsrc(rh, N * (long)sizeof(T)); rv = make_rsrc(v, N * (long)sizeof(T)); }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<T>(rr, vo, so)}; }
        __device__ void apply(int, unsigned so, unsigned vo, T ax, T, const Pre &p)
        {
            if (!on) return;
            bstore(ax, rv, vo, so);
            acc[0] = acc[0] + (double)p.rh * (double)ax;
        }
    } epi{rh, v, {0.0}, c.t.valid};
    march_pipe(op, g, c.t, ph, c.k0, c.k1, epi);
    reduce_finalize<1, kColWaves>(epi.acc, partials, &st->cnt[1], c.L, c.nb,
                                  [&](const double *tot) { finish_sigma(st, tot); });
}

// t = A s_hat and the five dots (FP32 line path)
template <class Op>
__global__ __launch_bounds__(kColThreads) void k_mv_omega_c(Op op, Geo g, KState *st, double *partials,
                                                            const typename Op::type *__restrict__ sh,
                                                            const typename Op::type *__restrict__ s,
                                                            const typename Op::type *__restrict__ rh,
                                                            typename Op::type *__restrict__ tt_out)
{
    using T = typename Op::type;
    if (st->done) return;
    const ChunkPos c = chunk_pos(g);
    struct Epi {
        const T *s, *rh;
        T *out;
        double acc[5];
        bool on;
        rsrc_t rs, rr, ro;
        struct Pre { T s, rh; };
        __device__ void setup(long N)
        {
            const long nbt = N * (long)sizeof(T);
            rs = make_rsrc(s, nbt); rr = make_rsrc(rh, nbt); ro = make_rsrc(out, nbt);
        }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<T>(rs, vo, so), bload<T>(rr, vo, so)}; }
        __device__ void apply(int, unsigned so, unsigned vo, T tv, T, const Pre &p)
        {
            if (!on) return;
            bstore(tv, ro, vo, so);
            const double td = tv, sd = p.s, rd = p.rh;
            acc[0] = acc[0] + td * sd;
            acc[1] = acc[1] + td * td;
            acc[2] = acc[2] + rd * sd;
            acc[3] = acc[3] + rd * td;
            acc[4] = acc[4] + sd * sd;
        }
    } epi{s, rh, tt_out, {0.0, 0.0, 0.0, 0.0, 0.0}, c.t.valid};
    march_pipe(op, g, c.t, sh, c.k0, c.k1, epi);
    reduce_finalize<5, kColWaves>(epi.acc, partials, &st->cnt[2], c.L, c.nb,
                                  [&](const double *tot) { finish_omega(st, tot); });
}

// r = in - A y (FP32 line path)
template <class Op>
__global__ __launch_bounds__(kColThreads) void k_resid_c(Op op, Geo g, const KState *st,
                                                         const typename Op::type *__restrict__ y,
                                                         const typename Op::type *__restrict__ in,
                                                         typename Op::type *__restrict__ r)
{
    using T = typename Op::type;
    if (st->done) return;
    const ChunkPos c = chunk_pos(g);
    struct Epi {
        const T *in;
        T *r;
        bool on;
        rsrc_t ri, rr;
        struct Pre { T in; };
        __device__ void setup(long N) { ri = make_rsrc(in, N * (long)sizeof(T)); rr = make_rsrc(r, N * (long)sizeof(T)); }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<T>(ri, vo, so)}; }
        __device__ void apply(int, unsigned so, unsigned vo, T ay, T, const Pre &p)
        {
            if (on) bstore(p.in - ay, rr, vo, so);
        }
    } epi{in, r, c.t.valid};
    march_pipe(op, g, c.t, y, c.k0, c.k1, epi);
}

// ---------------------------------------------------------------- FP64 outer residuals
// The FP64 stencil passes of the outer solves (FGMRES-IR and the true
// residual) as level-chunked column marches (kChunksK chunks per column,
// 1-D grid of chunk_blocks(g) blocks of kColThreads). Each row is the FP64
// row of Op in the reference's term order; only the order of the reductions'
// partial sums differs from the tiled FP64 BiCGStab kernels.

// ||b - A x||^2 in FP64 (true residual): st->outer_rnorm, st->outer_bnorm.
template <class Op>
__global__ __launch_bounds__(kColThreads) void k_true_resid(Op op, Geo g, KState *st, double *partials,
                                                            const double *__restrict__ x,
                                                            const double *__restrict__ b)
{
    const ChunkPos c = chunk_pos(g);
    struct Epi {
        const double *b;
        double acc[2];
        bool on;
        rsrc_t rb;
        struct Pre { double b; };
        __device__ void setup(long N) { rb = make_rsrc(b, N * 8); }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<double>(rb, vo, so)}; }
        __device__ void apply(int, unsigned, unsigned, double ax, double, const Pre &p)
        {
            if (!on) return;
            const double rv = p.b - ax;
            acc[0] = acc[0] + rv * rv;
            acc[1] = acc[1] + p.b * p.b;
        }
    } epi{b, {0.0, 0.0}, c.t.valid};
    march_pipe(op, g, c.t, x, c.k0, c.k1, epi);
    reduce_finalize<2, kColWaves>(epi.acc, partials, &st->cnt[3], c.L, c.nb, [&](const double *tot) {
        st->outer_rnorm = sqrt(tot[0]);
        st->outer_bnorm = sqrt(tot[1]);
    });
}

// r = b - A x (FP64) into v0, ||r||^2 and ||b||^2 -> st (outer_rnorm, outer_bnorm,
// outer_target, outer_done); h[0] = ||r||^2 for the first basis vector's scale.
template <class Op>
__global__ __launch_bounds__(kColThreads) void k_fg_resid(Op op, Geo g, KState *st, double *partials,
                                                          const double *__restrict__ x,
                                                          const double *__restrict__ b, double *__restrict__ v0,
                                                          double *h)
{
    const ChunkPos c = chunk_pos(g);
    struct Epi {
        const double *b;
        double *v0;
        double acc[2];
        bool on;
        rsrc_t rb, rv;
        struct Pre { double b; };
        __device__ void setup(long N) { rb = make_rsrc(b, N * 8); rv = make_rsrc(v0, N * 8); }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<double>(rb, vo, so)}; }
        __device__ void apply(int, unsigned so, unsigned vo, double ax, double, const Pre &p)
        {
            if (!on) return;
            const double rv_ = p.b - ax;
            bstore(rv_, rv, vo, so);
            acc[0] = acc[0] + rv_ * rv_;
            acc[1] = acc[1] + p.b * p.b;
        }
    } epi{b, v0, {0.0, 0.0}, c.t.valid};
    march_pipe(op, g, c.t, x, c.k0, c.k1, epi);
    reduce_finalize<2, kColWaves>(epi.acc, partials, &st->cnt[3], c.L, c.nb, [&](const double *tot) {
        st->outer_rnorm = sqrt(tot[0]);
        st->outer_bnorm = sqrt(tot[1]);
        st->outer_target = fmax(st->tol_abs, st->tol_rel * st->outer_bnorm);
        st->outer_done = st->outer_rnorm <= st->outer_target;
        h[0] = tot[0];
    });
}

// w = A z (FP64, z the FP32 inner solution widened on load) and h[0] =
// <w, v0>; z is kept in zk (FP32) for the final update.
template <class Op>
__global__ __launch_bounds__(kColThreads) void k_fg_matvec(Op op, Geo g, KState *st, double *partials,
                                                           const float *__restrict__ z, float *__restrict__ zk,
                                                           const double *__restrict__ v0, double *__restrict__ w,
                                                           double *h)
{
    const ChunkPos c = chunk_pos(g);
    struct Epi {
        const double *v0;
        double *w;
        float *zk;
        double acc[1];
        bool on;
        rsrc_t rv, rw, rz;
        struct Pre { double v; };
        __device__ void setup(long N) { rv = make_rsrc(v0, N * 8); rw = make_rsrc(w, N * 8); rz = make_rsrc(zk, N * 4); }
        __device__ Pre load(unsigned so, unsigned vo) const { return Pre{bload<double>(rv, vo, so)}; }
        __device__ void apply(int, unsigned so, unsigned vo, double ax, double xc, const Pre &p)
        {
            if (!on) return;
            bstore(ax, rw, vo, so);
            bstore((float)xc, rz, vo >> 1, so >> 1);  // exact: xc is the FP32 value widened
            acc[0] = acc[0] + ax * p.v;
        }
    } epi{v0, w, zk, {0.0}, c.t.valid};
    march_pipe(op, g, c.t, z, c.k0, c.k1, epi);
    reduce_finalize<1, kColWaves>(epi.acc, partials, &st->cnt[3], c.L, c.nb, [&](const double *tot) { h[0] = tot[0]; });
}

// out = y + w L^-1 in (y null: out = w L^-1 in), the Thomas algorithm per column
// with the loads of kLB levels issued together on each pass (the recursion
// itself is a few FMAs per level; its loads are what costs).
constexpr int kLB = 8;
__global__ __launch_bounds__(kColThreads) void k_line_solve_c(Geo g, const KState *st, const float *__restrict__ in,
                                                              const float *__restrict__ lo,
                                                              const float *__restrict__ cp,
                                                              const float *__restrict__ ip, float *__restrict__ dq,
                                                              const float *__restrict__ y, float *__restrict__ out,
                                                              float w)
{
    if (st->done) return;
    const ColPos t = col_pos(g, gridDim.x);
    const int mz = g.mz;
    const long N = g.P * mz;
    const unsigned voc = t.oc * 4u, pb = (unsigned)(g.P * 4);
    const long nbt = N * 4;
    const rsrc_t rin = make_rsrc(in, nbt), rlo = make_rsrc(lo, nbt), rip = make_rsrc(ip, nbt),
                 rdq = make_rsrc(dq, nbt), rcp = make_rsrc(cp, nbt), ro = make_rsrc(out, nbt);
    const bool hy = y != nullptr;
    const rsrc_t ry = make_rsrc(hy ? y : in, nbt);
    float dprev = 0.f;
    for (int k0 = 0; k0 < mz; k0 += kLB) {
        float a[kLB], l[kLB], q[kLB];
#pragma unroll
        for (int e = 0; e < kLB; ++e) {
            const unsigned so = (unsigned)(k0 + e < mz ? k0 + e : mz - 1) * pb;
            a[e] = bload<float>(rin, voc, so);
            l[e] = bload<float>(rlo, voc, so);
            q[e] = bload<float>(rip, voc, so);
        }
#pragma unroll
        for (int e = 0; e < kLB; ++e) {
            if (k0 + e < mz) {
                dprev = (a[e] - l[e] * dprev) * q[e];
                if (t.valid) bstore(dprev, rdq, voc, (unsigned)(k0 + e) * pb);
            }
        }
    }
    float delta = 0.f;
    for (int k1 = mz - 1; k1 >= 0; k1 -= kLB) {
        float d[kLB], c[kLB], yv[kLB];
#pragma unroll
        for (int e = 0; e < kLB; ++e) {
            const unsigned so = (unsigned)(k1 - e >= 0 ? k1 - e : 0) * pb;
            d[e] = bload<float>(rdq, voc, so);
            c[e] = bload<float>(rcp, voc, so);
            yv[e] = hy ? bload<float>(ry, voc, so) : 0.f;
        }
#pragma unroll
        for (int e = 0; e < kLB; ++e) {
            if (k1 - e >= 0) {
                delta = d[e] - c[e] * delta;
                if (t.valid) bstore(yv[e] + w * delta, ro, voc, (unsigned)(k1 - e) * pb);
            }
        }
    }
}

// The same solve for columns of at most KM levels with the forward results kept
// in registers: every level's loads are issued up front in batches, the
// forward recursion fills dq[], the back substitution reads it back, so the
// intermediate never goes to memory (5 arrays of traffic, 6 with y, instead
// of 7 and 8).
// factor rows in FP32 or bf16 (FT = unsigned short: the bits of a bf16, read
// with halved byte offsets and widened to FP32)
template <typename FT>
__device__ __forceinline__ float fload(rsrc_t r, unsigned vo, unsigned so)
{
    if constexpr (sizeof(FT) == 2) return OpAB::ld(r, vo, so);
    else return bload<float>(r, vo, so);
}
// The factor rows of k_line_solve_r: FT = float, three FP32 rows; FT =
// unsigned short, three bf16 rows; FT = unsigned, one {lo | ip << 16} bf16 pair
// per row (lo holds it) beside a bf16 cp row (cp), so the forward pass loads
// both of its factors with one 32-bit load.
template <typename FT>
__device__ __forceinline__ void fload_lq(rsrc_t rlo, rsrc_t rip, unsigned vo, unsigned so, float &l, float &q)
{
    if constexpr (std::is_same<FT, unsigned>::value) {
        const unsigned w = __builtin_amdgcn_raw_buffer_load_b32(rlo, (int)vo, uni(so), 0);
        l = OpAB::lo(w);
        q = OpAB::hi(w);
    } else {
        l = fload<FT>(rlo, vo, so);
        q = fload<FT>(rip, vo, so);
    }
}
template <typename FT>
__device__ __forceinline__ float fload_c(rsrc_t rcp, unsigned vo, unsigned so)
{
    if constexpr (std::is_same<FT, unsigned>::value) return OpAB::ld(rcp, vo, so);
    else return fload<FT>(rcp, vo, so);
}

template <int KM, typename FT = float, int LB = kLB>
__global__ __launch_bounds__(kColThreads) void k_line_solve_r(Geo g, const KState *st, const float *__restrict__ in,
                                                              const FT *__restrict__ lo, const FT *__restrict__ cp,
                                                              const FT *__restrict__ ip,
                                                              const float *__restrict__ y, float *__restrict__ out,
                                                              float w)
{
    static_assert(KM % LB == 0, "the level batches must tile KM");
    if (st->done) return;
    const ColPos t = col_pos(g, gridDim.x);
    const int mz = __builtin_amdgcn_readfirstlane(g.mz);
    const long N = g.P * mz;
    const unsigned voc = t.oc * 4u, pb = (unsigned)(g.P * 4);
    constexpr bool kPair = std::is_same<FT, unsigned>::value;
    const long nbt = N * 4, nbf = N * (long)sizeof(FT);
    const rsrc_t rin = make_rsrc(in, nbt), rlo = make_rsrc(lo, nbf), rip = kPair ? rlo : make_rsrc(ip, nbf),
                 rcp = make_rsrc(cp, kPair ? N * 2 : nbf), ro = make_rsrc(out, nbt);
    const bool hy = y != nullptr;
    const rsrc_t ry = make_rsrc(hy ? y : in, nbt);
    float dq[KM];
    float dprev = 0.f, delta = 0.f;
    if constexpr (LB < kLB) {
        // short batches, each loaded one batch ahead of its recursion; the
        // schedule barriers and the opaque use of the running value keep every
        // batch in its place, so the loads of later levels do not pile up in
        // registers (unpinned, the compiler issues all of a pass's loads first)
        float a[2][LB], l[2][LB], q[2][LB], c[2][LB], yv[2][LB];
        auto ldf = [&](int k0, int sl) {
#pragma unroll
            for (int e = 0; e < LB; ++e) {
                const unsigned so = (unsigned)(k0 + e < mz ? k0 + e : mz - 1) * pb;
                a[sl][e] = bload<float>(rin, voc, so);
                fload_lq<FT>(rlo, rip, voc, so, l[sl][e], q[sl][e]);
            }
        };
        auto ldb = [&](int k1, int sl) {
#pragma unroll
            for (int e = 0; e < LB; ++e) {
                const unsigned so = (unsigned)(k1 - e < mz ? k1 - e : mz - 1) * pb;
                c[sl][e] = fload_c<FT>(rcp, voc, so);
                yv[sl][e] = hy ? bload<float>(ry, voc, so) : 0.f;
            }
        };
        ldf(0, 0);
#pragma unroll
        for (int k0 = 0; k0 < KM; k0 += LB) {
            const int sl = (k0 / LB) & 1;
            __builtin_amdgcn_sched_barrier(0);
            if (k0 + LB < KM) ldf(k0 + LB, sl ^ 1);
            else ldb(KM - 1, 0);
#pragma unroll
            for (int e = 0; e < LB; ++e) {
                dprev = (a[sl][e] - l[sl][e] * dprev) * q[sl][e];
                dq[k0 + e] = dprev;  // levels >= mz hold values that are never used
            }
            asm volatile("" : "+v"(dprev));
        }
#pragma unroll
        for (int k1 = KM - 1; k1 >= 0; k1 -= LB) {
            const int sl = ((KM - 1 - k1) / LB) & 1;
            __builtin_amdgcn_sched_barrier(0);
            if (k1 - LB >= 0) ldb(k1 - LB, sl ^ 1);
#pragma unroll
            for (int e = 0; e < LB; ++e) {
                const int k = k1 - e;
                delta = k < mz ? dq[k] - c[sl][e] * delta : 0.f;
                if (t.valid && k < mz) bstore(yv[sl][e] + w * delta, ro, voc, (unsigned)k * pb);
            }
            asm volatile("" : "+v"(delta));
        }
        return;
    }
#pragma unroll
    for (int k0 = 0; k0 < KM; k0 += LB) {
        float a[LB], l[LB], q[LB];
#pragma unroll
        for (int e = 0; e < LB; ++e) {
            const unsigned so = (unsigned)(k0 + e < mz ? k0 + e : mz - 1) * pb;
            a[e] = bload<float>(rin, voc, so);
            fload_lq<FT>(rlo, rip, voc, so, l[e], q[e]);
        }
#pragma unroll
        for (int e = 0; e < LB; ++e) {
            dprev = (a[e] - l[e] * dprev) * q[e];
            dq[k0 + e] = dprev;  // levels >= mz hold values that are never used
        }
    }
#pragma unroll
    for (int k1 = KM - 1; k1 >= 0; k1 -= LB) {
        float c[LB], yv[LB];
#pragma unroll
        for (int e = 0; e < LB; ++e) {
            const unsigned so = (unsigned)(k1 - e < mz ? k1 - e : mz - 1) * pb;
            c[e] = fload_c<FT>(rcp, voc, so);
            yv[e] = hy ? bload<float>(ry, voc, so) : 0.f;
        }
#pragma unroll
        for (int e = 0; e < LB; ++e) {
            const int k = k1 - e;
            delta = k < mz ? dq[k] - c[e] * delta : 0.f;
            if (t.valid && k < mz) bstore(yv[e] + w * delta, ro, voc, (unsigned)k * pb);
        }
    }
}

// One damped line-relaxation step in a single pass, out = y + w L^-1 (in - A y):
// each level's residual (the row of Op on the five-column window of y, as in
// k_resid_c) feeds the forward recursion of the column's Thomas algorithm at
// once, the forward results stay in registers for the back substitution (as
// in k_line_solve_r), which reloads the centre column of y. Element by
// element the arithmetic of k_resid_c followed by k_line_solve_r, without the
// residual's round trip through memory and in one launch. Columns of at most
// KM levels, one column per thread, NT threads per block; the operands of a
// level are loaded PD levels ahead (y must carry the stencil guards).
constexpr int kSmoothSmallNT = 64;  // block of k_smooth_r on the coarse levels (fewer columns: more blocks)
template <class Op, int KM, typename FT, int NT = kColThreads, int PD = 1, int LB = 4>
__global__ __launch_bounds__(NT) void k_smooth_r(Op op, Geo g, const KState *st, const float *__restrict__ in,
                                                 const float *__restrict__ y, const FT *__restrict__ lo,
                                                 const FT *__restrict__ cp, const FT *__restrict__ ip,
                                                 float *__restrict__ out, float w)
{
    static_assert(KM % LB == 0, "the level batches must tile KM");
    using Cf = typename Op::Cf;
    if (st->done) return;
    const int L = logical_block(gridDim.x);
    const long col = (long)L * NT + threadIdx.x;
    const bool valid = col < g.P;
    const long c = valid ? col : 0;
    const int mx = __builtin_amdgcn_readfirstlane(g.mx);
    const int my = __builtin_amdgcn_readfirstlane(g.my);
    const int mz = __builtin_amdgcn_readfirstlane(g.mz);
    const long P = (long)mx * my;
    const long N = P * mz;
    const int i = (int)(c % mx), j = (int)(c / mx);
    const bool lat = i == 0 || i == mx - 1 || j == 0 || j == my - 1;
    const unsigned oc = (unsigned)c, voc = oc * 4u, pb = (unsigned)(P * 4);
    const long gd = mx + 1;
    const unsigned oS = 4u, oW = (unsigned)mx * 4u, oC = (unsigned)(mx + 1) * 4u, oE = (unsigned)(mx + 2) * 4u,
                   oN = (unsigned)(2 * mx + 1) * 4u;
    constexpr bool kPair = std::is_same<FT, unsigned>::value;
    const long nbf = N * (long)sizeof(FT);
    const rsrc_t ry = make_rsrc(y - gd, (N + 2 * gd) * 4), rin = make_rsrc(in, N * 4), rlo = make_rsrc(lo, nbf),
                 rip = kPair ? rlo : make_rsrc(ip, nbf), rcp = make_rsrc(cp, kPair ? N * 2 : nbf),
                 ro = make_rsrc(out, N * 4);
    const typename Op::Src rcf = op.src(N);
    const typename Op::Bc bc = op.bc(c);
    struct Slot {
        float c, w, e, s, n, b, l, q;
        Cf cf;
    };
    // the operands of step k: y at level k + 1, in, the row's coefficients and
    // the forward factors at level k (levels past the top re-read the top one)
    auto fill = [&](Slot &sl, int k) {
        const unsigned s1 = (unsigned)(k + 1 < mz ? k + 1 : mz - 1) * pb;
        load_x5(ry, voc, s1 + oS, s1 + oW, s1 + oC, s1 + oE, s1 + oN, sl.c, sl.w, sl.e, sl.s, sl.n);
        const int kk = k < mz ? k : mz - 1;
        const unsigned so = (unsigned)kk * pb;
        sl.b = bload<float>(rin, voc, so);
        sl.cf = op.load_src(rcf, (long)kk * P, so, oc, voc);
        fload_lq<FT>(rlo, rip, voc, so, sl.l, sl.q);
    };
    Win<float> win = {};
    load_x5(ry, voc, oS, oW, oC, oE, oN, win.c1, win.w1, win.e1, win.s1, win.n1);
    Slot sl[PD + 1];
#pragma unroll
    for (int k = 0; k < PD; ++k) fill(sl[k], k);
    float dq[KM];
    float dprev = 0.f;
    float cb[2][LB], yb[2][LB];
    auto ldb = [&](int k1, int s) {
#pragma unroll
        for (int e = 0; e < LB; ++e) {
            const unsigned so = (unsigned)(k1 - e < mz ? k1 - e : mz - 1) * pb;
            cb[s][e] = fload_c<FT>(rcp, voc, so);
            yb[s][e] = bload<float>(ry, voc, so + oC);
        }
    };
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        __builtin_amdgcn_sched_barrier(0);
        if (k + PD < KM) fill(sl[(k + PD) % (PD + 1)], k + PD);
        else if (k + PD == KM) ldb(KM - 1, 0);
        const Slot &cur = sl[k % (PD + 1)];
        if (k > 0) win.shift();
        win.c2 = cur.c; win.w2 = cur.w; win.e2 = cur.e; win.s2 = cur.s; win.n2 = cur.n;
        float ax;
        if (k == 0) ax = op.template row_c<kRowBottom>(lat, win, cur.cf, bc);
        else if (k == mz - 1) ax = op.template row_c<kRowTop>(lat, win, cur.cf, bc);
        else ax = op.template row_c<kRowInterior>(lat, win, cur.cf, bc);
        const float r = cur.b - ax;
        dprev = (r - cur.l * dprev) * cur.q;
        dq[k] = dprev;  // levels >= mz hold values that are never used
        asm volatile("" : "+v"(dprev));
    }
    float delta = 0.f;
#pragma unroll
    for (int k1 = KM - 1; k1 >= 0; k1 -= LB) {
        const int s = ((KM - 1 - k1) / LB) & 1;
        __builtin_amdgcn_sched_barrier(0);
        if (k1 - LB >= 0) ldb(k1 - LB, s ^ 1);
#pragma unroll
        for (int e = 0; e < LB; ++e) {
            const int k = k1 - e;
            delta = k < mz ? dq[k] - cb[s][e] * delta : 0.f;
            if (valid && k < mz) bstore(yb[s][e] + w * delta, ro, voc, (unsigned)k * pb);
        }
        asm volatile("" : "+v"(delta));
    }
}

// ------------------------------------------------------------------ FGMRES-IR
// Outer solve of the mixed-precision path: flexible GMRES in FP64 whose
// preconditioner is an FP32 BiCGStab solve (GMRES-based iterative refinement).
// V holds the FP64 Arnoldi basis, Z the FP32 preconditioned directions; the
// small Hessenberg least-squares problem is solved on the host.

constexpr int kFlatBlocks = 1024;  // grid of the flat (row-parallel) reductions

// The first FGMRES residual from a zero start: v0 = b (exactly b - A 0), its
// norm and ||b|| as k_fg_resid states them (v0 null: the norms only).
__global__ __launch_bounds__(kThreads) void k_fg_resid_zero(long N, KState *st, double *partials,
                                                            const double *__restrict__ b, double *__restrict__ v0,
                                                            double *h)
{
    double acc[2] = {0.0, 0.0};
    for (long n = (long)blockIdx.x * kThreads + threadIdx.x; n < N; n += (long)gridDim.x * kThreads) {
        const double bv = b[n];
        if (v0) v0[n] = bv;  // null: the norms only
        acc[0] = acc[0] + bv * bv;
    }
    acc[1] = acc[0];
    reduce_finalize<2>(acc, partials, &st->cnt[3], blockIdx.x, gridDim.x, [&](const double *tot) {
        st->outer_rnorm = sqrt(tot[0]);
        st->outer_bnorm = sqrt(tot[1]);
        st->outer_target = fmax(st->tol_abs, st->tol_rel * st->outer_bnorm);
        st->outer_done = st->outer_rnorm <= st->outer_target;
        h[0] = tot[0];
    });
}

// Start of one preconditioner application: scale the basis vector v by
// 1/sqrt(*vsq) in place (FP64) when normalise is set, and set up the FP32
// BiCGStab for A z = v from z = 0 (bicgstab_solve set-up,
// wind_iterative.F90:586-640, with r0 = v). x32 null: only r_hat is written;
// the first iteration of the line and multigrid inner solves reads r and p
// from it and does not read x (k_update_xrp FIRST).
__global__ __launch_bounds__(kThreads) void k_fg_setup(long N, KState *st_in, double *partials,
                                                       const double *vsq, int normalise, double *__restrict__ v,
                                                       const float *__restrict__ dinv32, float *__restrict__ x32,
                                                       float *__restrict__ r32, float *__restrict__ rh32,
                                                       float *__restrict__ p32, float *__restrict__ ph32)
{
    const double sc = normalise ? 1.0 / sqrt(*vsq) : 1.0;
    double acc[1] = {0.0};
    for (long n = (long)blockIdx.x * kThreads + threadIdx.x; n < N; n += (long)gridDim.x * kThreads) {
        const double vv = v[n] * sc;
        if (normalise) v[n] = vv;
        const float rf = (float)vv;
        rh32[n] = rf;
        if (x32) {  // null: the line and multigrid inner solves start from r_hat alone
            x32[n] = 0.0f;
            r32[n] = rf; p32[n] = rf;
        }
        if (dinv32) ph32[n] = dinv32[n] * rf;  // the first Jacobi sweep of p (no dinv32: unused)
        acc[0] = acc[0] + (double)rf * (double)rf;
    }
    reduce_finalize<1>(acc, partials, &st_in->cnt[3], blockIdx.x, gridDim.x, [&](const double *tot) {
        st_in->rho = tot[0];
        st_in->rnorm = sqrt(tot[0]);
        st_in->res0 = st_in->rnorm;
        st_in->bnorm = st_in->rnorm;
        st_in->target = fmax(1.0e-300, st_in->inner_tol * st_in->rnorm);
        st_in->iter = 0;
        st_in->omega = 1.0;
        st_in->alpha = 1.0;
        st_in->fmode = 0;
        if (st_in->rnorm <= st_in->target) { st_in->status = 0; st_in->done = 1; }
        else { st_in->status = 1; st_in->done = 0; }
    });
}

// out = a x + b y (out may alias x or y)
__global__ __launch_bounds__(kBlock) void k_axpby(long N, double a, const double *x, double b, const double *y,
                                                  double *out)
{
    for (long n = (long)blockIdx.x * kBlock + threadIdx.x; n < N; n += (long)gridDim.x * kBlock)
        out[n] = a * x[n] + b * y[n];
}

// One modified Gram-Schmidt step fused with the next inner product:
// w -= h[i] v_i, then h[i+1] = <w, u> (u = v_{i+1}; u = nullptr gives ||w||^2).
__global__ __launch_bounds__(kThreads) void k_fg_axpy_dot(long N, KState *st, double *partials,
                                                          double *__restrict__ w, const double *__restrict__ vi,
                                                          const double *__restrict__ u, double *h, int i)
{
    const double c = h[i];
    double acc[1] = {0.0};
    for (long n = (long)blockIdx.x * kThreads + threadIdx.x; n < N; n += (long)gridDim.x * kThreads) {
        const double wv = w[n] - c * vi[n];
        w[n] = wv;
        const double uv = u ? u[n] : wv;  // u == nullptr: ||w||^2
        acc[0] = acc[0] + wv * uv;
    }
    reduce_finalize<1>(acc, partials, &st->cnt[3], blockIdx.x, gridDim.x,
                       [&](const double *tot) { h[i + 1] = tot[0]; });
}

// The Arnoldi basis vectors one classical Gram-Schmidt pass reads.
constexpr int kFgMaxBasis = 33;
struct FgBasis {
    const double *v[kFgMaxBasis];
};

// Classical Gram-Schmidt, first pass: h[i0 + q] = <w, v_{i0+q}> for the C
// basis vectors q < C, all from one read of w (C a compile-time count, so
// every load is unconditional and a row's loads are in flight together).
template <int C>
__global__ __launch_bounds__(kThreads) void k_fg_dots(long N, KState *st, double *partials,
                                                      const double *__restrict__ w, FgBasis vb, int i0, double *h)
{
    const double *v[C];
#pragma unroll
    for (int q = 0; q < C; ++q) v[q] = vb.v[i0 + q];
    double acc[C];
#pragma unroll
    for (int q = 0; q < C; ++q) acc[q] = 0.0;
    for (long n = (long)blockIdx.x * kThreads + threadIdx.x; n < N; n += (long)gridDim.x * kThreads) {
        const double wv = w[n];
#pragma unroll
        for (int q = 0; q < C; ++q) acc[q] = acc[q] + wv * v[q][n];
    }
    reduce_finalize<C>(acc, partials, &st->cnt[3], blockIdx.x, gridDim.x, [&](const double *tot) {
#pragma unroll
        for (int q = 0; q < C; ++q) h[i0 + q] = tot[q];
    });
}

// Classical Gram-Schmidt, second pass, over the C basis vectors from i0:
// w -= sum_q h[i0 + q] v_{i0+q} (in that order); the pass that ends at v_j
// (last) also forms h[j+1] = ||w||^2.
template <int C>
__global__ __launch_bounds__(kThreads) void k_fg_orth(long N, KState *st, double *partials, double *__restrict__ w,
                                                      FgBasis vb, int i0, int j, int last, double *h)
{
    const double *v[C];
    double c[C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
        v[q] = vb.v[i0 + q];
        c[q] = h[i0 + q];
    }
    double acc[1] = {0.0};
    for (long n = (long)blockIdx.x * kThreads + threadIdx.x; n < N; n += (long)gridDim.x * kThreads) {
        double wv = w[n];
#pragma unroll
        for (int q = 0; q < C; ++q) wv = wv - c[q] * v[q][n];
        w[n] = wv;
        acc[0] = acc[0] + wv * wv;
    }
    if (last)
        reduce_finalize<1>(acc, partials, &st->cnt[3], blockIdx.x, gridDim.x,
                           [&](const double *tot) { h[j + 1] = tot[0]; });
}

// f(std::integral_constant<int, c>) for a count 1 <= c <= 8.
template <class F>
__host__ void with_count8(int c, F &&f)
{
    switch (c) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    default: f(std::integral_constant<int, 8>{}); break;
    }
}

// x += sum_q y[q] z_q (FP64 accumulation of the FP32 directions).
__global__ __launch_bounds__(kBlock) void k_fg_update(long N, double *__restrict__ x, const float *const *zs,
                                                      const double *y, int k)
{
    for (long n = (long)blockIdx.x * kBlock + threadIdx.x; n < N; n += (long)gridDim.x * kBlock) {
        double acc = x[n];
        for (int q = 0; q < k; ++q) acc = acc + y[q] * (double)zs[q][n];
        x[n] = acc;
    }
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
