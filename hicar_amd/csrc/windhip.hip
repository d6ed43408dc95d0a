// windhip.hip -- host side of the MI355X wind solver: the C-ABI of
// include/windhip.h, device state of one nest, and the orchestration of
// calc_iter_winds (src/physics/wind_iterative.F90:295-532).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "windhip.h"
#include "windhip_kernels.hpp"
#include "windhip_setup.hpp"
#include "windhip_mg.hpp"
#include "windhip_net.hpp"

using namespace windhip;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(WINDHIP_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

#define CHK(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != WINDHIP_OK)    \
            return rc_;           \
    } while (0)

int flat_grid(long n)
{
    long g = (n + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min(g, 16384L));
}

}  // namespace

// Scratch of the allocation-free entry point windhip_balance_uvw_noalloc: the
// FP64 horizontal divergence of one call, in the caller's mass-grid layout
// (ims:ime, kms:kme, jms:jme). A device global, so every GPU the library runs
// on holds its own copy (128 MiB) until the process ends. Its address on a
// device is taken once, by hipGetSymbolAddress with that device current, when
// the first handle on the device is created; host code uses only that address
// (the array's name is a host-side stand-in, not a device address). Calls that
// overlap on one GPU share it and race.
constexpr long kNoAllocScratchDoubles = 16L << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(WINDHIP_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(WINDHIP_ERR_HIP, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(WINDHIP_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return WINDHIP_OK;
}
}  // namespace

struct windhip_solver {
    windhip_desc desc{};
    Dims d{};
    Geo g{};
    long N = 0;
    long guard = 0;  // elements before and after every row array (mx + 1 rounded up to a multiple of 4)
    hipStream_t stream = nullptr;
    hipStream_t net_stream = nullptr;  // the learned guess, concurrent with the first inner solve
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_net0 = nullptr, ev_net1 = nullptr;
    double *partials2 = nullptr;       // reduction partials of the guess residual (second stream)
    size_t bytes = 0;
    std::vector<void *> allocs;

    // metrics on the row grid (FP64) and planes
    double *m_dzdx, *m_dzdy, *m_jaco, *m_jaco_w, *m_alpha, *m_adv_dz, *m_dz_if;
    double *m_dzdx_u, *m_jaco_u, *m_dzdy_v, *m_jaco_v;
    double *p_terrain, *p_sx, *p_sy;
    double *mf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // mx_u, my_u, my_v, mx_v, mxy
    bool mf_on[5] = {false, false, false, false, false};             // given by the last set_metrics
    // operator
    double *B, *C, *H, *I, *L, *M;
    float *B32, *C32, *H32, *I32, *L32, *M32;
    // bf16 copies of B32..M32 for the multigrid inner solve, refreshed after assembly
    unsigned *Qh = nullptr;  // bf16 coefficient rows of the multigrid inner solve, 3 words per row (OpAB)
    bool coef_bf16_stale = true;
    double *bcx, *bcy, *bcz, *dztop;
    float *bcx32, *bcy32, *bcz32, *dztop32;
    double *P15[15] = {};
    float *P15_32[15] = {};
    double *dinv;
    float *dinv32;
    // Krylov FP64
    double *x, *r, *rh, *p, *ph[3], *v, *s, *sh[3], *t, *b;
    // Krylov FP32 (inner solve of the refinement)
    float *x32 = nullptr, *r32, *rh32, *p32, *ph32[3], *v32, *s32, *sh32[3], *t32;
    // FGMRES-IR (mixed precision): basis, FP32 directions, Hessenberg columns
    static constexpr int kMaxRestart = 40;
    static constexpr int kRestart = 32;  // FGMRES cycle length (directions kept per cycle)
    static_assert(kRestart < kFgMaxBasis, "the Gram-Schmidt passes address the whole basis");
    int fg_m = 0;
    double *fg_v[kMaxRestart + 1] = {};
    float *fg_z[kMaxRestart] = {};
    double *fg_part = nullptr, *fg_h = nullptr, *fg_y = nullptr;
    float **fg_zs = nullptr;
    // reductions and state
    double *partials;
    double *scan_part, *scan_out;
    unsigned *scan_cnt;
    KState *st;       // [0] FP64 solve / outer, [1] FP32 inner, [2] learned guess residual (second stream)
    KState *st_host;  // pinned mirror
    double *scan_host;
    // staging for host-memory fields (caller layout, caller dtype)
    void *stage[6] = {};
    size_t stage_bytes[6] = {};
    double *scratch_uvw = nullptr;  // probing: u, v, w, T in natural layout
    unsigned long long *leak = nullptr;
    double **cf_dev = nullptr;

    // learned initial guess: attached network, its padded input/output volumes
    windhip_net *net = nullptr;
    float *net_out = nullptr;
    __bf16 *net_in = nullptr;  // the learned start's network input, bf16
    long net_vox = 0;
    int net_smooth = -1;  // line-relaxation steps applied to the network's guess (-1: 10 with the
                          // multigrid's one-pass steps, else 4)
    double *lr_c = nullptr, *lr_d = nullptr;   // line relaxation: Thomas coefficients
    // the learned guess's FP64 residual b - A guess and its squared norm, taken
    // on the guess's stream for the mixed solves: the first FGMRES residual
    // when the guess becomes the start
    double *lg_res = nullptr, *lg_beta = nullptr;
    double *zn_part = nullptr;                 // block partials of the norm of b (zero start)
    // x holds +0.0 in every row (set_lambda of zeros, probe_finalize): a start
    // from it needs no FP64 matvec
    bool x_zero_known = false;
    float *lr_y32 = nullptr, *lr_o32 = nullptr, *lr_b32 = nullptr;  // FP32 relaxation of the mixed solves' guess
    // MIXED_LINE: FP32 Thomas factors of the column tridiagonal and scratch
    float *lf_lo = nullptr, *lf_cp = nullptr, *lf_ip = nullptr, *lf_dq = nullptr, *lf_r = nullptr;
    // bf16 copies of the line factors for the multigrid fine level
    unsigned short *lf_lo16 = nullptr, *lf_cp16 = nullptr, *lf_ip16 = nullptr;
    unsigned *lf_lip16 = nullptr;  // {lo | ip << 16} bf16 pairs (columns of up to 44 levels)
    // MIXED_MG: the coarse levels of the semi-coarsening hierarchy (the fine
    // level is the solver's grid with the line factors lf_*): averaged
    // metrics, the rediscretised analytic operator (FP64 and FP32), its BC
    // planes and line factors, and the level's correction, rhs, residual and
    // spare vectors (x and t with stencil guards)
    struct MgLevel {
        MgGeo g;
        Geo gg;   // the level as a tile / column grid
        Dims d;   // set-up dimensions (dx of the level)
        double *dzdx, *dzdy, *jaco, *jaco_w, *alpha, *adv_dz, *dz_if, *dzdx_u, *jaco_u, *dzdy_v, *jaco_v;
        double *terrain, *sx, *sy;
        double *B, *C, *H, *I, *L, *M, *bcx, *bcy, *bcz, *dztop;
        float *B32, *C32, *H32, *I32, *L32, *M32, *bcx32, *bcy32, *bcz32, *dztop32;
        float *lo, *cp, *ip, *x, *b, *r, *t;
        // bf16 copies for the two-kernel smoothing step (levels of up to 44
        // rows per column): interleaved coefficient rows (OpAB), {lo | ip}
        // factor pairs and the cp factor row
        unsigned *Q, *lip;
        unsigned short *cp16;
        // levels smoothed a wavefront per column (at most mg_wave_cols columns
        // of up to 64 rows) keep their vectors b, r, x, t column-major (row
        // c mz + k) and column-major copies of the FP32 rows and line factors
        bool cm;
        float *cB, *cC, *cH, *cI, *cL, *cM, *clo, *ccp, *cip;
        // the level's operator is the Galerkin product of the finer level's
        // (OpG: 27 column-major planar coefficient rows in G; column-major
        // vectors, line factors clo, ccp, cip from its vertical part)
        bool gal;
        float *G;
        long S() const { return cm ? g.mz : 0; }  // MgIdx stride of the level's vectors
    };
    std::vector<MgLevel> mg;
    int mg_pre = 2, mg_post = 2;    // fine-level line-smoothing steps before and after the coarse correction
    int mg_cpre = 1, mg_cpost = 1;  // the same on the coarse levels
    int mg_depth = 0;               // coarse levels used (0: down to the last level with >= 4 columns each way)
    float mg_omega = 0.9f;          // damping of the line smoother
    // Coarse levels of at most this many columns (64 x 64 interior columns and
    // their ring) smooth in one launch, a wavefront per column. Larger levels
    // keep the residual + line-solve kernels: a wavefront reads 64 rows of one
    // column per load instruction, a cache line per lane, which costs more
    // than the two serial marches once a level has enough columns to hide
    // their latency (512 x 512 x 40: 66 x 66 columns 4 us faster per step,
    // 130 x 130 columns 17 us slower).
    long mg_wave_cols = 4608;
    // Coarse levels from mg[mg_gal_from] on take the Galerkin product of the
    // finer level's operator instead of the rediscretised one (0: none). The
    // first coarse level stays rediscretised: its 15-point bf16 rows keep the
    // largest coarse level's smoothing cheap, while the rediscretised
    // operators below it (averaged, smoothed terrain) make the cycle lose the
    // smooth error modes (a V-cycle alone diverges on steep terrain).
    int mg_gal_from = 1;
    OpG level_opg(const MgLevel &L) const { return OpG{L.G, L.g.N}; }
    template <typename T>
    OpA<T> level_op(const MgLevel &L) const;
    OpA<float> level_op_cm(const MgLevel &L) const;  // the column-major FP32 rows of a wavefront level
    bool metrics_set = false, structure_uploaded = false, probed = false;
    // Operator generation: moves whenever the operator or anything built from
    // it may change (first assembly after set_metrics, a bitwise change of
    // alpha, probing). The BC planes and diagonal, the line factors, their
    // bf16 copies and the multigrid hierarchy remember the generation they
    // were built for and are rebuilt only when it moved.
    unsigned long op_gen = 0, planes_gen = ~0ul, lf_gen = ~0ul, lf16_gen = ~0ul, mg_gen = ~0ul;
    bool alpha_changed = true;  // set by gather_alpha (scan_out[3]), read after the scan
    bool verbose = false;  // WINDHIP_VERBOSE=1: per-batch progress on stderr (verbose_solver, :274-275)
    int nsweeps = 2;  // precond_n_sweeps, BASE_PREC_SWEEPS (wind_iterative.F90:76)
    std::map<long, hipGraphExec_t> graphs;

    // ---------------------------------------------------------------- helpers
    template <typename T>
    int alloc(T **ptr, long count, bool zero = true, long guard_elems = -1)
    {
        // Arrays of N rows get `guard` zeroed elements before and after them: the
        // stencil kernels address the five columns C, W, E, S, N of every lane
        // from one shifted base (mx + 1 elements before the row), so lateral and
        // out-of-grid lanes touch at most mx + 1 <= guard elements beyond
        // either end; guard is a multiple of 4, which keeps the rows 16-byte
        // aligned for the flat kernels' 16-byte accesses.
        const long gd = guard_elems >= 0 ? guard_elems : (count == N ? guard : 0);
        size_t nb = sizeof(T) * (size_t)std::max(count + 2 * gd, 1L);
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, nb);
        if (e != hipSuccess) return fail(WINDHIP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", nb, hipGetErrorString(e));
        allocs.push_back(q);
        bytes += nb;
        if (zero || gd) HIPCHK(hipMemsetAsync(q, 0, nb, stream));
        *ptr = (T *)q + gd;
        return WINDHIP_OK;
    }

    size_t esize() const { return (size_t)desc.field_dtype; }
    int ni_mass() const { return desc.ime - desc.ims + 1; }
    int nk_mass() const { return desc.kme - desc.kms + 1; }
    int nj_mass() const { return desc.jme - desc.jms + 1; }
    // element counts of the caller's arrays
    long n_mass() const { return (long)ni_mass() * nk_mass() * nj_mass(); }
    long n_u() const { return (long)(ni_mass() + 1) * nk_mass() * nj_mass(); }
    long n_v() const { return (long)ni_mass() * nk_mass() * (nj_mass() + 1); }
    long n_2d(int kind) const
    {
        return kind == 1 ? (long)(ni_mass() + 1) * nj_mass() : kind == 2 ? (long)ni_mass() * (nj_mass() + 1)
                                                                      : (long)ni_mass() * nj_mass();
    }

    // Device view of a caller array. For host memory it is copied into staging
    // slot `slot`; `copy_in` false only sizes the slot (outputs).
    int view(const void *host_or_dev, long count, int slot, bool copy_in, void **dev)
    {
        if (desc.field_mem == WINDHIP_DEVICE) {
            *dev = const_cast<void *>(host_or_dev);
            return WINDHIP_OK;
        }
        size_t nb = (size_t)count * esize();
        if (stage_bytes[slot] < nb) {
            if (stage[slot]) {
                HIPCHK(hipStreamSynchronize(stream));
                HIPCHK(hipFree(stage[slot]));
                bytes -= stage_bytes[slot];
            }
            HIPCHK(hipMalloc(&stage[slot], nb));
            stage_bytes[slot] = nb;
            bytes += nb;
        }
        if (copy_in) HIPCHK(hipMemcpyAsync(stage[slot], host_or_dev, nb, hipMemcpyHostToDevice, stream));
        *dev = stage[slot];
        return WINDHIP_OK;
    }

    int copy_out(void *host_or_dev, long count, int slot)
    {
        if (desc.field_mem == WINDHIP_DEVICE) return WINDHIP_OK;
        HIPCHK(hipMemcpyAsync(host_or_dev, stage[slot], (size_t)count * esize(), hipMemcpyDeviceToHost, stream));
        return WINDHIP_OK;
    }

    template <typename F>
    Fld<F> fld(void *p, int kind) const
    {
        // kind 0 mass, 1 u-grid, 2 v-grid
        Fld<F> f;
        f.p = (F *)p;
        f.i0 = desc.ims; f.k0 = desc.kms; f.j0 = desc.jms;
        f.ni = ni_mass() + (kind == 1 ? 1 : 0);
        f.nk = nk_mass();
        return f;
    }

    Metrics metrics() const
    {
        Metrics m;
        m.dzdx = m_dzdx; m.dzdy = m_dzdy; m.jaco = m_jaco; m.jaco_w = m_jaco_w; m.alpha = m_alpha;
        m.adv_dz = m_adv_dz; m.dz_if = m_dz_if; m.dzdx_u = m_dzdx_u; m.jaco_u = m_jaco_u;
        m.dzdy_v = m_dzdy_v; m.jaco_v = m_jaco_v; m.sx = p_sx; m.sy = p_sy;
        m.mf_mx_u = mf_on[0] ? mf[0] : nullptr; m.mf_my_u = mf_on[1] ? mf[1] : nullptr;
        m.mf_my_v = mf_on[2] ? mf[2] : nullptr; m.mf_mx_v = mf_on[3] ? mf[3] : nullptr;
        m.mf_mxy = mf_on[4] ? mf[4] : nullptr;
        return m;
    }

    template <typename T>
    OpA<T> opa() const;
    OpAB opab() const;
    int refresh_coef_bf16();
    // the multigrid inner solve reads bf16 coefficient rows (OpAB) unless the
    // operator was probed or its 12-byte rows exceed one buffer resource
    bool use_bf16_rows() const { return !probed && N * 12 < (1L << 32); }
    template <typename T>
    OpP<T> opp() const;
    // call f with the operator in use (analytic or probed), chosen at compile time in the kernels
    template <typename T, class F>
    int with_op(F &&f) const
    {
        if (probed) return f(opp<T>());
        return f(opa<T>());
    }

    // gather a caller field into the row grid (FP64)
    int gather3(const void *src, int kind, int ia, int ib, int ka, int kb, int ja, int jb, double *dst,
                int slot, double scale = 1.0, double *changed = nullptr)
    {
        long cnt = kind == 1 ? n_u() : kind == 2 ? n_v() : n_mass();
        void *dv = nullptr;
        CHK(view(src, cnt, slot, true, &dv));
        long tot = (long)(ib - ia + 1) * (kb - ka + 1) * (jb - ja + 1);
        if (desc.field_dtype == WINDHIP_F32)
            k_gather3<float><<<flat_grid(tot), kBlock, 0, stream>>>(d, fld<float>(dv, kind), dst, ia, ib, ka, kb, ja, jb, scale, changed);
        else
            k_gather3<double><<<flat_grid(tot), kBlock, 0, stream>>>(d, fld<double>(dv, kind), dst, ia, ib, ka, kb, ja, jb, scale, changed);
        HIPCHK(hipGetLastError());
        return WINDHIP_OK;
    }

    int gather2(const void *src, int kind, int ia, int ib, int ja, int jb, double *dst, int slot)
    {
        void *dv = nullptr;
        CHK(view(src, n_2d(kind), slot, true, &dv));
        long tot = (long)(ib - ia + 1) * (jb - ja + 1);
        Fld<float> ff;
        Fld<double> fd;
        ff.p = (float *)dv; fd.p = (double *)dv;
        ff.i0 = fd.i0 = desc.ims; ff.j0 = fd.j0 = desc.jms; ff.k0 = fd.k0 = 0;
        ff.ni = fd.ni = ni_mass() + (kind == 1 ? 1 : 0);
        ff.nk = fd.nk = 1;
        if (desc.field_dtype == WINDHIP_F32)
            k_gather2<float><<<flat_grid(tot), kBlock, 0, stream>>>(d, ff, dst, ia, ib, ja, jb);
        else
            k_gather2<double><<<flat_grid(tot), kBlock, 0, stream>>>(d, fd, dst, ia, ib, ja, jb);
        HIPCHK(hipGetLastError());
        return WINDHIP_OK;
    }

    int init();
    int gather_alpha(const void *alpha);
    int assemble(bool full);
    int prepare_call(const void *alpha, const void *div);
    int refresh_operator(double amin, double amax);
    int scan(double *amin, double *amax, double *nans);
    int set_alpha(const void *alpha);
    int read_state(int idx);
    int run_plain(int max_iters, double tol_rel, double tol_abs, int batch, bool use_graph, int *status,
                  int *iters, double *res0, double *resf);
    int run_mixed(const windhip_solve_opts &o, windhip_solve_info *info, double *learned_guess = nullptr,
                  double plain_res = 0.0, bool x_zero = false);
    int true_residual(double *rel, const double *xv = nullptr);
    int net_dims(int *Dp, int *Hp, int *Wp) const;
    template <typename OT>
    int net_input(OT *in, double *scale, bool x_zero = false);
    int line_relax(double **y, double **scratch, int steps, hipStream_t strm);
    int line_relax32(double *y, int steps, hipStream_t strm, bool mg_rows = false);
    int learned_launch(double *res_plain, double **guess, bool f32_relax, bool mg_rows, bool x_zero,
                       const windhip_solve_opts &o);
    int prepare_fine_line(bool multigrid);
    int learned_join(double *res_guess);
    bool net_pending = false;
    template <typename T>
    int enqueue_iterations(int count);
    template <class Launch>
    int timed(int kind, Launch &&launch);
    bool prof_nested = false;  // a timed region is open (its inner launches get no events of their own)
    int harvest();
    template <typename T>
    void kvec(T **X, T **R, T **RH, T **P, T **V, T **S, T **Tv, T **PH, T **SH, const T **DI);
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    std::vector<int> ev_kind;
    int ev_next = 0;
    // Krylov launches after the solve has finished return at once: while
    // prof_gate is set, each timed launch first records the state's
    // done/fmode (k_prof_gate into prof_flag[event]) and the harvest leaves
    // such no-op launches out of the per-kind times and counts
    const KState *prof_gate = nullptr;
    int *prof_flag = nullptr;
    static constexpr int kProfFlags = 1 << 16;
    std::vector<char> ev_gated;
    struct GateScope {
        const KState *&g;
        GateScope(const KState *&g_, const KState *v) : g(g_) { g = v; }
        ~GateScope() { g = nullptr; }
    };
    double prof_ms[WINDHIP_PROFILE_KINDS] = {}, prof_calls[WINDHIP_PROFILE_KINDS] = {};
    int launch_batch(int precision, int count, bool use_graph, bool first = false);
    int enqueue_iterations_line(int count, bool multigrid, bool first);
    int build_mg();
    float *mg_coarse(size_t l, bool entered = false);
    bool mg_fused = true;  // fused descent/return kernels on the wavefront-per-column levels
    // one-pass kernels: smoothing steps with the residual and the line solve in
    // one kernel (k_smooth_r) on the fine level and the larger coarse levels
    // of up to 44 rows per column, and the fine level's residual with the
    // restriction (k_resid_restrict_c)
    bool mg_onepass = true;
    // The cycle from the first column-major level down as one persistent
    // launch (k_mg_tail) instead of a launch per step (off by default: slower
    // at 512 x 512 x 40): its plan (uploaded when the cycle's shape changes),
    // grid-barrier state, grid, the level it starts at (SIZE_MAX: off) and the
    // buffer that holds its result.
    bool mg_tail = false;
    MgTailPlan tail_host{};
    MgTailPlan *tail_plan = nullptr;
    MgGridBar *tail_bar = nullptr;
    int *tail_err_host = nullptr;
    int tail_blocks = 0;
    long tail_resident = 0;
    size_t tail_l0 = SIZE_MAX;
    float *tail_out = nullptr;
    size_t mg_used() const;
    float *plan_coarse(size_t l, bool entered, size_t l0, std::vector<MgTailStep> &ps);
    int prepare_tail(size_t l0);
    int check_tail();
    // form 0: residual + line solve (FP32 rows); 1: one launch, a wavefront
    // per column; 2: residual + line solve on the level's bf16 rows
    void mg_smooth(size_t l, int form, const float *y, float *out);
    OpAB level_opab(const MgLevel &L) const;
    int alloc_f32_krylov();
    int alloc_fgmres(int m);
    // the two classical Gram-Schmidt passes of FGMRES step j, over at most 8
    // basis vectors per launch (registers): the products with v_1..v_j (h_0
    // comes with the matvec), then the update with ||w||^2
    void fg_dots(const FgBasis &vb, int j, const double *w, double *h, int fb)
    {
        for (int i0 = 1; i0 <= j; i0 += 8)
            with_count8(std::min(8, j + 1 - i0), [&](auto c) {
                k_fg_dots<decltype(c)::value><<<fb, kThreads, 0, stream>>>(N, st, fg_part, w, vb, i0, h);
            });
    }
    void fg_orth(const FgBasis &vb, int j, double *w, double *h, int fb)
    {
        for (int i0 = 0; i0 <= j; i0 += 8) {
            const int cnt = std::min(8, j + 1 - i0);
            with_count8(cnt, [&](auto c) {
                k_fg_orth<decltype(c)::value><<<fb, kThreads, 0, stream>>>(N, st, fg_part, w, vb, i0, j,
                                                                           i0 + cnt > j, h);
            });
        }
    }
    int correction(void *u, void *v, void *w, const void *density, int adv_den);
    int divergence(const void *u, const void *v, const void *w, const void *density, int adv_den,
                   int horz_only, int use_dqdt, void *div_out);
    int probe(const void *alpha, const void *density, int adv_den, double *max_leak);
    int probe_alloc();
    int probe_finish();
    int probe_record_field(const void *T, int ca, int cb, int cc, double *max_leak);
    int balance_uvw(const void *u, const void *v, void *w, const void *density, int adv_den, int use_dqdt,
                    double *scr);
    int w_real(const void *u, const void *v, const void *w_grid, void *out);
    int idealized_wgrid(const void *u, const void *v, const void *w_real_in, void *w_out);
    double *div_scr = nullptr;  // balance_uvw: the horizontal divergence, FP64, caller mass layout
    double *noalloc_scr = nullptr;  // this device's `scratch` (windhip_balance_uvw_noalloc)
    void release();
};

template <>
OpA<double> windhip_solver::opa<double>() const
{
    OpA<double> o;
    o.B = B; o.C = C; o.H = H; o.I = I; o.L = L; o.M = M;
    o.bcx = bcx; o.bcy = bcy; o.bcz = bcz; o.dztop = dztop;
    o.inv = 1.0 / (desc.dx * desc.dx);
    o.neg4 = -4.0 * o.inv;
    return o;
}
template <>
OpA<float> windhip_solver::opa<float>() const
{
    OpA<float> o;
    o.B = B32; o.C = C32; o.H = H32; o.I = I32; o.L = L32; o.M = M32;
    o.bcx = bcx32; o.bcy = bcy32; o.bcz = bcz32; o.dztop = dztop32;
    double inv = 1.0 / (desc.dx * desc.dx);
    o.inv = (float)inv;
    o.neg4 = (float)(-4.0 * inv);
    return o;
}
OpAB windhip_solver::opab() const
{
    OpAB o;
    static_cast<OpA<float> &>(o) = opa<float>();
    o.Q = Qh;
    return o;
}
template <>
OpA<double> windhip_solver::level_op<double>(const MgLevel &L) const
{
    OpA<double> o;
    o.B = L.B; o.C = L.C; o.H = L.H; o.I = L.I; o.L = L.L; o.M = L.M;
    o.bcx = L.bcx; o.bcy = L.bcy; o.bcz = L.bcz; o.dztop = L.dztop;
    o.inv = 1.0 / (L.d.dx * L.d.dx);
    o.neg4 = -4.0 * o.inv;
    return o;
}
template <>
OpA<float> windhip_solver::level_op<float>(const MgLevel &L) const
{
    OpA<float> o;
    o.B = L.B32; o.C = L.C32; o.H = L.H32; o.I = L.I32; o.L = L.L32; o.M = L.M32;
    o.bcx = L.bcx32; o.bcy = L.bcy32; o.bcz = L.bcz32; o.dztop = L.dztop32;
    const double inv = 1.0 / (L.d.dx * L.d.dx);
    o.inv = (float)inv;
    o.neg4 = (float)(-4.0 * inv);
    return o;
}
OpA<float> windhip_solver::level_op_cm(const MgLevel &L) const
{
    OpA<float> o = level_op<float>(L);
    o.B = L.cB; o.C = L.cC; o.H = L.cH; o.I = L.cI; o.L = L.cL; o.M = L.cM;
    return o;
}
OpAB windhip_solver::level_opab(const MgLevel &L) const
{
    OpAB o;
    static_cast<OpA<float> &>(o) = level_op<float>(L);
    o.Q = L.Q;
    return o;
}
template <>
OpP<double> windhip_solver::opp<double>() const
{
    OpP<double> o;
    for (int q = 0; q < 15; ++q) o.cf[q] = P15[q];
    return o;
}
template <>
OpP<float> windhip_solver::opp<float>() const
{
    OpP<float> o;
    for (int q = 0; q < 15; ++q) o.cf[q] = P15_32[q];
    return o;
}

template <>
void windhip_solver::kvec<double>(double **X, double **R, double **RH, double **P, double **V, double **S,
                                  double **Tv, double **PH, double **SH, const double **DI)
{
    *X = x; *R = r; *RH = rh; *P = p; *V = v; *S = s; *Tv = t; *DI = dinv;
    for (int q = 0; q < 3; ++q) { PH[q] = ph[q]; SH[q] = sh[q]; }
}
template <>
void windhip_solver::kvec<float>(float **X, float **R, float **RH, float **P, float **V, float **S,
                                 float **Tv, float **PH, float **SH, const float **DI)
{
    *X = x32; *R = r32; *RH = rh32; *P = p32; *V = v32; *S = s32; *Tv = t32; *DI = dinv32;
    for (int q = 0; q < 3; ++q) { PH[q] = ph32[q]; SH[q] = sh32[q]; }
}

int windhip_solver::init()
{
    d.nx = desc.nx; d.ny = desc.ny; d.nz = desc.nz;
    d.mx = desc.nx + 2; d.my = desc.ny + 2; d.mz = desc.nz + 2;
    d.P = (long)d.mx * d.my;
    d.dx = desc.dx;
    g.mx = d.mx; g.my = d.my; g.mz = d.mz; g.P = d.P;
    g.nti = (d.mx + kTileI - 1) / kTileI;
    g.ntj = (d.my + kTileJ - 1) / kTileJ;
    g.nt = g.nti * g.ntj;
    N = d.P * d.mz;
    guard = (d.mx + 1 + 3) / 4 * 4;  // at least mx + 1, and row arrays 16-byte aligned
    if ((N + 2 * guard) * (long)sizeof(double) >= (1L << 32))
        return fail(WINDHIP_ERR_ARG, "windhip_create: %ld rows exceed the 4 GiB span of one buffer resource", N);
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&net_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    HIPCHK(hipEventCreateWithFlags(&ev_net0, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_net1, hipEventDisableTiming));
    double **rows64[] = {&m_dzdx, &m_dzdy, &m_jaco, &m_jaco_w, &m_alpha, &m_adv_dz, &m_dz_if,
                         &m_dzdx_u, &m_jaco_u, &m_dzdy_v, &m_jaco_v, &B, &C, &H, &I, &L, &M,
                         &dinv, &x, &r, &rh, &p, &ph[0], &ph[1], &ph[2], &v, &s, &sh[0], &sh[1], &sh[2],
                         &t, &b};
    for (double **q : rows64) CHK(alloc(q, N));
    float **rows32[] = {&B32, &C32, &H32, &I32, &L32, &M32, &dinv32};
    for (float **q : rows32) CHK(alloc(q, N));
    double **pl64[] = {&p_terrain, &p_sx, &p_sy, &bcx, &bcy, &bcz, &dztop};
    for (double **q : pl64) CHK(alloc(q, d.P));
    float **pl32[] = {&bcx32, &bcy32, &bcz32, &dztop32};
    for (float **q : pl32) CHK(alloc(q, d.P));
    CHK(alloc(&partials, (long)std::max(g.nt, chunk_blocks(g)) * 8));
    CHK(alloc(&partials2, (long)std::max(g.nt, chunk_blocks(g)) * 8));
    CHK(alloc(&scan_part, 3L * 1024));
    CHK(alloc(&scan_out, 4));
    CHK(alloc(&scan_cnt, 4));
    CHK(alloc(&st, 3));
    HIPCHK(hipHostMalloc((void **)&st_host, 3 * sizeof(KState), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&scan_host, 4 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(stream));
    return WINDHIP_OK;
}

void windhip_solver::release()
{
    // both streams drained before anything they may still read is freed
    if (stream) (void)hipStreamSynchronize(stream);
    if (net_stream) (void)hipStreamSynchronize(net_stream);
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
    for (void *q : allocs) (void)hipFree(q);
    allocs.clear();
    for (int q = 0; q < 6; ++q)
        if (stage[q]) (void)hipFree(stage[q]);
    if (st_host) (void)hipHostFree(st_host);
    if (scan_host) (void)hipHostFree(scan_host);
    if (tail_err_host) (void)hipHostFree(tail_err_host);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_net0) (void)hipEventDestroy(ev_net0);
    if (ev_net1) (void)hipEventDestroy(ev_net1);
    if (net_stream) (void)hipStreamDestroy(net_stream);
    if (stream) (void)hipStreamDestroy(stream);
}

int windhip_solver::alloc_f32_krylov()
{
    if (x32) return WINDHIP_OK;
    float **rows32[] = {&x32, &r32, &rh32, &p32, &ph32[0], &ph32[1], &ph32[2], &v32, &s32,
                        &sh32[0], &sh32[1], &sh32[2], &t32};
    for (float **q : rows32) CHK(alloc(q, N));
    return WINDHIP_OK;
}

// the bf16 coefficient rows of the multigrid inner solve, from the FP32 rows
int windhip_solver::refresh_coef_bf16()
{
    if (!coef_bf16_stale) return WINDHIP_OK;
    if (!Qh) CHK(alloc(&Qh, 3 * N));
    k_to_bf16_rows<<<flat_grid(N), kBlock, 0, stream>>>(B32, C32, H32, I32, L32, M32, Qh, N);
    HIPCHK(hipGetLastError());
    coef_bf16_stale = false;
    return WINDHIP_OK;
}

int windhip_solver::assemble(bool full)
{
    coef_bf16_stale = true;
    long tot = (long)d.nx * d.ny * d.nz;
    k_assemble<<<flat_grid(tot), kBlock, 0, stream>>>(d, metrics(), full ? 1 : 0, B, C, H, I, L, M, B32, C32,
                                                       H32, I32, L32, M32);
    HIPCHK(hipGetLastError());
    return WINDHIP_OK;
}

// alpha into m_alpha, noting in scan_out[3] (read back with the scan)
// whether any value changed bitwise since the last call.
int windhip_solver::gather_alpha(const void *alpha)
{
    HIPCHK(hipMemsetAsync(scan_out + 3, 0, sizeof(double), stream));
    return gather3(alpha, 0, 1, d.nx, 1, d.nz, 1, d.ny, m_alpha, 0, 1.0, scan_out + 3);
}

// Operator for the current alpha: first-call assembly (initialize_coefs) or
// the alpha-range refresh (update_coefs_gpu, wind_iterative.F90:410-427), then
// BC planes and the diagonal (extract_diagonal, :430). alpha must be gathered.
// The refresh runs when alpha varies (the reference's test) and differs from
// the alpha the coefficients were made from; with the same alpha it would
// reproduce them bit for bit, so it is skipped, and so are the BC planes and
// the diagonal while the operator generation has not moved.
int windhip_solver::refresh_operator(double amin, double amax)
{
    if (!structure_uploaded) {
        CHK(assemble(true));
        structure_uploaded = true;
        ++op_gen;
    } else if (alpha_changed) {
        if (amax > amin && !probed) CHK(assemble(false));
        ++op_gen;  // the BC rows and the coarse levels read alpha
    }
    if (planes_gen == op_gen) return WINDHIP_OK;
    planes_gen = op_gen;
    long pl = (long)d.nx * d.ny;
    k_bc_planes<<<flat_grid(pl), kBlock, 0, stream>>>(d, metrics(), bcx, bcy, bcz, dztop, bcx32, bcy32,
                                                       bcz32, dztop32);
    HIPCHK(hipGetLastError());
    k_diagonal<<<flat_grid(N), kBlock, 0, stream>>>(d, probed ? 1 : 0, P15[0], B, C, B32, C32, P15_32[0], bcz,
                                                    dztop, bcz32, dztop32, dinv, dinv32);
    HIPCHK(hipGetLastError());
    return WINDHIP_OK;
}

int windhip_solver::scan(double *amin, double *amax, double *nans)
{
    const int nblk = 1024;
    k_scan_inputs<<<nblk, kBlock, 0, stream>>>(d, m_alpha, b, scan_part, scan_cnt, scan_out, nblk);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(scan_host, scan_out, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    *amin = scan_host[0];
    *amax = scan_host[1];
    *nans = scan_host[2];
    alpha_changed = scan_host[3] != 0.0;
    return WINDHIP_OK;
}

int windhip_solver::set_alpha(const void *alpha)
{
    if (!metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_set_alpha: windhip_set_metrics has not been called");
    CHK(gather_alpha(alpha));
    double amin, amax, nans;
    CHK(scan(&amin, &amax, &nans));
    return refresh_operator(amin, amax);
}

// calc_iter_winds up to the solve (wind_iterative.F90:309-433)
int windhip_solver::prepare_call(const void *alpha, const void *div)
{
    if (!metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_solve: windhip_set_metrics has not been called");
    CHK(gather_alpha(alpha));
    void *dv = nullptr;
    CHK(view(div, n_mass(), 1, true, &dv));
    long tot = (long)d.nx * d.ny * d.nz;
    if (desc.field_dtype == WINDHIP_F32)
        k_rhs<float><<<flat_grid(tot), kBlock, 0, stream>>>(d, fld<const float>(dv, 0), b);
    else
        k_rhs<double><<<flat_grid(tot), kBlock, 0, stream>>>(d, fld<const double>(dv, 0), b);
    HIPCHK(hipGetLastError());
    double amin, amax, nans;
    CHK(scan(&amin, &amax, &nans));
    if (nans > 0) return fail(WINDHIP_ERR_NAN, "NaN in input divergence (count=%.0f)", nans);
    return refresh_operator(amin, amax);
}

int windhip_solver::read_state(int idx)
{
    (void)idx;
    HIPCHK(hipMemcpyAsync(st_host, st, 2 * sizeof(KState), hipMemcpyDeviceToHost, stream));  // [2] is read at the join
    HIPCHK(hipStreamSynchronize(stream));
    if (profiling) CHK(harvest());
    return WINDHIP_OK;
}

// Kernel kinds timed by the profiling mode (hipEvents around each launch on
// the handle's stream): base + 5 * precision (0 FP64, 1 FP32).
enum { PK_SWEEP = 0, PK_SIGMA = 1, PK_UPD_S = 2, PK_OMEGA = 3, PK_UPD_XRP = 4, PK_INIT = 10,
       PK_IR_RESID = 11, PK_IR_ACC = 12, PK_TRUE_RESID = 13, PK_FG_SETUP = 14, PK_FG_MATVEC = 15,
       PK_FG_GS = 16, PK_LINE_SOLVE = 17, PK_RESID = 18, PK_MV_SIGMA = 19, PK_MV_OMEGA = 20,
       PK_MG_TRANSFER = 21, PK_MG_COARSE = 22, PK_MG_SETUP = 23, PK_SMOOTH = 24, PK_RESID_RESTRICT = 25, PK_COUNT = 26 };

template <class Launch>
int windhip_solver::timed(int kind, Launch &&launch)
{
    if (!profiling || prof_nested) {  // inside a timed region: its launches are timed as a whole
        launch();
        return WINDHIP_OK;
    }
    if (ev_next >= (int)ev_pool.size()) {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        ev_pool.push_back({e0, e1});
    }
    auto &pr = ev_pool[ev_next];
    const bool gated = prof_gate && prof_flag && ev_next < kProfFlags;
    if (gated) k_prof_gate<<<1, 64, 0, stream>>>(prof_gate, prof_flag + ev_next);
    ev_gated.push_back(gated);
    HIPCHK(hipEventRecord(pr.first, stream));
    prof_nested = true;
    launch();
    prof_nested = false;
    HIPCHK(hipEventRecord(pr.second, stream));
    ev_kind.push_back(kind);
    ++ev_next;
    return WINDHIP_OK;
}

// Accumulate the timed launches after the stream has been synchronized.
int windhip_solver::harvest()
{
    std::vector<int> flag;
    if (prof_flag && ev_next > 0) {
        flag.resize(std::min(ev_next, kProfFlags));
        HIPCHK(hipMemcpy(flag.data(), prof_flag, flag.size() * sizeof(int), hipMemcpyDeviceToHost));
    }
    for (int q = 0; q < ev_next; ++q) {
        const int kind = ev_kind[q];
        if (ev_gated[q]) {
            // a no-op launch: the solve was done before it (bit 0), or, for the
            // x/r/p update, no update mode was set (bit 1 clear; a finished
            // solve's closing half-step still updates x)
            const bool xrp = kind == PK_UPD_XRP || kind == 5 + PK_UPD_XRP;
            if (xrp ? !(flag[q] & 2) : (flag[q] & 1)) continue;
        }
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev_pool[q].first, ev_pool[q].second));
        prof_ms[kind] += ms;
        prof_calls[kind] += 1;
    }
    ev_next = 0;
    ev_kind.clear();
    ev_gated.clear();
    return WINDHIP_OK;
}

// One BiCGStab iteration per repetition (wind_iterative.F90:645-786):
// sweeps 2..n of the preconditioner on p, v = A p_hat with sigma, s and its
// first sweep, sweeps on s, t = A s_hat with the five dots, the x/r/p update.
template <typename T>
int windhip_solver::enqueue_iterations(int count)
{
    const int prec = sizeof(T) == 8 ? 0 : 1;
    KState *sp = st + prec;
    T *X, *R, *RH, *Pv, *V, *S, *Tv, *PH[3], *SH[3];
    const T *DI;
    kvec<T>(&X, &R, &RH, &Pv, &V, &S, &Tv, PH, SH, &DI);
    const int fg = flat_grid(N);
    const int base = 5 * prec;
    GateScope gate(prof_gate, sp);
    return with_op<T>([&](auto op) -> int {
        using Op = decltype(op);
        for (int it = 0; it < count; ++it) {
            // sweeps 2..n-1 alone, the last sweep fused with v = A p_hat and sigma
            T *cur = PH[0];
            int q = 1;
            for (int sw = 2; sw < nsweeps; ++sw) {
                T *nxt = PH[q];
                CHK(timed(base + PK_SWEEP, [&] { k_sweep<Op><<<g.nt, kThreads, 0, stream>>>(op, g, sp, cur, Pv, nxt); }));
                cur = nxt;
                q = (q == 1) ? 2 : 1;
            }
            T *phat = PH[q];
            CHK(timed(base + PK_SIGMA, [&] {
                k_sweep_sigma<Op><<<g.nt, kThreads, 0, stream>>>(op, g, sp, partials, cur, Pv, phat, RH, V);
            }));
            CHK(timed(base + PK_UPD_S, [&] { k_update_s<T><<<fg, kBlock, 0, stream>>>(N, sp, R, V, DI, S, SH[0]); }));
            T *cs = SH[0];
            q = 1;
            for (int sw = 2; sw < nsweeps; ++sw) {
                T *nxt = SH[q];
                CHK(timed(base + PK_SWEEP, [&] { k_sweep<Op><<<g.nt, kThreads, 0, stream>>>(op, g, sp, cs, S, nxt); }));
                cs = nxt;
                q = (q == 1) ? 2 : 1;
            }
            T *shat = SH[q];
            CHK(timed(base + PK_OMEGA, [&] {
                k_sweep_omega<Op><<<g.nt, kThreads, 0, stream>>>(op, g, sp, partials, cs, S, shat, RH, Tv);
            }));
            CHK(timed(base + PK_UPD_XRP, [&] {
                k_update_xrp<T><<<fg, kBlock, 0, stream>>>(N, sp, X, phat, shat, S, Tv, R, Pv, V, DI, PH[0]);
            }));
        }
        HIPCHK(hipGetLastError());
        return WINDHIP_OK;
    });
}

// One preconditioned FP32 BiCGStab iteration per repetition: p_hat = M^-1 p,
// v = A p_hat with sigma, s = r - alpha v, s_hat = M^-1 s, t = A s_hat with
// the five dots, the x/r/p update. M^-1 is two vertical line-relaxation steps
// (line solve, residual, line solve) or, with multigrid, one V-cycle.
int windhip_solver::enqueue_iterations_line(int count, bool multigrid, bool first)
{
    KState *sp = st + 1;
    float *X, *R, *RH, *Pv, *V, *S, *Tv, *PH[3], *SH[3];
    const float *DI;
    kvec<float>(&X, &R, &RH, &Pv, &V, &S, &Tv, PH, SH, &DI);
    const int fg = flat_grid(N);
    const int cb = col_blocks(g), kb = chunk_blocks(g);
    GateScope gate(prof_gate, sp);
    auto body = [&](auto op) -> int {
        using Op = decltype(op);
        // M^-1 in: y1 = L^-1 in, r = in - A y1, out = y1 + L^-1 r
        // the multigrid fine level reads its factors as bf16 (a smoother of a
        // slightly perturbed column system), the line arm as FP32
        auto lsolve = [&](const float *in, const float *y, float *out, float w) {
            if (multigrid && g.mz <= 44)
                k_line_solve_r<44, unsigned, 4><<<cb, kColThreads, 0, stream>>>(
                    g, sp, in, lf_lip16, reinterpret_cast<const unsigned *>(lf_cp16), nullptr, y, out, w);
            else if (multigrid && g.mz <= 48)
                k_line_solve_r<48, unsigned short><<<cb, kColThreads, 0, stream>>>(g, sp, in, lf_lo16, lf_cp16,
                                                                                   lf_ip16, y, out, w);
            else if (multigrid && g.mz <= 64)
                k_line_solve_r<64, unsigned short><<<cb, kColThreads, 0, stream>>>(g, sp, in, lf_lo16, lf_cp16,
                                                                                   lf_ip16, y, out, w);
            else if (g.mz <= 48)
                k_line_solve_r<48><<<cb, kColThreads, 0, stream>>>(g, sp, in, lf_lo, lf_cp, lf_ip, y, out, w);
            else if (g.mz <= 64)
                k_line_solve_r<64><<<cb, kColThreads, 0, stream>>>(g, sp, in, lf_lo, lf_cp, lf_ip, y, out, w);
            else
                k_line_solve_c<<<cb, kColThreads, 0, stream>>>(g, sp, in, lf_lo, lf_cp, lf_ip, lf_dq, y, out, w);
        };
        auto resid = [&](const float *y, const float *in, float *r) {
            return timed(PK_RESID, [&] { k_resid_c<Op><<<kb, kColThreads, 0, stream>>>(op, g, sp, y, in, r); });
        };
        // one multigrid smoothing step out = y + w L^-1 (in - A y): one pass of
        // k_smooth_r where the columns fit it, else the residual and the line solve
        auto smooth = [&](const float *in, const float *y, float *out, float w) -> int {
            if (mg_onepass && g.mz <= 44)
                return timed(PK_SMOOTH, [&] {
                    k_smooth_r<Op, 44, unsigned><<<cb, kColThreads, 0, stream>>>(
                        op, g, sp, in, y, lf_lip16, reinterpret_cast<const unsigned *>(lf_cp16), nullptr, out, w);
                });
            CHK(resid(y, in, lf_r));
            return timed(PK_LINE_SOLVE, [&] { lsolve(lf_r, y, out, w); });
        };
        const MgGeo f0 = mg_geo(g.mx - 2, g.my - 2, g.mz);
        // out = M^-1 in, y1 and lf_dq as scratch
        auto precond = [&](const float *in, float *y1, float *out) -> int {
            if (!multigrid) {
                CHK(timed(PK_LINE_SOLVE, [&] { lsolve(in, nullptr, y1, 1.0f); }));
                CHK(resid(y1, in, lf_r));
                CHK(timed(PK_LINE_SOLVE, [&] { lsolve(lf_r, y1, out, 1.0f); }));
                return WINDHIP_OK;
            }
            // V-cycle from a zero guess: damped line smoothing, residual,
            // restriction, coarse levels, prolongation, line smoothing
            float *cur = y1, *alt = lf_dq;
            CHK(timed(PK_LINE_SOLVE, [&] { lsolve(in, nullptr, cur, mg_omega); }));
            for (int q = 1; q < mg_pre; ++q) {
                CHK(smooth(in, cur, alt, mg_omega));
                std::swap(cur, alt);
            }
            if (!mg.empty()) {
                if (mg_onepass) {  // residual and restriction in one pass
                    CHK(timed(PK_RESID_RESTRICT, [&] {
                        k_resid_restrict_c<Op><<<(int)((mg[0].g.P + kColThreads - 1) / kColThreads) * kRRChunks,
                                                 kColThreads, 0, stream>>>(op, g, mg[0].g, sp, cur, in, mg[0].b,
                                                                           mg[0].S());
                    }));
                } else {
                    CHK(resid(cur, in, lf_r));
                    CHK(timed(PK_MG_TRANSFER, [&] {
                        k_mg_restrict<<<mg_plane_grid(mg[0].g), kMgThreads, 0, stream>>>(f0, mg[0].g, sp, lf_r,
                                                                                          mg[0].b, 0, mg[0].S());
                    }));
                }
                // the coarse levels as one timed region (per-launch events would
                // add their own gaps to the microsecond launches there)
                float *e = nullptr;
                CHK(timed(PK_MG_COARSE, [&] { e = mg_coarse(0); }));
                CHK(timed(PK_MG_TRANSFER, [&] {
                    k_mg_prolong<<<mg_plane_grid(f0), kMgThreads, 0, stream>>>(f0, mg[0].g, sp, e, cur, 0, mg[0].S());
                }));
            }
            for (int q = 0; q < mg_post; ++q) {
                float *dst = q == mg_post - 1 ? out : alt;
                CHK(smooth(in, cur, dst, mg_omega));
                alt = cur;
                cur = dst;
            }
            return WINDHIP_OK;
        };
        for (int it = 0; it < count; ++it) {
            // the first iteration from the set-up: r = p = r_hat, x = 0
            const bool f = first && it == 0;
            CHK(precond(f ? RH : Pv, PH[1], PH[2]));
            CHK(timed(PK_MV_SIGMA, [&] {
                k_mv_sigma_c<Op><<<kb, kColThreads, 0, stream>>>(op, g, sp, partials, PH[2], RH, V);
            }));
            CHK(timed(5 + PK_UPD_S, [&] {
                k_update_s_only<float><<<fg, kBlock, 0, stream>>>(N, sp, f ? RH : R, V, S);
            }));
            CHK(precond(S, SH[1], SH[2]));
            CHK(timed(PK_MV_OMEGA, [&] {
                k_mv_omega_c<Op><<<kb, kColThreads, 0, stream>>>(op, g, sp, partials, SH[2], S, RH, Tv);
            }));
            CHK(timed(5 + PK_UPD_XRP, [&] {
                if (f)
                    k_update_xrp<float, true><<<fg, kBlock, 0, stream>>>(N, sp, X, PH[2], SH[2], S, Tv, R, Pv, V,
                                                                           nullptr, nullptr, RH);
                else
                    k_update_xrp<float><<<fg, kBlock, 0, stream>>>(N, sp, X, PH[2], SH[2], S, Tv, R, Pv, V,
                                                                     nullptr, nullptr);
            }));
        }
        HIPCHK(hipGetLastError());
        return WINDHIP_OK;
    };
    // the multigrid inner solve runs on the bf16-coefficient operator (the
    // rows are refreshed by run_mixed before any batch is enqueued or captured)
    if (multigrid && use_bf16_rows()) return body(opab());
    return with_op<float>(body);
}

// One damped line-relaxation step on coarse level l, out = y + w L^-1 (b - A y)
// with b = mg[l].b (y null: out = w L^-1 b). form 1: one launch, a wavefront
// per column (k_mg_smooth_w; on a column-major level, L.cm, its vectors are
// column-major); 0 and 2: a residual into mg[l].r and a line
// solve on the level's FP32 (0) or bf16 (2) rows and factors; 3 and 4: one
// pass of k_smooth_r on the bf16 (3) or FP32 (4) rows and factors (levels of
// up to 44 rows per column, y given; otherwise as 2 and 0).
void windhip_solver::mg_smooth(size_t l, int form, const float *y, float *out)
{
    MgLevel &L = mg[l];
    const KState *sp = st + 1;
    const Geo &gl = L.gg;
    if (form == 1 && gl.mz <= 64) {
        const int nb = (int)((L.g.P + kWaveCols - 1) / kWaveCols);
        timed(PK_MG_COARSE, [&] {
            if (L.gal)  // the Galerkin operator (column-major)
                k_mg_smooth_w<OpG, true><<<nb, 64 * kWaveCols, 0, stream>>>(
                    level_opg(L), gl, sp, L.b, y, L.clo, L.ccp, L.cip, out, mg_omega);
            else if (L.cm)  // the level's vectors are column-major
                k_mg_smooth_w<OpA<float>, true><<<nb, 64 * kWaveCols, 0, stream>>>(
                    level_op_cm(L), gl, sp, L.b, y, L.clo, L.ccp, L.cip, out, mg_omega);
            else
                k_mg_smooth_w<OpA<float>><<<nb, 64 * kWaveCols, 0, stream>>>(level_op<float>(L), gl, sp, L.b, y,
                                                                           L.lo, L.cp, L.ip, out, mg_omega);
        });
        return;
    }
    if (form >= 3 && y && gl.mz <= 44 && (form == 4 || L.Q)) {
        // one pass: the residual feeds the line solve in registers
        const int nb = (int)((L.g.P + kSmoothSmallNT - 1) / kSmoothSmallNT);
        timed(PK_MG_COARSE, [&] {
            if (form == 3)
                k_smooth_r<OpAB, 44, unsigned, kSmoothSmallNT><<<nb, kSmoothSmallNT, 0, stream>>>(
                    level_opab(L), gl, sp, L.b, y, L.lip, reinterpret_cast<const unsigned *>(L.cp16), nullptr, out,
                    mg_omega);
            else
                k_smooth_r<OpA<float>, 44, float, kSmoothSmallNT><<<nb, kSmoothSmallNT, 0, stream>>>(
                    level_op<float>(L), gl, sp, L.b, y, L.lo, L.cp, L.ip, out, mg_omega);
        });
        return;
    }
    const bool half = (form == 2 || form == 3) && L.Q;
    const float *in = L.b;
    if (y) {
        timed(PK_MG_COARSE, [&] {
            if (half)
                k_resid_c<OpAB><<<chunk_blocks(gl), kColThreads, 0, stream>>>(level_opab(L), gl, sp, y, L.b, L.r);
            else
                k_resid_c<OpA<float>><<<chunk_blocks(gl), kColThreads, 0, stream>>>(level_op<float>(L), gl, sp, y,
                                                                                   L.b, L.r);
        });
        in = L.r;
    }
    const int cb = col_blocks(gl);
    timed(PK_MG_COARSE, [&] {
        if (half)
            k_line_solve_r<44, unsigned, 4><<<cb, kColThreads, 0, stream>>>(
                gl, sp, in, L.lip, reinterpret_cast<const unsigned *>(L.cp16), nullptr, y, out, mg_omega);
        else if (gl.mz <= 48)
            k_line_solve_r<48><<<cb, kColThreads, 0, stream>>>(gl, sp, in, L.lo, L.cp, L.ip, y, out, mg_omega);
        else if (gl.mz <= 64)
            k_line_solve_r<64><<<cb, kColThreads, 0, stream>>>(gl, sp, in, L.lo, L.cp, L.ip, y, out, mg_omega);
        else
            k_mg_line<<<mg_blocks(L.g.P), kMgThreads, 0, stream>>>(L.g, sp, L.lo, L.cp, L.ip, in, y, out, mg_omega);
    });
}

// Coarse levels the V-cycle uses.
size_t windhip_solver::mg_used() const
{
    if (mg_depth > 0) return std::min(mg.size(), (size_t)mg_depth);
    // the deepest level that still has 4 interior columns each way: below it
    // the cell-centred levels misplace the lateral boundary by half a coarse
    // cell and their correction of the smoothest error hurts
    size_t q = 0;
    while (q < mg.size() && std::min(mg[q].g.nx, mg[q].g.ny) >= 4) ++q;
    return std::max<size_t>(q, 1);
}

// The steps mg_coarse enqueues from column-major level l down (every level
// from l0 on column-major, fused transfers, mg_cpost > 0), as the steps of
// k_mg_tail, in the same order with the same buffers; returns the buffer that
// holds the correction of level l.
float *windhip_solver::plan_coarse(size_t l, bool entered, size_t l0, std::vector<MgTailStep> &ps)
{
    MgLevel &L = mg[l];
    const int lv = (int)(l - l0);
    auto step = [&](int kind, float w, const float *y, const float *e, float *out) {
        MgTailStep t{};
        t.kind = kind; t.level = lv; t.w = w; t.y = y; t.e = e; t.out = out;
        ps.push_back(t);
    };
    if (l + 1 == mg.size()) {  // one interior column: its line solve is exact
        if (!entered) step(kTailSmooth, 1.0f, nullptr, nullptr, L.x);
        return L.x;
    }
    float *cur = L.x, *alt = L.t;
    if (!entered) step(kTailSmooth, mg_omega, nullptr, nullptr, cur);
    if (l + 1 == mg_used()) {
        for (int q = 1; q < mg_cpre + mg_cpost; ++q) {
            step(kTailSmooth, mg_omega, cur, nullptr, alt);
            std::swap(cur, alt);
        }
        return cur;
    }
    for (int q = 1; q < mg_cpre; ++q) {
        step(kTailSmooth, mg_omega, cur, nullptr, alt);
        std::swap(cur, alt);
    }
    const MgLevel &C = mg[l + 1];
    step(kTailDown, l + 2 == mg.size() ? 1.0f : mg_omega, cur, nullptr, C.x);
    float *e = plan_coarse(l + 1, true, l0, ps);
    step(kTailUp, mg_omega, cur, e, alt);
    std::swap(cur, alt);
    for (int q = 1; q < mg_cpost; ++q) {
        step(kTailSmooth, mg_omega, cur, nullptr, alt);
        std::swap(cur, alt);
    }
    return cur;
}

// Make mg_coarse(l0) one k_mg_tail launch when the cycle from l0 down fits
// it, else leave the per-step launches (tail_l0 = SIZE_MAX). Uploads the plan
// when it changed; never called while the stream is captured.
int windhip_solver::prepare_tail(size_t l0)
{
    tail_l0 = SIZE_MAX;
    if (!mg_tail || !mg_fused || mg_cpost <= 0 || l0 >= mg.size() || mg.size() - l0 > (size_t)kTailMaxLevels)
        return WINDHIP_OK;
    for (size_t l = l0; l < mg.size(); ++l)
        if (!mg[l].cm || mg[l].gal) return WINDHIP_OK;  // the persistent steps read the analytic rows
    std::vector<MgTailStep> ps;
    float *res = plan_coarse(l0, false, l0, ps);
    if (ps.size() > (size_t)kTailMaxSteps) return WINDHIP_OK;
    if (!tail_plan) {
        int dev = 0, cus = 0, per_cu = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mg_tail, 64 * kTailWaves, 0));
        // blocks that can all be resident at once (the grid barrier waits for every block)
        tail_resident = (long)per_cu * cus;
        if (tail_resident < 1) return WINDHIP_OK;
        CHK(alloc(&tail_plan, 1));
        CHK(alloc(&tail_bar, 1));
        HIPCHK(hipHostMalloc((void **)&tail_err_host, sizeof(int), hipHostMallocDefault));
        *tail_err_host = 0;
        std::memset(&tail_host, 0, sizeof tail_host);
    }
    MgTailPlan h;
    std::memset(&h, 0, sizeof h);  // padding too: the plans are compared bytewise
    h.nsteps = (int)ps.size();
    h.nlevels = (int)(mg.size() - l0);
    for (size_t q = 0; q < ps.size(); ++q) h.step[q] = ps[q];
    for (int q = 0; q < h.nlevels; ++q) {
        const MgLevel &L = mg[l0 + (size_t)q];
        MgTailLevel &v = h.lv[q];
        v.op = level_op_cm(L);
        v.g = L.gg;
        v.m = L.g;
        v.lo = L.clo; v.cp = L.ccp; v.ip = L.cip;
        v.b = L.b;
    }
    if (std::memcmp(&h, &tail_host, sizeof h) != 0) {
        tail_host = h;
        HIPCHK(hipMemcpyAsync(tail_plan, &tail_host, sizeof tail_host, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    // a wavefront per column of the largest level, at most the resident grid
    const long need = (mg[l0].g.P + kTailWaves - 1) / kTailWaves;
    tail_blocks = (int)std::min<long>(std::min<long>(need, tail_resident), kTailMaxBlocks);
    if (tail_blocks < 1) return WINDHIP_OK;
    tail_out = res;
    tail_l0 = l0;
    return WINDHIP_OK;
}

// A grid barrier of k_mg_tail that ran out of time (sticky) is an error.
int windhip_solver::check_tail()
{
    if (!tail_bar) return WINDHIP_OK;
    HIPCHK(hipMemcpyAsync(tail_err_host, &tail_bar->err, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (*tail_err_host) {
        *tail_err_host = 0;
        HIPCHK(hipMemsetAsync(tail_bar, 0, sizeof(MgGridBar), stream));  // a fresh barrier for the next solve
        HIPCHK(hipStreamSynchronize(stream));
        return fail(WINDHIP_ERR_HIP, "multigrid: a grid barrier of the persistent coarse-level launch timed out");
    }
    return WINDHIP_OK;
}

// Enqueue the V-cycle below the fine level from coarse level l (rhs in
// mg[l].b); returns the buffer that holds its correction. entered: the level
// above has already written the rhs and the first smoothing step into mg[l].x
// (k_mg_down_w).
float *windhip_solver::mg_coarse(size_t l, bool entered)
{
    MgLevel &L = mg[l];
    const KState *sp = st + 1;
    const Geo &gl = L.gg;
    const int cb = col_blocks(gl), kb = chunk_blocks(gl);
    const OpA<float> op = level_op<float>(L);
    const bool wave = L.g.P <= mg_wave_cols || L.gal;
    auto line = [&](const float *in, const float *y, float *out, float w) {
        timed(PK_MG_COARSE, [&] {
            if (gl.mz <= 48)
                k_line_solve_r<48><<<cb, kColThreads, 0, stream>>>(gl, sp, in, L.lo, L.cp, L.ip, y, out, w);
            else if (gl.mz <= 64)
                k_line_solve_r<64><<<cb, kColThreads, 0, stream>>>(gl, sp, in, L.lo, L.cp, L.ip, y, out, w);
            else
                k_mg_line<<<mg_blocks(L.g.P), kMgThreads, 0, stream>>>(L.g, sp, L.lo, L.cp, L.ip, in, y, out, w);
        });
    };
    // the two-kernel levels smooth on their bf16 rows when they have them, in
    // one pass where their columns fit k_smooth_r
    const int form = wave ? 1 : (L.Q ? (mg_onepass ? 3 : 2) : 0);
    auto resid = [&](const float *xv) {
        timed(PK_MG_COARSE, [&] {
            if (L.gal)
                k_mg_resid_w<OpG, true><<<(int)((L.g.P + kWaveCols - 1) / kWaveCols), 64 * kWaveCols, 0, stream>>>(
                    level_opg(L), gl, sp, xv, L.b, L.r);
            else if (L.cm)
                k_mg_resid_w<OpA<float>, true><<<(int)((L.g.P + kWaveCols - 1) / kWaveCols), 64 * kWaveCols, 0,
                                                  stream>>>(level_op_cm(L), gl, sp, xv, L.b, L.r);
            else if (form >= 2)
                k_resid_c<OpAB><<<kb, kColThreads, 0, stream>>>(level_opab(L), gl, sp, xv, L.b, L.r);
            else
                k_resid_c<OpA<float>><<<kb, kColThreads, 0, stream>>>(op, gl, sp, xv, L.b, L.r);
        });
    };
    if (!entered && l == tail_l0) {  // this level and all below in one launch
        timed(PK_MG_COARSE, [&] {
            k_mg_tail<<<tail_blocks, 64 * kTailWaves, 0, stream>>>(tail_plan, sp, tail_bar);
        });
        return tail_out;
    }
    const size_t used = mg_used();
    if (l + 1 == mg.size()) {  // one interior column: its line solve is exact
        if (!entered) {
            if (L.cm)
                timed(PK_MG_COARSE, [&] {
                    k_mg_smooth_w<OpA<float>, true><<<(int)((L.g.P + kWaveCols - 1) / kWaveCols), 64 * kWaveCols, 0,
                                                       stream>>>(level_op_cm(L), gl, sp, L.b, nullptr, L.clo, L.ccp,
                                                                 L.cip, L.x, 1.0f);
                });
            else
                line(L.b, nullptr, L.x, 1.0f);
        }
        return L.x;
    }
    float *cur = L.x, *alt = L.t;
    if (!entered) mg_smooth(l, form, nullptr, cur);
    if (l + 1 == used) {  // the coarsest level in use: smoothing only
        for (int q = 1; q < mg_cpre + mg_cpost; ++q) {
            mg_smooth(l, form, cur, alt);
            std::swap(cur, alt);
        }
        return cur;
    }
    for (int q = 1; q < mg_cpre; ++q) {
        mg_smooth(l, form, cur, alt);
        std::swap(cur, alt);
    }
    const MgLevel &C = mg[l + 1];
    // between two levels of one wavefront per column (both column-major) the
    // descent (residual, restriction, the coarse level's first step) and the
    // return (prolongation, the first post-smoothing step) are one launch
    // each; the descent from a two-kernel level keeps its separate kernels
    // (measured at 512 x 512 x 40 with plane-major wavefront levels: 45 us
    // against 35 us for the three kernels)
    const bool fuse_down = mg_fused && L.cm && C.cm;
    const bool fuse_up = mg_fused && L.cm && C.cm && mg_cpost > 0;
    if (fuse_down) {
        const float wc = l + 2 == mg.size() ? 1.0f : mg_omega;  // the one-column level is solved exactly
        // both levels are column-major (L.cm, C.cm)
        timed(PK_MG_COARSE, [&] {
            const int nb = (int)((C.g.P + kWaveCols - 1) / kWaveCols);
            if (L.gal)
                k_mg_down_w<OpG, true><<<nb, 64 * kWaveCols, 0, stream>>>(level_opg(L), gl, C.g, sp, cur, L.b, C.clo,
                                                                         C.ccp, C.cip, C.b, C.x, wc);
            else
                k_mg_down_w<OpA<float>, true><<<nb, 64 * kWaveCols, 0, stream>>>(
                    level_op_cm(L), gl, C.g, sp, cur, L.b, C.clo, C.ccp, C.cip, C.b, C.x, wc);
        });
    } else if (mg_onepass && form == 3) {
        // a plane-major level with bf16 rows: residual and restriction in one pass
        timed(PK_MG_COARSE, [&] {
            k_resid_restrict_c<OpAB><<<(int)((C.g.P + kColThreads - 1) / kColThreads) * kRRChunks, kColThreads, 0,
                                       stream>>>(level_opab(L), gl, C.g, sp, cur, L.b, C.b, C.S());
        });
    } else {
        resid(cur);
        timed(PK_MG_COARSE, [&] {
            k_mg_restrict<<<mg_plane_grid(C.g), kMgThreads, 0, stream>>>(L.g, C.g, sp, L.r, C.b, L.S(), C.S());
        });
    }
    float *e = mg_coarse(l + 1, fuse_down);
    int post = mg_cpost;
    if (fuse_up) {
        timed(PK_MG_COARSE, [&] {  // both levels are column-major
            const int nb = (int)((L.g.P + kWaveCols - 1) / kWaveCols);
            if (L.gal)
                k_mg_up_w<OpG, true><<<nb, 64 * kWaveCols, 0, stream>>>(level_opg(L), gl, C.g, sp, cur, e, L.b, L.clo,
                                                                       L.ccp, L.cip, alt, mg_omega);
            else
                k_mg_up_w<OpA<float>, true><<<nb, 64 * kWaveCols, 0, stream>>>(
                    level_op_cm(L), gl, C.g, sp, cur, e, L.b, L.clo, L.ccp, L.cip, alt, mg_omega);
        });
        std::swap(cur, alt);
        --post;
    } else {
        timed(PK_MG_COARSE, [&] {
            k_mg_prolong<<<mg_plane_grid(L.g), kMgThreads, 0, stream>>>(L.g, C.g, sp, e, cur, L.S(), C.S());
        });
    }
    for (int q = 0; q < post; ++q) {
        mg_smooth(l, form, cur, alt);
        std::swap(cur, alt);
    }
    return cur;
}

// The coarse levels (allocated once) and their operators, rebuilt by every
// solve: the metrics and alpha averaged level by level, then the set-up chain
// of the fine operator at each level's size and dx (init_module_vars' dz_if
// and surface slopes, the analytic assembly, the BC planes) and the line
// factors.
int windhip_solver::build_mg()
{
    const MgGeo f0 = mg_geo(g.mx - 2, g.my - 2, g.mz);
    if (mg.empty()) {
        MgGeo f = f0;
        double dx = desc.dx;
        while (f.nx > 1 || f.ny > 1) {
            MgLevel L{};
            L.g = mg_geo((f.nx + 1) / 2, (f.ny + 1) / 2, g.mz);
            dx *= 2.0;
            L.gg.mx = L.g.mx; L.gg.my = L.g.my; L.gg.mz = L.g.mz; L.gg.P = L.g.P;
            L.gg.nti = (L.g.mx + kTileI - 1) / kTileI;
            L.gg.ntj = (L.g.my + kTileJ - 1) / kTileJ;
            L.gg.nt = L.gg.nti * L.gg.ntj;
            L.d.nx = L.g.nx; L.d.ny = L.g.ny; L.d.nz = L.g.mz - 2;
            L.d.mx = L.g.mx; L.d.my = L.g.my; L.d.mz = L.g.mz; L.d.P = L.g.P; L.d.dx = dx;
            const long n = L.g.N, pl = L.g.P;
            L.gal = mg_gal_from > 0 && (long)mg.size() >= mg_gal_from && L.g.mz <= 64;
            L.cm = (L.g.P <= mg_wave_cols || L.gal) && L.g.mz <= 64;
            // stencil guards of the level's iterates: mx + 1 columns
            const long gd = (L.g.mx + 1) * (L.cm ? (long)L.g.mz : 1L);
            for (double **a : {&L.dzdx, &L.dzdy, &L.jaco, &L.jaco_w, &L.alpha, &L.adv_dz, &L.dz_if, &L.dzdx_u,
                               &L.jaco_u, &L.dzdy_v, &L.jaco_v, &L.B, &L.C, &L.H, &L.I, &L.L, &L.M})
                CHK(alloc(a, n));
            for (double **a : {&L.terrain, &L.sx, &L.sy, &L.bcx, &L.bcy, &L.bcz, &L.dztop}) CHK(alloc(a, pl));
            for (float **a : {&L.B32, &L.C32, &L.H32, &L.I32, &L.L32, &L.M32, &L.lo, &L.cp, &L.ip, &L.b, &L.r})
                CHK(alloc(a, n));
            for (float **a : {&L.bcx32, &L.bcy32, &L.bcz32, &L.dztop32}) CHK(alloc(a, pl));
            CHK(alloc(&L.x, n, true, gd));  // stencil inputs of the level's residual
            CHK(alloc(&L.t, n, true, gd));
            if (L.cm)
                for (float **a : {&L.cB, &L.cC, &L.cH, &L.cI, &L.cL, &L.cM, &L.clo, &L.ccp, &L.cip}) CHK(alloc(a, n));
            if (L.gal) CHK(alloc(&L.G, 27 * n));
            mg.push_back(L);
            f = L.g;
        }
    }
    if (mg.empty()) return WINDHIP_OK;
    for (size_t l = 0; l < mg.size(); ++l) {
        MgLevel &L = mg[l];
        if (L.gal) {  // the Galerkin product of the finer level's operator, and its line factors
            const MgLevel &F = mg[l - 1];
            CHK(timed(PK_MG_SETUP, [&] {
                if (F.gal)
                    k_mg_rap<FineGal><<<mg_blocks(L.g.N), kMgThreads, 0, stream>>>(
                        FineGal{level_opg(F), F.g.mx, F.g.my, F.g.mz}, F.g, L.g, L.G);
                else
                    k_mg_rap<FineRedisc><<<mg_blocks(L.g.N), kMgThreads, 0, stream>>>(
                        FineRedisc{level_op<float>(F), F.g.P, F.g.mx, F.g.my, F.g.mz}, F.g, L.g, L.G);
                k_mg_factor_g<<<mg_blocks(L.g.P), kMgThreads, 0, stream>>>(L.g, level_opg(L), L.clo, L.ccp, L.cip);
            }));
            continue;
        }
        MgMetricsIn in{};
        if (l == 0) {
            in = MgMetricsIn{m_dzdx, m_dzdy, m_jaco, m_jaco_w, m_alpha, m_adv_dz, m_dzdx_u, m_jaco_u,
                             m_dzdy_v, m_jaco_v, p_terrain};
        } else {
            const MgLevel &F = mg[l - 1];
            in = MgMetricsIn{F.dzdx, F.dzdy, F.jaco, F.jaco_w, F.alpha, F.adv_dz, F.dzdx_u, F.jaco_u,
                             F.dzdy_v, F.jaco_v, F.terrain};
        }
        const MgMetricsOut out{L.dzdx, L.dzdy, L.jaco, L.jaco_w, L.alpha, L.adv_dz, L.dzdx_u, L.jaco_u,
                               L.dzdy_v, L.jaco_v, L.terrain};
        const MgGeo fg = l == 0 ? f0 : mg[l - 1].g;
        Metrics m{};
        m.dzdx = L.dzdx; m.dzdy = L.dzdy; m.jaco = L.jaco; m.jaco_w = L.jaco_w; m.alpha = L.alpha;
        m.adv_dz = L.adv_dz; m.dz_if = L.dz_if; m.dzdx_u = L.dzdx_u; m.jaco_u = L.jaco_u;
        m.dzdy_v = L.dzdy_v; m.jaco_v = L.jaco_v; m.sx = L.sx; m.sy = L.sy;
        const long tot = (long)L.d.nx * L.d.ny * L.d.nz, pl = (long)L.d.nx * L.d.ny;
        CHK(timed(PK_MG_SETUP, [&] {
            k_mg_coarsen<<<mg_blocks(L.g.N), kMgThreads, 0, stream>>>(fg, L.g, in, out, 0);
            k_module_vars<<<flat_grid(pl), kBlock, 0, stream>>>(L.d, L.adv_dz, L.dzdx, L.dzdy, L.terrain, L.dz_if,
                                                                 L.sx, L.sy);
            k_assemble<<<flat_grid(tot), kBlock, 0, stream>>>(L.d, m, 1, L.B, L.C, L.H, L.I, L.L, L.M, L.B32,
                                                               L.C32, L.H32, L.I32, L.L32, L.M32);
            k_bc_planes<<<flat_grid(pl), kBlock, 0, stream>>>(L.d, m, L.bcx, L.bcy, L.bcz, L.dztop, L.bcx32,
                                                               L.bcy32, L.bcz32, L.dztop32);
            k_line_factor<OpA<double>><<<L.gg.nt, kThreads, 0, stream>>>(level_op<double>(L), L.gg, L.lo, L.cp,
                                                                         L.ip);
        }));
        if (L.cm) {  // column-major copies for the wavefront kernels
            const long n = L.g.N;
            CHK(timed(PK_MG_SETUP, [&] {
                const float *src[] = {L.B32, L.C32, L.H32, L.I32, L.L32, L.M32, L.lo, L.cp, L.ip};
                float *dst[] = {L.cB, L.cC, L.cH, L.cI, L.cL, L.cM, L.clo, L.ccp, L.cip};
                for (int q = 0; q < 9; ++q)
                    k_mg_to_cm<<<flat_grid(n), kBlock, 0, stream>>>(L.g.P, L.g.mz, src[q], dst[q]);
            }));
        }
        if (L.g.mz <= 44) {  // bf16 copies for the two-kernel smoothing step
            const long n = L.g.N;
            if (!L.Q) {
                CHK(alloc(&L.Q, 3 * n));
                CHK(alloc(&L.lip, n));
                CHK(alloc(&L.cp16, n));
            }
            CHK(timed(PK_MG_SETUP, [&] {
                k_to_bf16_rows<<<flat_grid(n), kBlock, 0, stream>>>(L.B32, L.C32, L.H32, L.I32, L.L32, L.M32, L.Q, n);
                k_to_bf16_pair<<<flat_grid(n), kBlock, 0, stream>>>(L.lo, L.ip, L.lip, n);
                k_to_bf16<<<flat_grid(n), kBlock, 0, stream>>>(L.cp, L.cp16, n);
            }));
        }
    }
    HIPCHK(hipGetLastError());
    return WINDHIP_OK;
}

// Launch `count` iterations, from a cached hipGraph when use_graph.
// first: the batch opens an inner solve of the line or multigrid mode (its
// first iteration reads r and p from r_hat, x from zero)
int windhip_solver::launch_batch(int precision, int count, bool use_graph, bool first)
{
    first = first && (precision == 2 || precision == 3);
    auto enqueue = [&]() {
        if (precision == 0) return enqueue_iterations<double>(count);
        if (precision == 2 || precision == 3) return enqueue_iterations_line(count, precision == 3, first);
        return enqueue_iterations<float>(count);
    };
    if (!use_graph || profiling) return enqueue();
    const long key = ((long)count << 9) | ((long)first << 8) | ((long)nsweeps << 4) | ((long)precision << 1) |
                     (probed ? 1 : 0);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue();
        hipError_t e = hipStreamEndCapture(stream, &graph);
        if (rc != WINDHIP_OK) return rc;
        if (e != hipSuccess) return fail(WINDHIP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        hipGraphExec_t exec;
        HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(graph));
        it = graphs.emplace(key, exec).first;
    }
    HIPCHK(hipGraphLaunch(it->second, stream));
    return WINDHIP_OK;
}

// bicgstab_solve (wind_iterative.F90:555-791) on the device. The host only
// launches batches of iterations and reads the state after each batch.
int windhip_solver::run_plain(int max_iters, double tol_rel, double tol_abs, int batch, bool use_graph,
                              int *status, int *iters, double *res0, double *resf)
{
    k_state_reset<<<1, 1, 0, stream>>>(st, max_iters, tol_rel, tol_abs, 0.0);
    CHK(with_op<double>([&](auto op) {
        return timed(PK_INIT, [&] {
            k_init<decltype(op)><<<g.nt, kThreads, 0, stream>>>(op, g, st, partials, x, b, dinv, r, rh, p, ph[0]);
        });
    }));
    HIPCHK(hipGetLastError());
    CHK(read_state(0));
    int done_iters = 0;
    while (!st_host[0].done) {
        int n = std::min(batch, std::max(1, max_iters - done_iters));
        CHK(launch_batch(0, n, use_graph));
        CHK(read_state(0));
        done_iters += n;
        if (done_iters > max_iters + batch) break;  // defensive: the device sets done at max_iters
    }
    *status = st_host[0].status;
    *iters = st_host[0].iter;
    *res0 = st_host[0].res0;
    *resf = st_host[0].rnorm;
    return WINDHIP_OK;
}

int windhip_solver::true_residual(double *rel, const double *xv)
{
    const double *xx = xv ? xv : x;
    CHK(with_op<double>([&](auto op) {
        return timed(PK_TRUE_RESID, [&] {
            k_true_resid<decltype(op)><<<chunk_blocks(g), kColThreads, 0, stream>>>(op, g, st, partials, xx, b);
        });
    }));
    HIPCHK(hipGetLastError());
    CHK(read_state(0));
    double bn = st_host[0].outer_bnorm;
    *rel = bn > 0 ? st_host[0].outer_rnorm / bn : st_host[0].outer_rnorm;
    return WINDHIP_OK;
}

// ------------------------------------------------------------------ learned start
// Padded network volume of the grid (multiples of 8).
int windhip_solver::net_dims(int *Dp, int *Hp, int *Wp) const
{
    *Dp = (d.nz + 7) / 8 * 8;
    *Hp = (d.ny + 7) / 8 * 8;
    *Wp = (d.nx + 7) / 8 * 8;
    return WINDHIP_OK;
}

// Network input from the current rhs b, metrics and alpha. The rhs is scaled
// by 1 / rms(b) over the cells (the solution is linear in b); *scale returns
// the factor that turns the network output back into lambda, rms(b) dx^2.
template <typename OT>
int windhip_solver::net_input(OT *in, double *scale, bool x_zero)
{
    int Dp, Hp, Wp;
    net_dims(&Dp, &Hp, &Wp);
    if (x_zero) {  // the residual of a zero lambda is b: its norm only
        if (!zn_part) CHK(alloc(&zn_part, 2L * kFlatBlocks + 1));  // 2 partials per block, then ||b||^2
        const int fb = (int)std::min<long>(kFlatBlocks, (N + kThreads - 1) / kThreads);
        CHK(timed(PK_TRUE_RESID, [&] {
            k_fg_resid_zero<<<fb, kThreads, 0, stream>>>(N, st, zn_part, b, nullptr, zn_part + 2L * kFlatBlocks);
        }));
        HIPCHK(hipGetLastError());
        CHK(read_state(0));
    } else {
        double rel;
        CHK(true_residual(&rel, nullptr));  // also ||b||
    }
    const double bn = st_host[0].outer_bnorm;
    const double rms = bn / sqrt((double)d.nx * d.ny * d.nz);
    const double bs = rms > 0 ? 1.0 / rms : 0.0;
    k_net_input<OT><<<flat_grid((long)Hp * Wp), kBlock, 0, stream>>>(d, Dp, Hp, Wp, b, bs, m_dzdx, m_dzdy, m_jaco,
                                                                 m_alpha, m_adv_dz, in);
    HIPCHK(hipGetLastError());
    *scale = rms * d.dx * d.dx;
    return WINDHIP_OK;
}

// `steps` line-relaxation steps on *y against the current rhs (FP64 operator),
// alternating between *y and *scratch; *y points at the result.
int windhip_solver::line_relax(double **y, double **scratch, int steps, hipStream_t strm)
{
    if (steps <= 0) return WINDHIP_OK;
    if (!lr_c) {
        CHK(alloc(&lr_c, N));
        CHK(alloc(&lr_d, N));
    }
    for (int q = 0; q < steps; ++q) {
        CHK(with_op<double>([&](auto op) {
            k_line_relax<decltype(op)><<<g.nt, kThreads, 0, strm>>>(op, g, *y, b, lr_c, lr_d, *scratch);
            return WINDHIP_OK;
        }));
        HIPCHK(hipGetLastError());
        std::swap(*y, *scratch);
    }
    return WINDHIP_OK;
}

// The same relaxation in FP32 on the FP32 operator rows, for the guess of the
// mixed solves (whose own inner solves are FP32): y (FP64) is rounded, relaxed
// `steps` times and widened back in place.
int windhip_solver::line_relax32(double *y, int steps, hipStream_t strm, bool mg_rows)
{
    if (steps <= 0) return WINDHIP_OK;
    if (!lr_c) {
        CHK(alloc(&lr_c, N));
        CHK(alloc(&lr_d, N));
    }
    if (!lr_y32) {
        CHK(alloc(&lr_y32, N));  // stencil inputs: guarded and zeroed like every row array
        CHK(alloc(&lr_o32, N));
        CHK(alloc(&lr_b32, N));
    }
    // the FP64 Thomas buffers hold the FP32 ones (N doubles each)
    float *cq = reinterpret_cast<float *>(lr_c), *dq = reinterpret_cast<float *>(lr_d);
    k_round_f32<<<flat_grid(N), kBlock, 0, strm>>>(N, y, lr_y32);
    k_round_f32<<<flat_grid(N), kBlock, 0, strm>>>(N, b, lr_b32);
    float *cur = lr_y32, *nxt = lr_o32;
    if (mg_rows) {
        // the multigrid fine level's bf16 rows and factors (prepare_fine_line,
        // queued on the solver's stream before this stream was forked): one
        // pass of k_smooth_r per step, undamped, on the guess's own live state
        k_state_reset<<<1, 1, 0, strm>>>(st + 2, 1, 0.0, 0.0, 0.0);
        auto body = [&](auto op) -> int {
            for (int q = 0; q < steps; ++q) {
                k_smooth_r<decltype(op), 44, unsigned><<<col_blocks(g), kColThreads, 0, strm>>>(
                    op, g, st + 2, lr_b32, cur, lf_lip16, reinterpret_cast<const unsigned *>(lf_cp16), nullptr, nxt,
                    1.0f);
                std::swap(cur, nxt);
            }
            return WINDHIP_OK;
        };
        if (use_bf16_rows()) CHK(body(opab()));
        else CHK(with_op<float>(body));
    } else {
        for (int q = 0; q < steps; ++q) {
            CHK(with_op<float>([&](auto op) {
                k_line_relax<decltype(op)><<<g.nt, kThreads, 0, strm>>>(op, g, cur, lr_b32, cq, dq, nxt);
                return WINDHIP_OK;
            }));
            std::swap(cur, nxt);
        }
    }
    k_widen_f64<<<flat_grid(N), kBlock, 0, strm>>>(N, cur, y);
    HIPCHK(hipGetLastError());
    return WINDHIP_OK;
}

// The learned guess: the network's lambda after net_smooth vertical
// line-relaxation steps (FP64 for the plain solve, FP32 for the mixed ones),
// and the FP64 true residuals of that guess and
// of the plain start (the current lambda: the previous solution, or zero).
// learned_launch takes the plain start's residual on the solver's stream, then
// queues the network, the relaxation and the guess's residual on a second
// stream (its own state slot and partials), so the mixed solves can build
// their line factors and coarse levels meanwhile; learned_join makes the
// solver's stream wait for it and reads the guess's residual. The guess lands in t or s (FP64 vectors
// of the plain solve, unused by the mixed solves).
int windhip_solver::learned_launch(double *res_plain, double **guess, bool f32_relax, bool mg_rows, bool x_zero,
                                   const windhip_solve_opts &o)
{
    int Dp, Hp, Wp;
    net_dims(&Dp, &Hp, &Wp);
    const long vox = (long)Dp * Hp * Wp;
    if (vox > net_vox) {
        CHK(alloc(&net_in, vox * wnet::kInCh));
        CHK(alloc(&net_out, vox));
        net_vox = vox;
    }
    if (!lr_c) {
        CHK(alloc(&lr_c, N));
        CHK(alloc(&lr_d, N));
    }
    // the multigrid solve's relaxation of the guess reads the fine level's
    // bf16 rows and factors: queued first (nothing when they are current)
    if (mg_rows) CHK(prepare_fine_line(true));
    double scale = 0.0;
    CHK(net_input(net_in, &scale, x_zero));
    *res_plain = st_host[0].outer_rnorm;  // of the current lambda, from net_input's residual
    HIPCHK(hipEventRecord(ev_net0, stream));
    HIPCHK(hipStreamWaitEvent(net_stream, ev_net0, 0));
    const int rc = wnet_forward_dev_bf16(net, net_stream, net_in, wnet::Shape{1, Dp, Hp, Wp}, net_out);
    if (rc) return fail(rc, "learned start: network forward failed");
    k_net_output<<<flat_grid(d.P), kBlock, 0, net_stream>>>(d, Dp, Hp, Wp, net_out, scale, t);
    HIPCHK(hipGetLastError());
    double *gy = t, *gs = s;
    const int steps = net_smooth >= 0 ? net_smooth : (mg_rows ? 10 : 4);
    if (f32_relax) CHK(line_relax32(gy, steps, net_stream, mg_rows));
    else CHK(line_relax(&gy, &gs, steps, net_stream));
    if (f32_relax) {
        // the mixed solves: the guess's residual vector too, the first FGMRES
        // residual if the guess becomes the start (k_fg_resid, as there)
        if (!lg_res) {
            CHK(alloc(&lg_res, N));
            CHK(alloc(&lg_beta, 1));
        }
        k_state_reset<<<1, 1, 0, net_stream>>>(st + 2, 0, o.tol_rel, o.tol_abs, 0.0);
        CHK(with_op<double>([&](auto op) {
            k_fg_resid<decltype(op)><<<chunk_blocks(g), kColThreads, 0, net_stream>>>(op, g, st + 2, partials2, gy, b,
                                                                                      lg_res, lg_beta);
            return WINDHIP_OK;
        }));
    } else {
        CHK(with_op<double>([&](auto op) {
            k_true_resid<decltype(op)><<<chunk_blocks(g), kColThreads, 0, net_stream>>>(op, g, st + 2, partials2, gy, b);
            return WINDHIP_OK;
        }));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev_net1, net_stream));
    // per-kernel timing (opts.profile) brackets each launch with events on the
    // solver's stream: with the guess still running beside it those would
    // count the wait for the other stream's kernels, so a profiled solve
    // finishes the guess first
    if (profiling) HIPCHK(hipStreamWaitEvent(stream, ev_net1, 0));
    net_pending = true;
    *guess = gy;
    return WINDHIP_OK;
}

int windhip_solver::learned_join(double *res_guess)
{
    if (!net_pending) return WINDHIP_OK;
    HIPCHK(hipStreamWaitEvent(stream, ev_net1, 0));
    HIPCHK(hipMemcpyAsync(st_host + 2, st + 2, sizeof(KState), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    net_pending = false;
    if (res_guess) *res_guess = st_host[2].outer_rnorm;
    return WINDHIP_OK;
}

// The fine level's line-relaxation factors and, for the multigrid inner
// solve, the bf16 coefficient rows and bf16 factors, rebuilt on the solver's
// stream only when the operator generation moved (same metrics, alpha and
// operator: kept).
int windhip_solver::prepare_fine_line(bool multigrid)
{
    if (!lf_lo) {
        CHK(alloc(&lf_lo, N));
        CHK(alloc(&lf_cp, N));
        CHK(alloc(&lf_ip, N));
        CHK(alloc(&lf_dq, N));
        CHK(alloc(&lf_r, N));
    }
    if (lf_gen != op_gen) {
        CHK(with_op<double>([&](auto op) {
            k_line_factor<decltype(op)><<<g.nt, kThreads, 0, stream>>>(op, g, lf_lo, lf_cp, lf_ip);
            return WINDHIP_OK;
        }));
        HIPCHK(hipGetLastError());
        lf_gen = op_gen;
    }
    if (!multigrid) return WINDHIP_OK;
    if (use_bf16_rows()) CHK(refresh_coef_bf16());
    // bf16 factors of the fine-level line solve: a {lo | ip} pair row and a cp
    // row for columns of up to 44 levels, three rows above
    if (lf16_gen != op_gen) {
        if (!lf_cp16) CHK(alloc(&lf_cp16, N));
        k_to_bf16<<<flat_grid(N), kBlock, 0, stream>>>(lf_cp, lf_cp16, N);
        if (g.mz <= 44) {
            if (!lf_lip16) CHK(alloc(&lf_lip16, N));
            k_to_bf16_pair<<<flat_grid(N), kBlock, 0, stream>>>(lf_lo, lf_ip, lf_lip16, N);
        } else {
            if (!lf_lo16)
                for (unsigned short **a : {&lf_lo16, &lf_ip16}) CHK(alloc(a, N));
            k_to_bf16<<<flat_grid(N), kBlock, 0, stream>>>(lf_lo, lf_lo16, N);
            k_to_bf16<<<flat_grid(N), kBlock, 0, stream>>>(lf_ip, lf_ip16, N);
        }
        HIPCHK(hipGetLastError());
        lf16_gen = op_gen;
    }
    return WINDHIP_OK;
}

// Mixed precision: FGMRES-IR. Flexible GMRES(m) in FP64 on A lambda = b whose
// preconditioner is an FP32 BiCGStab solve of A z = v_j to inner_tol (the
// reference's iteration, wind_iterative.F90:555-791, on the FP32 operator and
// vectors). The Arnoldi basis, the Hessenberg least squares, the lambda update
// and the true residual are FP64; FGMRES minimises the FP64 residual over all
// directions of a cycle, so an inaccurate inner solve slows it down but cannot
// make the residual grow.
int windhip_solver::run_mixed(const windhip_solve_opts &o, windhip_solve_info *info, double *learned_guess,
                              double plain_res, bool x_zero)
{
    CHK(alloc_f32_krylov());
    const bool multigrid = o.precision == WINDHIP_MIXED_MG;
    const bool line = o.precision == WINDHIP_MIXED_LINE || multigrid;
    if (line) CHK(prepare_fine_line(multigrid));
    if (multigrid && mg_gen != op_gen) {
        CHK(build_mg());
        mg_gen = op_gen;
    }
    if (multigrid) {  // the column-major levels as one launch where they fit it
        size_t l0 = 0;
        while (l0 < mg.size() && !mg[l0].cm) ++l0;
        CHK(prepare_tail(l0));
    }
    // the learned guess, computed on the second stream while the set-up above
    // ran, becomes the start when its residual is below the plain start's
    if (learned_guess) {
        double rg = 0.0;
        CHK(learned_join(&rg));
        info->guess_res = rg;
        info->plain_res = plain_res;
        if (rg < plain_res) {
            HIPCHK(hipMemcpyAsync(x, learned_guess, N * sizeof(double), hipMemcpyDeviceToDevice, stream));
            info->start_used = WINDHIP_START_LEARNED;
        }
        if (verbose)
            fprintf(stderr, "[windhip] learned start: residual %.3e vs plain start %.3e -> %s\n", rg, plain_res,
                    rg < plain_res ? "network" : "plain");
    }
    // 0 (or out of range): one V-cycle preconditioned iteration mostly reaches
    // 0.3; the other inner solves run to 0.5
    const double inner_tol = (o.inner_tol > 0 && o.inner_tol < 1) ? o.inner_tol : (multigrid ? 0.3 : 0.5);
    const int m = std::max(1, std::min(o.max_outer, kRestart));
    CHK(alloc_fgmres(m));
    const int H = m + 2;  // stride of a Hessenberg column in fg_h
    const int fb = (int)std::min<long>(kFlatBlocks, (N + kThreads - 1) / kThreads);
    k_state_reset<<<1, 1, 0, stream>>>(st, 0, o.tol_rel, o.tol_abs, 0.0);
    double *beta2 = fg_h + (long)m * H;  // ||r||^2 of the cycle's residual
    auto resid = [&]() -> int {
        CHK(with_op<double>([&](auto op) {
            return timed(PK_IR_RESID, [&] {
                k_fg_resid<decltype(op)><<<chunk_blocks(g), kColThreads, 0, stream>>>(op, g, st, partials, x, b, fg_v[0], beta2);
            });
        }));
        HIPCHK(hipGetLastError());
        return read_state(0);
    };
    if (info->start_used == WINDHIP_START_LEARNED) {
        // the guess's residual, taken beside the network on its stream by the
        // same kernel and state set-up as resid() would use here
        std::swap(fg_v[0], lg_res);
        HIPCHK(hipMemcpyAsync(beta2, lg_beta, sizeof(double), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(st, st + 2, sizeof(KState), hipMemcpyDeviceToDevice, stream));
        st_host[0] = st_host[2];
    } else if (x_zero) {
        // from a zero lambda the first residual is b itself: no FP64 matvec
        CHK(timed(PK_IR_RESID, [&] { k_fg_resid_zero<<<fb, kThreads, 0, stream>>>(N, st, fg_part, b, fg_v[0], beta2); }));
        HIPCHK(hipGetLastError());
        CHK(read_state(0));
    } else {
        CHK(resid());
    }
    info->res0 = st_host[0].outer_rnorm;
    info->bnorm = st_host[0].outer_bnorm;
    int total_inner = 0, steps = 0, status = 0;
    std::vector<double> Hm((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), col(H);
    // With the multigrid preconditioner an inner solve usually ends within its
    // first batch of iterations. Its direction's FGMRES step (below) is then
    // enqueued right behind that batch, and one synchronisation reads both the
    // inner state and the Hessenberg column. When the inner solve had not
    // ended, the step's results are discarded and the step runs again once it
    // has (the same kernels on the same z_j, so the same values). Ahead only
    // while the previous direction's inner solve ended within one batch.
    bool ahead = multigrid;
    while (!st_host[0].outer_done && steps < o.max_outer && total_inner < o.max_iters) {
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = st_host[0].outer_rnorm;
        int k = 0;
        bool breakdown = false;
        for (int j = 0; j < m && steps < o.max_outer && total_inner < o.max_iters; ++j) {
            double *hc = fg_h + (long)j * H;
            // w = A z_j (with h_0 = <w, v_0>), orthogonalised against v_0..v_j by
            // classical Gram-Schmidt: the other inner products from one read of
            // w, then w -= sum h_i v_i with ||w||^2 in a second pass (2j + 4
            // vector passes against 4j + 3 for modified Gram-Schmidt; FP64, and
            // the true residual closes every cycle). The Hessenberg column j and
            // both states to the host.
            auto enqueue_step = [&]() -> int {
                CHK(with_op<double>([&](auto op) {
                    return timed(PK_FG_MATVEC, [&] {
                        k_fg_matvec<decltype(op)><<<chunk_blocks(g), kColThreads, 0, stream>>>(
                            op, g, st, partials, x32, fg_z[j], fg_v[0], fg_v[j + 1], hc);
                    });
                }));
                FgBasis vb{};
                for (int i = 0; i <= j; ++i) vb.v[i] = fg_v[i];
                CHK(timed(PK_FG_GS, [&] {
                    if (j > 0) fg_dots(vb, j, fg_v[j + 1], hc, fb);
                    fg_orth(vb, j, fg_v[j + 1], hc, fb);
                }));
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(col.data(), hc, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, stream));
                HIPCHK(hipMemcpyAsync(st_host, st, 2 * sizeof(KState), hipMemcpyDeviceToHost, stream));
                HIPCHK(hipStreamSynchronize(stream));
                return harvest();
            };
            bool stepped = false;  // col holds this direction's step
            {
            // z_j ~ A^-1 v_j in FP32 (v_j normalised in place first). An inner
            // solve whose FP32 recurrence overflowed is run again from z = 0 with
            // half the iterations it had reached, keeping a finite iterate.
            const double *vsq = j == 0 ? beta2 : fg_h + (long)(j - 1) * H + j;
            // inexact-Krylov relaxation: once the residual estimate is within
            // a few times the target, the inner solve of this direction only
            // needs to reduce its rhs by about target / estimate (with a
            // factor 2 margin; never tighter than inner_tol, never above 0.5)
            const double est = j == 0 ? st_host[0].outer_rnorm : fabs(gv[j]);
            const double tol_j =
                est > 0 ? std::max(inner_tol, std::min(0.5, 0.5 * st_host[0].outer_target / est)) : inner_tol;
            int cap = o.inner_max_iters, spent = 0;
            for (int attempt = 0;; ++attempt) {
                // every inner solve starts from a fresh state with this solve's
                // cap (o.inner_max_iters, halved on an overflow rerun)
                k_state_reset<<<1, 1, 0, stream>>>(st + 1, cap, 0.0, 0.0, tol_j);
                CHK(timed(PK_FG_SETUP, [&] {
                    k_fg_setup<<<fb, kThreads, 0, stream>>>(N, st + 1, fg_part, vsq, attempt == 0, fg_v[j],
                                                            line ? nullptr : dinv32, line ? nullptr : x32,
                                                            r32, rh32, p32, ph32[0]);
                }));
                HIPCHK(hipGetLastError());
                int done_iters = 0;
                if (ahead && attempt == 0) {
                    done_iters = std::min(o.iters_per_sync, std::max(1, cap));
                    CHK(launch_batch(3, done_iters, o.use_graph != 0, true));
                    CHK(enqueue_step());
                    // an inner solve that ended before its first iteration left
                    // z_j unwritten: its step is made again from z_j = 0 (below)
                    stepped = st_host[1].done != 0 && st_host[1].iter > 0;
                } else {
                    CHK(read_state(1));
                }
                while (!st_host[1].done) {
                    const int n = std::min(o.iters_per_sync, std::max(1, cap - done_iters));
                    CHK(launch_batch(multigrid ? 3 : (line ? 2 : 1), n, o.use_graph != 0, done_iters == 0));
                    CHK(read_state(1));
                    done_iters += n;
                    if (done_iters > cap + o.iters_per_sync) break;
                }
                if (multigrid) ahead = done_iters <= o.iters_per_sync;
                spent += st_host[1].iter;
                const double rn = st_host[1].rnorm;
                if (rn == rn && rn < 1e300) break;
                stepped = false;
                if (verbose)
                    fprintf(stderr, "[windhip] FP32 inner solve overflowed after %d iterations; rerun with %d\n",
                            st_host[1].iter, st_host[1].iter / 2);
                cap = st_host[1].iter / 2;
                if (cap < 1 || attempt >= 3) { breakdown = true; break; }
            }
            if (breakdown) { total_inner += spent; ++steps; break; }
            total_inner += spent;
            ++steps;
            }
            if (!stepped) {
                // the line and multigrid set-ups leave z_j (x32) to the first
                // iteration: a solve that ended before it has z_j = 0
                if (line && st_host[1].iter == 0) HIPCHK(hipMemsetAsync(x32, 0, N * sizeof(float), stream));
                CHK(enqueue_step());
            }
            // Hessenberg column j, previous Givens rotations, new rotation
            for (int i = 0; i <= j; ++i) Hm[(size_t)i * m + j] = col[i];
            const double hn = sqrt(std::max(col[j + 1], 0.0));
            for (int i = 0; i < j; ++i) {
                const double a = Hm[(size_t)i * m + j], c2 = Hm[(size_t)(i + 1) * m + j];
                Hm[(size_t)i * m + j] = cs[i] * a + sn[i] * c2;
                Hm[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * c2;
            }
            const double a = Hm[(size_t)j * m + j];
            const double rr = std::hypot(a, hn);
            cs[j] = rr > 0 ? a / rr : 1.0;
            sn[j] = rr > 0 ? hn / rr : 0.0;
            Hm[(size_t)j * m + j] = rr;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            k = j + 1;
            if (verbose)
                fprintf(stderr, "[windhip] FGMRES step %d: %d FP32 iterations (status %d), residual estimate %.3e "
                                "(target %.3e)\n", steps, st_host[1].iter, st_host[1].status, fabs(gv[j + 1]),
                        st_host[0].outer_target);
            if (!(rr > 0) || !(hn == hn) || !(rr == rr)) { k = j; breakdown = true; break; }  // drop column j
            if (fabs(gv[j + 1]) <= st_host[0].outer_target) break;
            if (hn <= 1e-14 * rr) break;  // happy breakdown: the basis spans the solution
        }
        if (k == 0) break;
        // y = R^-1 g, lambda += sum y_q z_q (FP64)
        for (int i = k - 1; i >= 0; --i) {
            double acc = gv[i];
            for (int q = i + 1; q < k; ++q) acc -= Hm[(size_t)i * m + q] * y[q];
            y[i] = Hm[(size_t)i * m + i] != 0 ? acc / Hm[(size_t)i * m + i] : 0.0;
        }
        HIPCHK(hipMemcpyAsync(fg_y, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, stream));
        CHK(timed(PK_IR_ACC, [&] {
            k_fg_update<<<flat_grid(N), kBlock, 0, stream>>>(N, x, (const float *const *)fg_zs, fg_y, k);
        }));
        HIPCHK(hipGetLastError());
        const double before = st_host[0].outer_rnorm;
        CHK(resid());
        if (verbose)
            fprintf(stderr, "[windhip] FGMRES cycle of %d steps: true residual %.3e -> %.3e\n", k, before,
                    st_host[0].outer_rnorm);
        // a cycle that made no progress (or broke down without progress) ends the solve
        if (!(st_host[0].outer_rnorm < before)) { status = 2; break; }
        (void)breakdown;
    }
    if (multigrid && tail_l0 != SIZE_MAX) CHK(check_tail());
    info->status = st_host[0].outer_done ? 0 : (status ? status : 1);
    info->iters = total_inner;
    info->outer_iters = steps;
    info->res_final = st_host[0].outer_rnorm;
    info->true_rel_res = info->bnorm > 0 ? st_host[0].outer_rnorm / info->bnorm : st_host[0].outer_rnorm;
    return WINDHIP_OK;
}

// Arnoldi basis (m + 1 FP64 vectors), m FP32 directions and the small arrays of
// FGMRES(m); kept for later solves.
int windhip_solver::alloc_fgmres(int m)
{
    if (fg_m >= m) return WINDHIP_OK;
    for (int q = fg_m + 1; q <= m; ++q) CHK(alloc(&fg_v[q], N));
    if (fg_m == 0) CHK(alloc(&fg_v[0], N));
    for (int q = fg_m; q < m; ++q) CHK(alloc(&fg_z[q], N));
    if (!fg_part) {
        CHK(alloc(&fg_part, (long)kFlatBlocks * 8));  // up to 8 products per block (Gram-Schmidt)
        CHK(alloc(&fg_h, (long)(kMaxRestart + 1) * (kMaxRestart + 2)));
        CHK(alloc(&fg_y, kMaxRestart));
        CHK(alloc(&fg_zs, kMaxRestart));
    }
    HIPCHK(hipMemcpyAsync(fg_zs, fg_z, sizeof(float *) * m, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    fg_m = m;
    return WINDHIP_OK;
}

int windhip_solver::correction(void *u, void *v, void *w, const void *density, int adv_den)
{
    void *du, *dv, *dw, *dr = nullptr;
    CHK(view(u, n_u(), 2, true, &du));
    CHK(view(v, n_v(), 3, true, &dv));
    CHK(view(w, n_mass(), 4, true, &dw));
    const int use_rho = (adv_den && density) ? 1 : 0;
    if (use_rho) CHK(view(density, n_mass(), 5, true, &dr));
    const Metrics m = metrics();
    long tu = (long)(d.nx + 1) * d.ny, tv = (long)d.nx * (d.ny + 1), tw = (long)d.nx * d.ny;
    if (desc.field_dtype == WINDHIP_F32) {
        k_correct_u<float><<<flat_grid(tu), kBlock, 0, stream>>>(d, m, x, fld<float>(du, 1), fld<const float>(dr, 0), use_rho);
        k_correct_v<float><<<flat_grid(tv), kBlock, 0, stream>>>(d, m, x, fld<float>(dv, 2), fld<const float>(dr, 0), use_rho);
        k_correct_w<float><<<flat_grid(tw), kBlock, 0, stream>>>(d, m, x, fld<float>(dw, 0), fld<const float>(dr, 0), use_rho);
    } else {
        k_correct_u<double><<<flat_grid(tu), kBlock, 0, stream>>>(d, m, x, fld<double>(du, 1), fld<const double>(dr, 0), use_rho);
        k_correct_v<double><<<flat_grid(tv), kBlock, 0, stream>>>(d, m, x, fld<double>(dv, 2), fld<const double>(dr, 0), use_rho);
        k_correct_w<double><<<flat_grid(tw), kBlock, 0, stream>>>(d, m, x, fld<double>(dw, 0), fld<const double>(dr, 0), use_rho);
    }
    HIPCHK(hipGetLastError());
    CHK(copy_out(u, n_u(), 2));
    CHK(copy_out(v, n_v(), 3));
    CHK(copy_out(w, n_mass(), 4));
    HIPCHK(hipStreamSynchronize(stream));
    return WINDHIP_OK;
}

int windhip_solver::divergence(const void *u, const void *v, const void *w, const void *density, int adv_den,
                               int horz_only, int use_dqdt, void *div_out)
{
    void *du, *dv, *dw, *dr = nullptr, *dd;
    CHK(view(u, n_u(), 2, true, &du));
    CHK(view(v, n_v(), 3, true, &dv));
    CHK(view(w, n_mass(), 4, true, &dw));
    const int use_rho = (adv_den && density) ? 1 : 0;
    if (use_rho) CHK(view(density, n_mass(), 5, true, &dr));
    CHK(view(div_out, n_mass(), 1, desc.field_mem == WINDHIP_HOST, &dd));
    long tw = (long)d.nx * d.ny;
    if (desc.field_dtype == WINDHIP_F32)
        k_divergence<float><<<flat_grid(tw), kBlock, 0, stream>>>(d, metrics(), fld<const float>(du, 1),
            fld<const float>(dv, 2), fld<const float>(dw, 0), fld<const float>(dr, 0), use_rho, adv_den,
            horz_only, use_dqdt, fld<float>(dd, 0));
    else
        k_divergence<double><<<flat_grid(tw), kBlock, 0, stream>>>(d, metrics(), fld<const double>(du, 1),
            fld<const double>(dv, 2), fld<const double>(dw, 0), fld<const double>(dr, 0), use_rho, adv_den,
            horz_only, use_dqdt, fld<double>(dd, 0));
    HIPCHK(hipGetLastError());
    CHK(copy_out(div_out, n_mass(), 1));
    HIPCHK(hipStreamSynchronize(stream));
    return WINDHIP_OK;
}

// balance_uvw (wind.F90:123-191): the horizontal divergence of (u, v)
// (calc_divergence with horz_only, :250-523; density-weighted whatever adv_den
// is, since balance_uvw leaves advect_density at its default) into FP64
// scratch, then calc_w (:193-234) writes w; adv_den selects calc_w's form.
int windhip_solver::balance_uvw(const void *u, const void *v, void *w, const void *density, int adv_den,
                                int use_dqdt, double *scr)
{
    void *du, *dv, *dw, *dr = nullptr;
    CHK(view(u, n_u(), 2, true, &du));
    CHK(view(v, n_v(), 3, true, &dv));
    CHK(view(w, n_mass(), 4, true, &dw));
    const int use_rho = density ? 1 : 0;
    if (use_rho) CHK(view(density, n_mass(), 5, true, &dr));
    const long tw = (long)d.nx * d.ny;
    // k_divergence writes every domain cell of the scratch that k_calc_w
    // reads, so the scratch needs no clearing between calls
    const Fld<double> dd = fld<double>(scr, 0);
    const Fld<const double> ddc{scr, dd.i0, dd.k0, dd.j0, dd.ni, dd.nk};
    auto run = [&](auto tag) {
        using F = decltype(tag);
        k_divergence<F, double><<<flat_grid(tw), kBlock, 0, stream>>>(d, metrics(), fld<const F>(du, 1),
            fld<const F>(dv, 2), fld<const F>(dw, 0), fld<const F>(dr, 0), use_rho, 1, 1, use_dqdt, dd);
        k_calc_w<F><<<flat_grid(tw), kBlock, 0, stream>>>(d, metrics(), ddc, fld<const F>(dr, 0), use_rho, adv_den,
                                                          fld<F>(dw, 0));
    };
    if (desc.field_dtype == WINDHIP_F32) run(float{});
    else run(double{});
    HIPCHK(hipGetLastError());
    CHK(copy_out(w, n_mass(), 4));
    HIPCHK(hipStreamSynchronize(stream));
    return WINDHIP_OK;
}

// calc_w_real (wind.F90:1175-1227); w_grid may be null (zero).
int windhip_solver::w_real(const void *u, const void *v, const void *w_grid, void *out)
{
    void *du, *dv, *dg = nullptr, *dout;
    CHK(view(u, n_u(), 2, true, &du));
    CHK(view(v, n_v(), 3, true, &dv));
    if (w_grid) CHK(view(w_grid, n_mass(), 4, true, &dg));
    CHK(view(out, n_mass(), 1, desc.field_mem == WINDHIP_HOST, &dout));
    const long tot = (long)d.nx * d.ny * d.nz;
    auto run = [&](auto tag) {
        using F = decltype(tag);
        k_w_real<F><<<flat_grid(tot), kBlock, 0, stream>>>(d, metrics(), fld<const F>(du, 1), fld<const F>(dv, 2),
                                                           fld<const F>(dg, 0), fld<F>(dout, 0));
    };
    if (desc.field_dtype == WINDHIP_F32) run(float{});
    else run(double{});
    HIPCHK(hipGetLastError());
    CHK(copy_out(out, n_mass(), 1));
    HIPCHK(hipStreamSynchronize(stream));
    return WINDHIP_OK;
}

// calc_idealized_wgrid (wind.F90:940-998).
int windhip_solver::idealized_wgrid(const void *u, const void *v, const void *w_real_in, void *w_out)
{
    void *du, *dv, *dwr, *dout;
    CHK(view(u, n_u(), 2, true, &du));
    CHK(view(v, n_v(), 3, true, &dv));
    CHK(view(w_real_in, n_mass(), 4, true, &dwr));
    CHK(view(w_out, n_mass(), 1, desc.field_mem == WINDHIP_HOST, &dout));
    const long tw = (long)d.nx * d.ny;
    auto run = [&](auto tag) {
        using F = decltype(tag);
        k_idealized_wgrid<F><<<flat_grid(tw), kBlock, 0, stream>>>(d, metrics(), fld<const F>(du, 1),
                                                                   fld<const F>(dv, 2), fld<const F>(dwr, 0),
                                                                   fld<F>(dout, 0));
    };
    if (desc.field_dtype == WINDHIP_F32) run(float{});
    else run(double{});
    HIPCHK(hipGetLastError());
    CHK(copy_out(w_out, n_mass(), 1));
    HIPCHK(hipStreamSynchronize(stream));
    return WINDHIP_OK;
}

// calibrate_projection_operator (wind.F90:1010-1099) with the probe hooks of
// wind_iterative.F90:1096-1248, on the device and on internal FP64 scratch
// (u, v, w, the response T and the density in the natural (i,k,j) layout).
// The 15 probed coefficient arrays (FP64 and their FP32 rounding), the leak
// counter and the natural-layout scratch of probing, made on first use.
int windhip_solver::probe_alloc()
{
    if (!P15[0]) {
        for (int q = 0; q < 15; ++q) CHK(alloc(&P15[q], N));
        for (int q = 0; q < 15; ++q) CHK(alloc(&P15_32[q], N));
        CHK(alloc(&cf_dev, 15));
        HIPCHK(hipMemcpyAsync(cf_dev, P15, 15 * sizeof(double *), hipMemcpyHostToDevice, stream));
        CHK(alloc(&leak, 1));
    }
    const long nu = (long)(d.nx + 1) * d.nz * d.ny, nv = (long)d.nx * d.nz * (d.ny + 1);
    const long nw = (long)d.nx * d.nz * d.ny;
    if (!scratch_uvw) CHK(alloc(&scratch_uvw, nu + nv + 3 * nw));
    return WINDHIP_OK;
}

// probe_finalize (wind_iterative.F90:1232-1248): the recorded operator is the
// solver's from now on; x_sol (the last pattern) is reset.
int windhip_solver::probe_finish()
{
    for (int q = 0; q < 15; ++q) k_round_f32<<<flat_grid(N), kBlock, 0, stream>>>(N, P15[q], P15_32[q]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(x, 0, N * sizeof(double), stream));
    HIPCHK(hipStreamSynchronize(stream));
    probed = true;
    structure_uploaded = true;
    ++op_gen;
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
    return WINDHIP_OK;
}

// probe_record (wind_iterative.F90:1171-1230): 2 T of the caller's response
// field into the coefficients its colouring reaches; *max_leak is raised to
// the largest |2 T| off the 15-point footprint.
int windhip_solver::probe_record_field(const void *T, int ca, int cb, int cc, double *max_leak)
{
    CHK(probe_alloc());
    const long nu = (long)(d.nx + 1) * d.nz * d.ny, nv = (long)d.nx * d.nz * (d.ny + 1);
    const long nw = (long)d.nx * d.nz * d.ny;
    double *sT = scratch_uvw + nu + nv + nw;
    void *dT = nullptr;
    CHK(view(T, n_mass(), 1, true, &dT));
    if (desc.field_dtype == WINDHIP_F32)
        k_to_natural<float><<<flat_grid(nw), kBlock, 0, stream>>>(d, fld<float>(dT, 0), sT);
    else
        k_to_natural<double><<<flat_grid(nw), kBlock, 0, stream>>>(d, fld<double>(dT, 0), sT);
    HIPCHK(hipMemsetAsync(leak, 0, sizeof(unsigned long long), stream));
    k_probe_record<<<flat_grid(nw), kBlock, 0, stream>>>(d, ca, cb, cc, sT, cf_dev, leak);
    HIPCHK(hipGetLastError());
    unsigned long long lk = 0;
    HIPCHK(hipMemcpyAsync(&lk, leak, sizeof(lk), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double lv;
    memcpy(&lv, &lk, sizeof(lv));
    if (max_leak && lv > *max_leak) *max_leak = lv;
    return WINDHIP_OK;
}

int windhip_solver::probe(const void *alpha, const void *density, int adv_den, double *max_leak)
{
    x_zero_known = false;  // the colourings are written into lambda
    if (!metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_probe_operator: metrics not set");
    CHK(gather3(alpha, 0, 1, d.nx, 1, d.nz, 1, d.ny, m_alpha, 0));
    CHK(probe_alloc());
    const long nu = (long)(d.nx + 1) * d.nz * d.ny, nv = (long)d.nx * d.nz * (d.ny + 1);
    const long nw = (long)d.nx * d.nz * d.ny;
    double *su = scratch_uvw, *sv = su + nu, *sw = sv + nv, *sT = sw + nw, *srho = sT + nw;
    const int use_rho = (adv_den && density) ? 1 : 0;
    if (use_rho) {
        void *dr = nullptr;
        CHK(view(density, n_mass(), 5, true, &dr));
        if (desc.field_dtype == WINDHIP_F32)
            k_to_natural<float><<<flat_grid(nw), kBlock, 0, stream>>>(d, fld<float>(dr, 0), srho);
        else
            k_to_natural<double><<<flat_grid(nw), kBlock, 0, stream>>>(d, fld<double>(dr, 0), srho);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemsetAsync(leak, 0, sizeof(unsigned long long), stream));
    const Fld<double> fu{su, 1, 1, 1, d.nx + 1, d.nz}, fv{sv, 1, 1, 1, d.nx, d.nz}, fw{sw, 1, 1, 1, d.nx, d.nz};
    const Fld<const double> cu{su, 1, 1, 1, d.nx + 1, d.nz}, cv{sv, 1, 1, 1, d.nx, d.nz}, cw{sw, 1, 1, 1, d.nx, d.nz};
    const Fld<const double> fr{srho, 1, 1, 1, d.nx, d.nz};
    const Fld<double> fT{sT, 1, 1, 1, d.nx, d.nz};
    const Metrics m = metrics();
    const long tu = (long)(d.nx + 1) * d.ny, tv = (long)d.nx * (d.ny + 1), tw = (long)d.nx * d.ny;
    for (int cc = 0; cc < 3; ++cc)
        for (int cb = 0; cb < 3; ++cb)
            for (int ca = 0; ca < 3; ++ca) {
                k_probe_pattern<<<flat_grid(N), kBlock, 0, stream>>>(d, ca, cb, cc, x);
                HIPCHK(hipMemsetAsync(su, 0, (nu + nv + nw) * sizeof(double), stream));
                k_correct_u<double><<<flat_grid(tu), kBlock, 0, stream>>>(d, m, x, fu, fr, use_rho);
                k_correct_v<double><<<flat_grid(tv), kBlock, 0, stream>>>(d, m, x, fv, fr, use_rho);
                k_correct_w<double><<<flat_grid(tw), kBlock, 0, stream>>>(d, m, x, fw, fr, use_rho);
                k_divergence<double><<<flat_grid(tw), kBlock, 0, stream>>>(d, m, cu, cv, cw, fr, use_rho, adv_den,
                                                                            0, 1, fT);
                k_probe_record<<<flat_grid(nw), kBlock, 0, stream>>>(d, ca, cb, cc, sT, cf_dev, leak);
                HIPCHK(hipGetLastError());
            }
    unsigned long long lk = 0;
    HIPCHK(hipMemcpyAsync(&lk, leak, sizeof(lk), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double lv;
    memcpy(&lv, &lk, sizeof(lv));
    if (max_leak) *max_leak = lv;
    return probe_finish();
}

// =================================================================== C-ABI
extern "C" {

const char *windhip_last_error(void) { return g_err.c_str(); }

void windhip_default_opts(windhip_solve_opts *o)
{
    if (!o) return;
    o->tol_rel = 1.0e-5;   // bicg_tol_rel (wind_iterative.F90:60)
    o->tol_abs = 1.0e-10;  // bicg_tol_abs (:59)
    o->max_iters = 1500;   // wind_solver_max_iters default (:57)
    o->start = WINDHIP_START_WARM;
    o->precision = WINDHIP_PLAIN_F64;
    o->inner_tol = 0.0;  // automatic: 0.3 with MIXED_MG, else 0.5
    o->inner_max_iters = 2000;
    o->max_outer = 60;
    o->iters_per_sync = 0;
    o->use_graph = 1;
    o->retry = 1;
    o->profile = 0;
}

int windhip_create(const windhip_desc *desc, windhip_solver **out)
{
    g_err.clear();
    if (!desc || !out) return fail(WINDHIP_ERR_ARG, "windhip_create: null argument");
    *out = nullptr;
    const windhip_desc &s = *desc;
    if (s.nx < 3 || s.ny < 3 || s.nz < 2) return fail(WINDHIP_ERR_ARG, "windhip_create: need nx, ny >= 3 and nz >= 2");
    if (!(s.dx > 0)) return fail(WINDHIP_ERR_ARG, "windhip_create: dx must be positive");
    if (s.ims > 1 || s.ime < s.nx || s.kms > 1 || s.kme < s.nz || s.jms > 1 || s.jme < s.ny)
        return fail(WINDHIP_ERR_ARG, "windhip_create: memory bounds must contain 1..nx, 1..nz, 1..ny");
    if (s.field_dtype != WINDHIP_F32 && s.field_dtype != WINDHIP_F64)
        return fail(WINDHIP_ERR_ARG, "windhip_create: field_dtype must be WINDHIP_F32 or WINDHIP_F64");
    if (s.field_mem != WINDHIP_HOST && s.field_mem != WINDHIP_DEVICE)
        return fail(WINDHIP_ERR_ARG, "windhip_create: field_mem must be WINDHIP_HOST or WINDHIP_DEVICE");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(WINDHIP_ERR_HIP, "windhip_create: no HIP device available");
    windhip_solver *h = new windhip_solver();
    h->desc = s;
    const char *vb = getenv("WINDHIP_VERBOSE");
    h->verbose = vb && vb[0] && vb[0] != '0';
    int rc = h->init();
    if (rc != WINDHIP_OK) {
        std::string keep = g_err;
        h->release();
        delete h;
        g_err = keep;
        return rc;
    }
    rc = scratch_address(&h->noalloc_scr);
    if (rc != WINDHIP_OK) {
        std::string keep = g_err;
        h->release();
        delete h;
        g_err = keep;
        return rc;
    }
    *out = h;
    return WINDHIP_OK;
}

int windhip_destroy(windhip_solver *s)
{
    if (!s) return WINDHIP_OK;
    s->release();
    delete s;
    return WINDHIP_OK;
}

int windhip_set_metrics(windhip_solver *s, const void *dzdx, const void *dzdy, const void *jacobian,
                        const void *jacobian_u, const void *jacobian_v, const void *jacobian_w,
                        const void *dzdx_u, const void *dzdy_v, const void *advection_dz, const void *terrain,
                        const void *mapfac_mx_u, const void *mapfac_my_u, const void *mapfac_my_v,
                        const void *mapfac_mx_v, const void *mapfac_mxy)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "null handle");
    const void *req[] = {dzdx, dzdy, jacobian, jacobian_u, jacobian_v, jacobian_w, dzdx_u, dzdy_v, advection_dz, terrain};
    for (const void *q : req)
        if (!q) return fail(WINDHIP_ERR_ARG, "windhip_set_metrics: metric arrays must not be NULL");
    const int nx = s->d.nx, ny = s->d.ny, nz = s->d.nz;
    CHK(s->gather3(dzdx, 0, 1, nx, 1, nz, 1, ny, s->m_dzdx, 0));
    CHK(s->gather3(dzdy, 0, 1, nx, 1, nz, 1, ny, s->m_dzdy, 1));
    CHK(s->gather3(jacobian, 0, 1, nx, 1, nz, 1, ny, s->m_jaco, 2));
    CHK(s->gather3(jacobian_w, 0, 1, nx, 1, nz, 1, ny, s->m_jaco_w, 3));
    CHK(s->gather3(advection_dz, 0, 1, nx, 1, nz, 1, ny, s->m_adv_dz, 4));
    HIPCHK(hipStreamSynchronize(s->stream));
    CHK(s->gather3(jacobian_u, 1, 1, nx + 1, 1, nz, 1, ny, s->m_jaco_u, 0));
    CHK(s->gather3(dzdx_u, 1, 1, nx + 1, 1, nz, 1, ny, s->m_dzdx_u, 1));
    CHK(s->gather3(jacobian_v, 2, 1, nx, 1, nz, 1, ny + 1, s->m_jaco_v, 2));
    CHK(s->gather3(dzdy_v, 2, 1, nx, 1, nz, 1, ny + 1, s->m_dzdy_v, 3));
    CHK(s->gather2(terrain, 0, 1, nx, 1, ny, s->p_terrain, 4));
    HIPCHK(hipStreamSynchronize(s->stream));
    const void *mfs[5] = {mapfac_mx_u, mapfac_my_u, mapfac_my_v, mapfac_mx_v, mapfac_mxy};
    const int kinds[5] = {1, 1, 2, 2, 0};
    for (int q = 0; q < 5; ++q) {
        s->mf_on[q] = mfs[q] != nullptr;  // a null map factor is 1, also after an earlier non-null one
        if (!mfs[q]) continue;
        if (!s->mf[q]) CHK(s->alloc(&s->mf[q], s->d.P));
        int ib = kinds[q] == 1 ? nx + 1 : nx, jb = kinds[q] == 2 ? ny + 1 : ny;
        CHK(s->gather2(mfs[q], kinds[q], 1, ib, 1, jb, s->mf[q], q));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    long pl = (long)nx * ny;
    k_module_vars<<<flat_grid(pl), kBlock, 0, s->stream>>>(s->d, s->m_adv_dz, s->m_dzdx, s->m_dzdy, s->p_terrain,
                                                          s->m_dz_if, s->p_sx, s->p_sy);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->metrics_set = true;
    s->structure_uploaded = false;
    s->probed = false;
    return WINDHIP_OK;
}

int windhip_solve(windhip_solver *s, const void *alpha, const void *div, const windhip_solve_opts *opts,
                  windhip_solve_info *info)
{
    if (!s || !alpha || !div) return fail(WINDHIP_ERR_ARG, "windhip_solve: null argument");
    windhip_solve_opts o;
    if (opts) o = *opts;
    else windhip_default_opts(&o);
    if (o.max_iters < 0 || o.inner_max_iters < 1 || o.max_outer < 1)
        return fail(WINDHIP_ERR_ARG, "windhip_solve: iteration caps must be positive");
    if (o.start == WINDHIP_START_LEARNED && !s->net)
        return fail(WINDHIP_ERR_ARG, "windhip_solve: the learned start needs a network (windhip_set_net)");
    windhip_solve_info inf{};
    if (o.iters_per_sync <= 0) {
        long per = 64000000L / std::max(1L, s->N);
        o.iters_per_sync = (int)std::max(4L, std::min(64L, per));
        // a V-cycle preconditioned inner solve mostly ends after one iteration:
        // a longer batch would end in launches that return at once
        if (o.precision == WINDHIP_MIXED_MG) o.iters_per_sync = 1;
    }
    s->profiling = o.profile != 0;
    if (s->profiling && !s->prof_flag) CHK(s->alloc(&s->prof_flag, windhip_solver::kProfFlags, false, 0));
    HIPCHK(hipEventRecord(s->ev0, s->stream));
    CHK(s->prepare_call(alpha, div));
    if (o.start == WINDHIP_START_ZERO) HIPCHK(hipMemsetAsync(s->x, 0, s->N * sizeof(double), s->stream));
    inf.start_used = o.start;
    // a start from a lambda known to be zero (a zero start, or the current
    // lambda when it was set to zeros) takes its first residual without a matvec
    const bool x_zero = o.start == WINDHIP_START_ZERO || (o.start == WINDHIP_START_LEARNED && s->x_zero_known);
    s->x_zero_known = false;  // the solve writes lambda
    const bool mixed = o.precision == WINDHIP_MIXED || o.precision == WINDHIP_MIXED_LINE ||
                       o.precision == WINDHIP_MIXED_MG;
    double *guess = nullptr, *mixed_guess = nullptr, rp = 0.0;
    if (o.start == WINDHIP_START_LEARNED) {
        CHK(s->learned_launch(&rp, &guess, mixed, o.precision == WINDHIP_MIXED_MG && s->g.mz <= 44, x_zero, o));
        inf.start_used = WINDHIP_START_WARM;
        if (mixed) {
            mixed_guess = guess;  // the start, if better, once the mixed set-up is queued
        } else {
            double rg = 0.0;
            CHK(s->learned_join(&rg));
            inf.guess_res = rg;
            inf.plain_res = rp;
            const bool used = rg < rp;
            if (used) HIPCHK(hipMemcpyAsync(s->x, guess, s->N * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
            inf.start_used = used ? WINDHIP_START_LEARNED : WINDHIP_START_WARM;
            if (s->verbose)
                fprintf(stderr, "[windhip] learned start: residual %.3e vs plain start %.3e -> %s\n", rg, rp,
                        used ? "network" : "plain");
        }
    }
    if (mixed) {
        CHK(s->run_mixed(o, &inf, mixed_guess, rp, x_zero));
        inf.nsweeps = s->nsweeps;
    } else {
        int status = 1, iters = 0;
        double res0 = 0, resf = 0;
        CHK(s->run_plain(o.max_iters, o.tol_rel, o.tol_abs, o.iters_per_sync, o.use_graph != 0, &status, &iters,
                         &res0, &resf));
        inf.nsweeps = s->nsweeps;
        int rc = WINDHIP_OK;
        // Adaptive preconditioner retry (wind_iterative.F90:474-504)
        if (o.retry && status != 0 && iters > 0 && resf > 0.01 * res0) {
            while (status != 0 && resf > 0.01 * res0 && s->nsweeps < 4) {
                s->nsweeps += 1;
                HIPCHK(hipMemsetAsync(s->x, 0, s->N * sizeof(double), s->stream));
                CHK(s->run_plain(o.max_iters, o.tol_rel, o.tol_abs, o.iters_per_sync, o.use_graph != 0, &status,
                                 &iters, &res0, &resf));
                inf.nsweeps = s->nsweeps;
            }
            s->nsweeps = 2;
            if (status != 0 && resf > res0)
                rc = fail(WINDHIP_ERR_DIVERGED, "BiCGStab diverged after the retry ladder (res0=%g, res=%g)", res0, resf);
        }
        inf.status = status;
        inf.iters = iters;
        inf.res0 = res0;
        inf.res_final = resf;
        CHK(s->true_residual(&inf.true_rel_res));
        inf.bnorm = s->st_host[0].outer_bnorm;
        if (rc != WINDHIP_OK) {
            if (info) *info = inf;
            return rc;
        }
    }
    HIPCHK(hipEventRecord(s->ev1, s->stream));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    inf.ms = ms;
    if (info) *info = inf;
    return WINDHIP_OK;
}

int windhip_apply_correction(windhip_solver *s, void *u, void *v, void *w, const void *density, int adv_den)
{
    if (!s || !u || !v || !w) return fail(WINDHIP_ERR_ARG, "windhip_apply_correction: null argument");
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_apply_correction: metrics not set");
    return s->correction(u, v, w, density, adv_den);
}

int windhip_calc_iter_winds(windhip_solver *s, const void *alpha, const void *div, void *u, void *v, void *w,
                            const void *density, int adv_den, const windhip_solve_opts *opts,
                            windhip_solve_info *info)
{
    CHK(windhip_solve(s, alpha, div, opts, info));
    return windhip_apply_correction(s, u, v, w, density, adv_den);
}

int windhip_divergence(windhip_solver *s, const void *u, const void *v, const void *w, const void *density,
                       int adv_den, int horz_only, int use_dqdt, void *div_out)
{
    if (!s || !u || !v || !w || !div_out) return fail(WINDHIP_ERR_ARG, "windhip_divergence: null argument");
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_divergence: metrics not set");
    return s->divergence(u, v, w, density, adv_den, horz_only, use_dqdt, div_out);
}

int windhip_balance_uvw(windhip_solver *s, const void *u, const void *v, void *w, const void *density, int adv_den,
                        int use_dqdt)
{
    if (!s || !u || !v || !w) return fail(WINDHIP_ERR_ARG, "windhip_balance_uvw: null argument");
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_balance_uvw: metrics not set");
    if (!s->div_scr) CHK(s->alloc(&s->div_scr, s->n_mass()));
    return s->balance_uvw(u, v, w, density, adv_den, use_dqdt, s->div_scr);
}

int windhip_balance_uvw_noalloc(windhip_solver *s, const void *u, const void *v, void *w, const void *density,
                                int adv_den, int use_dqdt)
{
    if (!s || !u || !v || !w) return fail(WINDHIP_ERR_ARG, "windhip_balance_uvw_noalloc: null argument");
    if (s->desc.field_mem != WINDHIP_DEVICE)
        return fail(WINDHIP_ERR_ARG, "windhip_balance_uvw_noalloc: fields must be in device memory "
                                     "(host fields would be staged through allocated buffers)");
    if (s->n_mass() > kNoAllocScratchDoubles)
        return fail(WINDHIP_ERR_ARG, "windhip_balance_uvw_noalloc: the mass grid has %ld elements, more than the "
                                     "%ld of the static scratch", s->n_mass(), kNoAllocScratchDoubles);
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_balance_uvw_noalloc: metrics not set");
    if (!s->noalloc_scr) return fail(WINDHIP_ERR_HIP, "windhip_balance_uvw_noalloc: no scratch address");
    return s->balance_uvw(u, v, w, density, adv_den, use_dqdt, s->noalloc_scr);
}

int windhip_calc_w_real(windhip_solver *s, const void *u, const void *v, const void *w_grid, void *w_real)
{
    if (!s || !u || !v || !w_real) return fail(WINDHIP_ERR_ARG, "windhip_calc_w_real: null argument");
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_calc_w_real: metrics not set");
    return s->w_real(u, v, w_grid, w_real);
}

int windhip_idealized_wgrid(windhip_solver *s, const void *u, const void *v, const void *w_real, void *w)
{
    if (!s || !u || !v || !w_real || !w) return fail(WINDHIP_ERR_ARG, "windhip_idealized_wgrid: null argument");
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_idealized_wgrid: metrics not set");
    return s->idealized_wgrid(u, v, w_real, w);
}

// lambda rows (0:nx+1, 0:nz+1, 0:ny+1) Fortran order <-> [k][j][i]
int windhip_get_lambda(windhip_solver *s, double *lam)
{
    if (!s || !lam) return fail(WINDHIP_ERR_ARG, "windhip_get_lambda: null argument");
    std::vector<double> tmp(s->N);
    HIPCHK(hipMemcpyAsync(tmp.data(), s->x, s->N * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int mx = s->d.mx, my = s->d.my, mz = s->d.mz;
    for (int j = 0; j < my; ++j)
        for (int k = 0; k < mz; ++k)
            for (int i = 0; i < mx; ++i)
                lam[i + (long)mx * (k + (long)mz * j)] = tmp[(long)k * s->d.P + (long)j * mx + i];
    return WINDHIP_OK;
}

int windhip_set_lambda(windhip_solver *s, const double *lam)
{
    if (!s || !lam) return fail(WINDHIP_ERR_ARG, "windhip_set_lambda: null argument");
    std::vector<double> tmp(s->N);
    const int mx = s->d.mx, my = s->d.my, mz = s->d.mz;
    unsigned long long bits = 0;  // +0.0 everywhere iff no bit is set
    for (int j = 0; j < my; ++j)
        for (int k = 0; k < mz; ++k)
            for (int i = 0; i < mx; ++i) {
                const double v = lam[i + (long)mx * (k + (long)mz * j)];
                tmp[(long)k * s->d.P + (long)j * mx + i] = v;
                bits |= __builtin_bit_cast(unsigned long long, v);
            }
    s->x_zero_known = false;
    HIPCHK(hipMemcpyAsync(s->x, tmp.data(), s->N * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->x_zero_known = bits == 0;
    return WINDHIP_OK;
}

int windhip_probe_operator(windhip_solver *s, const void *alpha, const void *density, int adv_den, double *max_leak)
{
    if (!s || !alpha) return fail(WINDHIP_ERR_ARG, "windhip_probe_operator: null argument");
    return s->probe(alpha, density, adv_den, max_leak);
}

int windhip_probe_lambda_pattern(windhip_solver *s, int ca, int cb, int cc)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "windhip_probe_lambda_pattern: null handle");
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_probe_lambda_pattern: metrics not set");
    if (ca < 0 || ca > 2 || cb < 0 || cb > 2 || cc < 0 || cc > 2)
        return fail(WINDHIP_ERR_ARG, "windhip_probe_lambda_pattern: colours must be 0, 1 or 2");
    CHK(s->probe_alloc());
    s->x_zero_known = false;
    k_probe_pattern<<<flat_grid(s->N), kBlock, 0, s->stream>>>(s->d, ca, cb, cc, s->x);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return WINDHIP_OK;
}

int windhip_probe_record(windhip_solver *s, const void *T, int ca, int cb, int cc, double *max_leak)
{
    if (!s || !T || !max_leak) return fail(WINDHIP_ERR_ARG, "windhip_probe_record: null argument");
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_probe_record: metrics not set");
    if (ca < 0 || ca > 2 || cb < 0 || cb > 2 || cc < 0 || cc > 2)
        return fail(WINDHIP_ERR_ARG, "windhip_probe_record: colours must be 0, 1 or 2");
    return s->probe_record_field(T, ca, cb, cc, max_leak);
}

int windhip_probe_finalize(windhip_solver *s)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "windhip_probe_finalize: null handle");
    if (!s->P15[0]) return fail(WINDHIP_ERR_ARG, "windhip_probe_finalize: nothing was recorded");
    return s->probe_finish();
}

int windhip_time_kernel(windhip_solver *s, int which, int reps, double *avg_ms)
{
    if (!s || !avg_ms || reps < 1) return fail(WINDHIP_ERR_ARG, "windhip_time_kernel: bad argument");
    if (!s->structure_uploaded) return fail(WINDHIP_ERR_ARG, "windhip_time_kernel: solve once first");
    auto launch = [&]() {
        if (which == 0)
            s->with_op<double>([&](auto op) {
                k_spmv<decltype(op)><<<s->g.nt, kThreads, 0, s->stream>>>(op, s->g, s->ph[1], s->v);
                return 0;
            });
        else
            s->with_op<float>([&](auto op) {
                k_spmv<decltype(op)><<<s->g.nt, kThreads, 0, s->stream>>>(op, s->g, s->ph32[1], s->v32);
                return 0;
            });
    };
    if (which == 1 && !s->x32) CHK(s->alloc_f32_krylov());
    launch();
    HIPCHK(hipEventRecord(s->ev0, s->stream));
    for (int q = 0; q < reps; ++q) launch();
    HIPCHK(hipEventRecord(s->ev1, s->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *avg_ms = ms / reps;
    return WINDHIP_OK;
}

int windhip_set_alpha(windhip_solver *s, const void *alpha)
{
    if (!s || !alpha) return fail(WINDHIP_ERR_ARG, "windhip_set_alpha: null argument");
    return s->set_alpha(alpha);
}

static void rows_to_dev_order(const windhip_solver *s, const double *f, double *dv)
{
    const int mx = s->d.mx, my = s->d.my, mz = s->d.mz;
    for (int j = 0; j < my; ++j)
        for (int k = 0; k < mz; ++k)
            for (int i = 0; i < mx; ++i) dv[(long)k * s->d.P + (long)j * mx + i] = f[i + (long)mx * (k + (long)mz * j)];
}

static void rows_from_dev_order(const windhip_solver *s, const double *dv, double *f)
{
    const int mx = s->d.mx, my = s->d.my, mz = s->d.mz;
    for (int j = 0; j < my; ++j)
        for (int k = 0; k < mz; ++k)
            for (int i = 0; i < mx; ++i) f[i + (long)mx * (k + (long)mz * j)] = dv[(long)k * s->d.P + (long)j * mx + i];
}

int windhip_apply_operator(windhip_solver *s, int precision, const double *x_rows, double *y_rows)
{
    if (!s || !x_rows || !y_rows) return fail(WINDHIP_ERR_ARG, "windhip_apply_operator: null argument");
    if (!s->structure_uploaded) return fail(WINDHIP_ERR_ARG, "windhip_apply_operator: no operator yet (set_alpha or solve first)");
    std::vector<double> tmp(s->N);
    rows_to_dev_order(s, x_rows, tmp.data());
    double *dx_ = nullptr, *dy_ = nullptr;
    HIPCHK(hipMalloc(&dx_, s->N * sizeof(double)));
    HIPCHK(hipMalloc(&dy_, s->N * sizeof(double)));
    HIPCHK(hipMemcpyAsync(dx_, tmp.data(), s->N * sizeof(double), hipMemcpyHostToDevice, s->stream));
    if (precision == WINDHIP_MIXED) {
        float *fx = nullptr, *fy = nullptr;
        HIPCHK(hipMalloc(&fx, s->N * sizeof(float)));
        HIPCHK(hipMalloc(&fy, s->N * sizeof(float)));
        k_round_f32<<<flat_grid(s->N), kBlock, 0, s->stream>>>(s->N, dx_, fx);
        s->with_op<float>([&](auto op) {
            k_spmv<decltype(op)><<<s->g.nt, kThreads, 0, s->stream>>>(op, s->g, fx, fy);
            return 0;
        });
        k_widen_f64<<<flat_grid(s->N), kBlock, 0, s->stream>>>(s->N, fy, dy_);
        HIPCHK(hipStreamSynchronize(s->stream));
        (void)hipFree(fx);
        (void)hipFree(fy);
    } else {
        s->with_op<double>([&](auto op) {
            k_spmv<decltype(op)><<<s->g.nt, kThreads, 0, s->stream>>>(op, s->g, dx_, dy_);
            return 0;
        });
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(tmp.data(), dy_, s->N * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    (void)hipFree(dx_);
    (void)hipFree(dy_);
    rows_from_dev_order(s, tmp.data(), y_rows);
    return WINDHIP_OK;
}

int windhip_get_rows(windhip_solver *s, int which, double *out)
{
    if (!s || !out) return fail(WINDHIP_ERR_ARG, "windhip_get_rows: null argument");
    const double *src[] = {s->b, s->dinv, s->B, s->C, s->H, s->I, s->L, s->M, s->x};
    if (which < 0 || which > 8) return fail(WINDHIP_ERR_ARG, "windhip_get_rows: which must be 0..8");
    std::vector<double> tmp(s->N);
    HIPCHK(hipMemcpyAsync(tmp.data(), src[which], s->N * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    rows_from_dev_order(s, tmp.data(), out);
    return WINDHIP_OK;
}

int windhip_set_net(windhip_solver *s, windhip_net *net)
{
    if (!s) return WINDHIP_ERR_ARG;
    s->net = net;
    return WINDHIP_OK;
}

int windhip_set_guess_smoothing(windhip_solver *s, int steps)
{
    if (!s || steps < 0 || steps > 100) return fail(WINDHIP_ERR_ARG, "windhip_set_guess_smoothing: 0..100 steps");
    s->net_smooth = steps;
    return WINDHIP_OK;
}

int windhip_set_multigrid(windhip_solver *s, int pre, int post, int coarse_pre, int coarse_post, int depth,
                          double omega)
{
    auto ok = [](int n) { return n >= 1 && n <= 8; };
    if (!s || !ok(pre) || !ok(post) || !ok(coarse_pre) || coarse_post < 0 || coarse_post > 8 || depth < 0 ||
        !(omega > 0 && omega <= 1.5))
        return fail(WINDHIP_ERR_ARG, "windhip_set_multigrid: smoothing steps in 1..8 (coarse_post 0..8), depth >= 0, "
                                     "omega in (0, 1.5]");
    s->mg_pre = pre;
    s->mg_post = post;
    s->mg_cpre = coarse_pre;
    s->mg_cpost = coarse_post;
    s->mg_depth = depth;
    s->mg_omega = (float)omega;
    for (auto &kv : s->graphs) (void)hipGraphExecDestroy(kv.second);  // the cached V-cycles change shape
    s->graphs.clear();
    return WINDHIP_OK;
}

int windhip_set_mg_fusion(windhip_solver *s, int fused_transfers, int onepass)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "windhip_set_mg_fusion: null handle");
    if (fused_transfers < 0 || fused_transfers > 2)
        return fail(WINDHIP_ERR_ARG, "windhip_set_mg_fusion: fused_transfers %d outside 0..2", fused_transfers);
    s->mg_fused = fused_transfers != 0;
    s->mg_tail = fused_transfers == 2;
    s->mg_onepass = onepass != 0;
    for (auto &kv : s->graphs) (void)hipGraphExecDestroy(kv.second);  // the cached V-cycles change shape
    s->graphs.clear();
    return WINDHIP_OK;
}

static void mg_rows_from_cm(const MgGeo &g, const float *src, float *dst);

int windhip_set_mg_galerkin(windhip_solver *s, int first_level)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "windhip_set_mg_galerkin: null handle");
    if (first_level != 0 && first_level < 2)
        return fail(WINDHIP_ERR_ARG, "windhip_set_mg_galerkin: first level %d (0: none, >= 2; level 1 stays "
                                     "rediscretised)", first_level);
    const int from = first_level == 0 ? 0 : first_level - 1;
    if (!s->mg.empty() && from != s->mg_gal_from)
        return fail(WINDHIP_ERR_ARG, "windhip_set_mg_galerkin: the hierarchy is built; set it before the first "
                                     "MIXED_MG solve");
    s->mg_gal_from = from;
    return WINDHIP_OK;
}

int windhip_mg_stencil(windhip_solver *s, int level, float *q27_host)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "windhip_mg_stencil: null handle");
    if (level < 1 || level > (int)s->mg.size())
        return fail(WINDHIP_ERR_ARG, "windhip_mg_stencil: level %d outside 1..%d (run a MIXED_MG solve first)", level,
                    (int)s->mg.size());
    const auto &L = s->mg[level - 1];
    if (!L.gal) return fail(WINDHIP_ERR_ARG, "windhip_mg_stencil: level %d is rediscretised", level);
    if (!q27_host) return WINDHIP_OK;
    HIPCHK(hipStreamSynchronize(s->stream));
    std::vector<float> t((size_t)L.g.N);
    for (int o = 0; o < 27; ++o) {
        HIPCHK(hipMemcpy(t.data(), L.G + (long)o * L.g.N, sizeof(float) * L.g.N, hipMemcpyDeviceToHost));
        mg_rows_from_cm(L.g, t.data(), q27_host + (long)o * L.g.N);
    }
    return WINDHIP_OK;
}

int windhip_mg_level(windhip_solver *s, int level, int *dims_out, float *coef_host)
{
    if (!s || !dims_out) return fail(WINDHIP_ERR_ARG, "windhip_mg_level: null argument");
    if (level < 1 || level > (int)s->mg.size())
        return fail(WINDHIP_ERR_ARG, "windhip_mg_level: level %d outside 1..%d (run a MIXED_MG solve first)", level,
                    (int)s->mg.size());
    const auto &L = s->mg[level - 1];
    dims_out[0] = L.g.nx; dims_out[1] = L.g.ny; dims_out[2] = L.g.mz;
    if (!coef_host) return WINDHIP_OK;
    if (L.gal)
        return fail(WINDHIP_ERR_ARG, "windhip_mg_level: level %d holds a Galerkin operator (windhip_mg_stencil)",
                    level);
    HIPCHK(hipStreamSynchronize(s->stream));
    const float *arr[6] = {L.B32, L.C32, L.H32, L.I32, L.L32, L.M32};
    for (int q = 0; q < 6; ++q)
        HIPCHK(hipMemcpy(coef_host + (long)q * L.g.N, arr[q], sizeof(float) * L.g.N, hipMemcpyDeviceToHost));
    return WINDHIP_OK;
}

// Host rows of a level between the plane-major layout of the C-ABI (rows
// [k][j][i]) and the column-major layout of the wavefront levels (row c mz + k).
static std::vector<float> mg_rows_to_cm(const MgGeo &g, const float *src)
{
    std::vector<float> d((size_t)g.N);
    for (long c = 0; c < g.P; ++c)
        for (int k = 0; k < g.mz; ++k) d[(size_t)(c * g.mz + k)] = src[(long)k * g.P + c];
    return d;
}
static void mg_rows_from_cm(const MgGeo &g, const float *src, float *dst)
{
    for (long c = 0; c < g.P; ++c)
        for (int k = 0; k < g.mz; ++k) dst[(long)k * g.P + c] = src[c * g.mz + k];
}

int windhip_mg_smooth(windhip_solver *s, int level, int fused, const float *b, const float *y, float *out)
{
    if (!s || !b || !out) return fail(WINDHIP_ERR_ARG, "windhip_mg_smooth: null argument");
    if (level < 1 || level > (int)s->mg.size())
        return fail(WINDHIP_ERR_ARG, "windhip_mg_smooth: level %d outside 1..%d (run a MIXED_MG solve first)", level,
                    (int)s->mg.size());
    if (fused < 0 || fused > 4) return fail(WINDHIP_ERR_ARG, "windhip_mg_smooth: form %d outside 0..4", fused);
    auto &L = s->mg[level - 1];
    if (L.gal && fused != 1)
        return fail(WINDHIP_ERR_ARG, "windhip_mg_smooth: level %d holds a Galerkin operator: form 1 only", level);
    const size_t bytes = sizeof(float) * L.g.N;
    // the one-launch form works on the level's own layout (column-major on
    // the wavefront levels), the others on the plane-major rows
    const bool cm = fused == 1 && L.cm;
    std::vector<float> tb, ty, to;
    if (cm) {
        tb = mg_rows_to_cm(L.g, b);
        if (y) ty = mg_rows_to_cm(L.g, y);
        to.resize(L.g.N);
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(L.b, cm ? tb.data() : b, bytes, hipMemcpyHostToDevice));
    if (y) HIPCHK(hipMemcpy(L.x, cm ? ty.data() : y, bytes, hipMemcpyHostToDevice));
    windhip::k_state_reset<<<1, 1, 0, s->stream>>>(s->st + 1, 1, 0.0, 0.0, 0.0);  // a live state: done = 0
    HIPCHK(hipGetLastError());
    s->mg_smooth(level - 1, fused, y ? L.x : nullptr, L.t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(cm ? to.data() : out, L.t, bytes, hipMemcpyDeviceToHost));
    if (cm) mg_rows_from_cm(L.g, to.data(), out);
    return WINDHIP_OK;
}

int windhip_mg_cycle(windhip_solver *s, int level, int fused, const float *b, float *out)
{
    if (!s || !b || !out) return fail(WINDHIP_ERR_ARG, "windhip_mg_cycle: null argument");
    if (fused < 0 || fused > 2) return fail(WINDHIP_ERR_ARG, "windhip_mg_cycle: fused %d outside 0..2", fused);
    if (level < 1 || level > (int)s->mg.size())
        return fail(WINDHIP_ERR_ARG, "windhip_mg_cycle: level %d outside 1..%zu (run a MIXED_MG solve first)", level,
                    s->mg.size());
    windhip_solver::MgLevel &L = s->mg[(size_t)level - 1];
    const long n = L.g.N;
    // the level's vectors are column-major on the wavefront levels
    std::vector<float> tb, to;
    if (L.cm) {
        tb = mg_rows_to_cm(L.g, b);
        to.resize(n);
    }
    // the V-cycle's kernels return at once when the inner state says done
    k_state_reset<<<1, 1, 0, s->stream>>>(s->st + 1, 1, 0.0, 0.0, 0.5);
    HIPCHK(hipMemcpyAsync(L.b, L.cm ? tb.data() : b, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
    const bool keep = s->mg_fused, keep_tail = s->mg_tail, prof = s->profiling;
    s->mg_fused = fused != 0;
    s->mg_tail = fused == 2;
    s->profiling = false;
    size_t l0 = (size_t)level - 1;  // the persistent launch starts at the first column-major level
    while (l0 < s->mg.size() && !s->mg[l0].cm) ++l0;
    int rc = s->prepare_tail(fused == 2 ? l0 : SIZE_MAX);
    float *e = rc == WINDHIP_OK ? s->mg_coarse((size_t)level - 1) : nullptr;
    const bool tail = s->tail_l0 != SIZE_MAX;
    s->tail_l0 = SIZE_MAX;  // the next solve plans its own
    s->mg_fused = keep;
    s->mg_tail = keep_tail;
    s->profiling = prof;
    CHK(rc);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(L.cm ? to.data() : out, e, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (tail) CHK(s->check_tail());
    if (L.cm) mg_rows_from_cm(L.g, to.data(), out);
    return WINDHIP_OK;
}

int windhip_learned_input(windhip_solver *s, float *in_dev, int *dims_out, double *scale)
{
    if (!s || !dims_out) return WINDHIP_ERR_ARG;
    int Dp, Hp, Wp;
    s->net_dims(&Dp, &Hp, &Wp);
    dims_out[0] = Dp; dims_out[1] = Hp; dims_out[2] = Wp;
    if (!in_dev) return WINDHIP_OK;
    if (!s->metrics_set) return fail(WINDHIP_ERR_ARG, "windhip_learned_input: metrics not set");
    double sc = 0.0;
    CHK(s->net_input(in_dev, &sc));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (scale) *scale = sc;
    return WINDHIP_OK;
}

int windhip_profile(windhip_solver *s, int reset, double *calls, double *ms)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "null handle");
    for (int q = 0; q < WINDHIP_PROFILE_KINDS; ++q) {
        if (calls) calls[q] = s->prof_calls[q];
        if (ms) ms[q] = s->prof_ms[q];
        if (reset) s->prof_calls[q] = s->prof_ms[q] = 0.0;
    }
    return WINDHIP_OK;
}

int windhip_memory(windhip_solver *s, double *bytes_handle, double *bytes_device_used)
{
    if (!s) return fail(WINDHIP_ERR_ARG, "null handle");
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    if (bytes_handle) *bytes_handle = (double)s->bytes;
    if (bytes_device_used) *bytes_device_used = (double)(tot - fr);
    return WINDHIP_OK;
}

}  // extern "C"
