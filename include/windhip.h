/*
 * windhip.h -- C-ABI of the MI355X mass-conserving wind solver.
 *
 * Drop-in for the Krylov solve of HICAR's `wind_iterative` module
 * (src/physics/wind_iterative.F90). Fortran host code binds it through
 * ISO_C_BINDING (hicar_amd/fortran/windhip_interface.F90); Python through
 * ctypes (hicar_amd/_native.py). Plain pointers and sizes only.
 *
 * Conventions
 *  - Every entry point returns int: WINDHIP_OK (0) or a nonzero code; the text
 *    of the last error is in windhip_last_error(). (The reference prints and
 *    `stop`s instead, wind_iterative.F90:506-528.)
 *  - One opaque handle per nest (the reference keeps one module state per nest,
 *    wind_iterative.F90:134-181). A handle owns its HIP stream and device state
 *    and is not thread-safe; different handles are independent.
 *  - Field arrays (metrics, alpha, div, u, v, w, density) are the caller's:
 *    Fortran column-major (i,k,j) with the memory bounds of windhip_desc,
 *    element type desc.field_dtype (HICAR's `real` is WINDHIP_F32), in host or
 *    device memory per desc.field_mem. Mass-grid arrays span
 *    (ims:ime, kms:kme, jms:jme); u-grid arrays (ims:ime+1, ...); v-grid arrays
 *    (..., jms:jme+1); 2-D arrays (ims:ime, jms:jme), u-grid 2-D (ims:ime+1,
 *    jms:jme), v-grid 2-D (ims:ime, jms:jme+1).
 *  - Solver rows (lambda) are (0:nx+1, 0:nz+1, 0:ny+1), Fortran order, FP64,
 *    as x_sol of wind_iterative.F90:359-370.
 *  - One GPU: the domain is a single tile, ids = its = 1, ide = ite = nx, etc.
 */
#ifndef WINDHIP_H
#define WINDHIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct windhip_solver windhip_solver;

enum {
    WINDHIP_OK = 0,
    WINDHIP_ERR_ARG = 1,      /* invalid argument or call order */
    WINDHIP_ERR_HIP = 2,      /* HIP runtime error (message names the call) */
    WINDHIP_ERR_NAN = 3,      /* NaN in the input divergence (wind_iterative.F90:319-349) */
    WINDHIP_ERR_DIVERGED = 4, /* still diverged after the retry ladder (:506-528) */
    WINDHIP_ERR_NOMEM = 5
};

enum { WINDHIP_F32 = 4, WINDHIP_F64 = 8 };     /* field element type */
enum { WINDHIP_HOST = 0, WINDHIP_DEVICE = 1 }; /* field memory */

/* Start value of lambda for a solve. WARM is the reference behaviour: x_sol
 * persists between calls (wind_iterative.F90:435-442). */
enum { WINDHIP_START_WARM = 0, WINDHIP_START_ZERO = 1, WINDHIP_START_GIVEN = 2,
       WINDHIP_START_LEARNED = 3 };

/* PLAIN_F64: the reference's BiCGStab in FP64 throughout. MIXED: FP32 inner
 * BiCGStab (matvec, vectors, coefficients; the reference's iteration with its
 * 2 Jacobi-Richardson sweeps) as the preconditioner of an FP64 flexible GMRES
 * refinement; the true residual, its norm, the rho/omega dot products and the
 * lambda accumulation stay FP64. MIXED_LINE: the same with the FP32 BiCGStab
 * preconditioned by two vertical line-relaxation steps instead; MIXED_MG:
 * preconditioned by one multigrid V-cycle (horizontal 2 x 2 semi-coarsening,
 * the first coarse operator rediscretised from the averaged metrics at its
 * spacing, the deeper ones Galerkin products R A P (windhip_set_mg_galerkin),
 * damped vertical line smoothing; windhip_set_multigrid). Neither has a
 * counterpart in the reference. */
enum { WINDHIP_PLAIN_F64 = 0, WINDHIP_MIXED = 1, WINDHIP_MIXED_LINE = 2, WINDHIP_MIXED_MG = 3 };

typedef struct windhip_desc {
    int nx, ny, nz;              /* physical cells (ide, jde, kde) */
    double dx;                   /* domain%dx */
    int ims, ime, kms, kme, jms, jme; /* memory bounds of mass-grid field arrays */
    int field_dtype;             /* WINDHIP_F32 or WINDHIP_F64 */
    int field_mem;               /* WINDHIP_HOST or WINDHIP_DEVICE */
} windhip_desc;

typedef struct windhip_solve_opts {
    double tol_rel;        /* bicg_tol_rel (reference 1e-5, wind_iterative.F90:60) */
    double tol_abs;        /* bicg_tol_abs (1e-10, :59) */
    int max_iters;         /* wind_solver_iterations (1500, :57); MIXED: cap on all FP32 iterations */
    int start;             /* WINDHIP_START_* */
    int precision;         /* WINDHIP_PLAIN_F64, WINDHIP_MIXED, WINDHIP_MIXED_LINE or WINDHIP_MIXED_MG */
    double inner_tol;      /* mixed modes: relative reduction asked of each FP32 inner solve
                              (0 = automatic: 0.3 with MIXED_MG, else 0.5) */
    int inner_max_iters;   /* MIXED: cap per inner solve */
    int max_outer;         /* MIXED: cap on FGMRES steps (= inner solves); cycles restart every 32 */
    int iters_per_sync;    /* iterations launched between host convergence checks; 0 = auto */
    int use_graph;         /* replay the iteration batch as a captured hipGraph */
    int retry;             /* apply the reference's precond retry ladder (:474-504) */
    int profile;           /* time every hot kernel launch with hipEvents (disables graphs) */
} windhip_solve_opts;

typedef struct windhip_solve_info {
    int status;            /* 0 converged, 1 iteration cap, 2 breakdown (bicgstab_solve) */
    int iters;             /* BiCGStab iterations (MIXED: total inner iterations) */
    int nsweeps;           /* preconditioner sweeps of the last solve */
    int outer_iters;       /* MIXED: FGMRES steps (FP32 inner solves) */
    double res0;           /* ||r0|| (recurrence norm of the reference, absolute) */
    double res_final;      /* final recurrence residual norm (MIXED: true residual) */
    double bnorm;          /* ||b|| */
    double true_rel_res;   /* ||b - A lambda|| / ||b||, FP64, after the solve */
    double ms;             /* device time of the solve (hipEvents on the handle's stream) */
    int start_used;        /* WINDHIP_START_* actually used (the learned start falls back) */
    int reserved;
    double guess_res;      /* LEARNED: ||b - A lambda_net|| of the network's guess */
    double plain_res;      /* LEARNED: ||b - A lambda|| of the plain start it was compared with */
} windhip_solve_info;

/* Library / handle life cycle. create ~ init_iter_winds (wind_iterative.F90:215-291):
 * validates bounds, allocates device state, creates the stream. */
void windhip_default_opts(windhip_solve_opts *opts);
int windhip_create(const windhip_desc *desc, windhip_solver **out);
/* finalize_iter_winds (wind_iterative.F90:2142-2267) */
int windhip_destroy(windhip_solver *s);
const char *windhip_last_error(void);

/* init_module_vars (wind_iterative.F90:2274-2421): terrain metrics of the tile.
 * terrain is neighbor_terrain (2-D). Map factors may be NULL (1.0):
 * mapfac_mx_u, mapfac_my_u on the u grid; mapfac_my_v, mapfac_mx_v on the v
 * grid; mapfac_mxy on the mass grid (calc_divergence, wind.F90:288-290, and
 * calc_updated_winds, wind_iterative.F90:1797-1798). */
int windhip_set_metrics(windhip_solver *s, const void *dzdx, const void *dzdy,
                        const void *jacobian, const void *jacobian_u, const void *jacobian_v,
                        const void *jacobian_w, const void *dzdx_u, const void *dzdy_v,
                        const void *advection_dz, const void *terrain,
                        const void *mapfac_mx_u, const void *mapfac_my_u,
                        const void *mapfac_my_v, const void *mapfac_mx_v,
                        const void *mapfac_mxy);

/* The solve of calc_iter_winds (wind_iterative.F90:295-528): alpha/div copy,
 * NaN scan, first-call assembly or alpha-range refresh, diagonal, rhs,
 * BiCGStab (bicgstab_solve :555-791), retry ladder. Lambda stays on the device
 * for windhip_apply_correction and as the next warm start. opts may be NULL. */
int windhip_solve(windhip_solver *s, const void *alpha, const void *div,
                  const windhip_solve_opts *opts, windhip_solve_info *info);

/* calc_updated_winds (wind_iterative.F90:1742-2071): u, v (dqdt) and the
 * grid-relative w (dqdt) corrected in place from grad(lambda). density may be
 * NULL (1.0); it is read only when adv_den is nonzero. */
int windhip_apply_correction(windhip_solver *s, void *u, void *v, void *w,
                             const void *density, int adv_den);

/* calc_iter_winds(domain, alpha, div, adv_den) (wind_iterative.F90:295): solve
 * followed by the correction. opts and info may be NULL. */
int windhip_calc_iter_winds(windhip_solver *s, const void *alpha, const void *div,
                            void *u, void *v, void *w, const void *density, int adv_den,
                            const windhip_solve_opts *opts, windhip_solve_info *info);

/* calc_divergence (wind.F90:250-523), the operator D: div_out on the mass grid.
 * density may be NULL (1.0). */
int windhip_divergence(windhip_solver *s, const void *u, const void *v, const void *w,
                       const void *density, int adv_den, int horz_only, int use_dqdt,
                       void *div_out);

/* The caller's steps around the solve in update_winds (wind.F90:790-860), on
 * the same fields and metrics (every field in the caller's layout, dtype and
 * memory as for windhip_calc_iter_winds; domain cells only):
 * windhip_balance_uvw = balance_uvw (wind.F90:123-191): w from the horizontal
 *   divergence of (u, v) by calc_w's upward integration (:193-234); w is
 *   written. The caller passes the data or the dqdt arrays, as update_in
 *   selects in the reference (use_dqdt mirrors that flag; the divergence is
 *   the density-weighted form of calc_divergence either way, :159/:164).
 *   adv_den selects calc_w's form; density may be null (rho = 1).
 * windhip_calc_w_real = calc_w_real (:1175-1227): the real vertical velocity
 *   on the mass grid from u, v and the grid-relative w_grid (null: zero).
 * windhip_idealized_wgrid = calc_idealized_wgrid (:940-998): the grid-relative
 *   w that reproduces a prescribed w_real for the given u, v. */
int windhip_balance_uvw(windhip_solver *s, const void *u, const void *v, void *w, const void *density, int adv_den,
                        int use_dqdt);
/* No counterpart in the reference: windhip_balance_uvw for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results; it calls no
 * hipMalloc/hipFree: its FP64 divergence scratch is a static device array of
 * 16,777,216 doubles (128 MiB per GPU, held until the process ends), so the
 * caller's mass grid ((ime-ims+1) (kme-kms+1) (jme-jms+1) elements) may hold
 * at most that many elements (else WINDHIP_ERR_ARG), and the handle must
 * have been created with field_mem = WINDHIP_DEVICE (host fields would be
 * staged through allocated buffers; else WINDHIP_ERR_ARG). Every call on one
 * GPU uses the same scratch: calls that overlap on a GPU race. */
int windhip_balance_uvw_noalloc(windhip_solver *s, const void *u, const void *v, void *w, const void *density,
                                int adv_den, int use_dqdt);
int windhip_calc_w_real(windhip_solver *s, const void *u, const void *v, const void *w_grid, void *w_real);
int windhip_idealized_wgrid(windhip_solver *s, const void *u, const void *v, const void *w_real, void *w);

/* Warm start / learned guess: lambda over the solver rows, host FP64. */
int windhip_get_lambda(windhip_solver *s, double *lambda_rows);
int windhip_set_lambda(windhip_solver *s, const double *lambda_rows);

/* Operator probing (wind_iterative.F90:1078-1248, wind.F90:1010-1099): build
 * A = 2 D(G(.)) from 27 lattice colourings on the device and switch the solver
 * to the probed 15-array operator. alpha is the wind_alpha field in use. */
int windhip_probe_operator(windhip_solver *s, const void *alpha, const void *density,
                           int adv_den, double *max_leak);

/* The reference's probing hooks one by one, for a host that drives the
 * 27-colour loop of calibrate_projection_operator (wind.F90:1010-1099)
 * itself with its own divergence operator:
 *   windhip_probe_lambda_pattern = probe_lambda_pattern(ca, cb, cc)
 *     (wind_iterative.F90:1096-1116): lambda = 1 on the physical cells with
 *     (i mod 3, k mod 3, j mod 3) = (ca, cb, cc), 0 elsewhere;
 *   probe_zero_corrections (:1118-1153) is the caller zeroing its u, v, w;
 *   probe_apply_corrections (:1155-1161) is windhip_apply_correction;
 *   windhip_probe_record = probe_record(domain, Tfield, ca, cb, cc, max_leak)
 *     (:1163-1230): 2 T of the caller's mass-grid field T (its calc_divergence
 *     of the corrected winds) into the 15 coefficients the colouring reaches;
 *     *max_leak (in/out) is raised to the largest |2 T| off the stencil;
 *   windhip_probe_finalize = probe_finalize (:1232-1248): the solver uses the
 *     recorded operator from now on and lambda is reset to 0.
 * Colours are 0..2. */
int windhip_probe_lambda_pattern(windhip_solver *s, int ca, int cb, int cc);
int windhip_probe_record(windhip_solver *s, const void *T, int ca, int cb, int cc, double *max_leak);
int windhip_probe_finalize(windhip_solver *s);

/* Per-kernel timing of the solve path: average device time (ms) of `reps`
 * back-to-back launches of one kernel on the handle's stream, bracketed by
 * hipEvents. which: 0 = FP64 matvec, 1 = FP32 matvec. */
int windhip_time_kernel(windhip_solver *s, int which, int reps, double *avg_ms);

/* ---- Diagnostics (tests, benchmarks) ----
 * set_alpha: the alpha copy, first-call assembly or alpha-range refresh, BC
 * planes and diagonal of calc_iter_winds (wind_iterative.F90:309-430) without
 * the solve. apply_operator: y = A x with the current operator (spmv,
 * :813-1027), precision WINDHIP_PLAIN_F64 or WINDHIP_MIXED (the FP32 operator);
 * rows in Fortran order as for windhip_get_lambda. get_rows: which = 0 rhs b,
 * 1 D^-1, 2..7 the stored coefficients B, C, H, I, L, M, 8 lambda. */
int windhip_set_alpha(windhip_solver *s, const void *alpha);
int windhip_apply_operator(windhip_solver *s, int precision, const double *x_rows, double *y_rows);
int windhip_get_rows(windhip_solver *s, int which, double *out_rows);

/* ------------------------------------------------------------------ learned guess
 * No counterpart in the reference: the learned initial guess for lambda. A
 * 3-D convolutional encoder-decoder (4 scales, 16/32/64/64 channels, 3x3x3
 * kernels, bf16 weights, FP32 accumulation on MFMA) maps the rhs, dz/dx, dz/dy,
 * the Jacobian, alpha and the height above ground to lambda / (rms(b) dx^2).
 * Volumes are channels-last FP32 device arrays [B][D][H][W][8] (channels 6, 7
 * zero), D, H, W multiples of 8. Parameters are one flat FP32 array, per
 * convolution W[co][27][ci] (tap = 9 (dz+1) + 3 (dy+1) + (dx+1)) then bias[co],
 * in the order c0a c0b c1a c1b c2a c2b c3a c3b d2a d2b d1a d1b d0a d0b, then
 * the 1x1x1 head (16 weights, 1 bias). */
typedef struct windhip_net windhip_net;
long windhip_net_param_count(void);
int windhip_net_create(windhip_net **out);
int windhip_net_destroy(windhip_net *net);
const char *windhip_net_last_error(windhip_net *net);
int windhip_net_set_params(windhip_net *net, const float *params_host);
int windhip_net_get_params(windhip_net *net, float *params_host);
/* out[B][D][H][W] = network(in) */
int windhip_net_forward(windhip_net *net, const float *in_dev, int B, int D, int H, int W, float *out_dev);
/* Mean squared error against target [B][D][H][W] and the gradient of every
 * parameter (copied to grad_host when non-null) */
int windhip_net_loss_grad(windhip_net *net, const float *in_dev, const float *target_dev, int B, int D, int H,
                          int W, double *loss, float *grad_host);
/* One Adam step (Kingma & Ba) with the gradient of the last loss_grad */
int windhip_net_adam_step(windhip_net *net, float lr, float beta1, float beta2, float eps);
/* Mean device time of the forward pass over a B x D x H x W volume */
int windhip_net_time_forward(windhip_net *net, int B, int D, int H, int W, int reps, double *ms);
/* Attach a network to a solver: opts.start = WINDHIP_START_LEARNED then
 * starts from the network's guess, after `steps` (default: 10 in MIXED_MG,
 * whose steps are one-pass kernels on bf16 rows, else 4) vertical
 * line-relaxation steps lambda += M^-1 (b - A lambda) with M the tridiagonal
 * column part of A, when its FP64 residual is below the plain start's (the
 * current lambda), and from the plain start otherwise. A network whose
 * parameters are all zero gives the relaxation of the zero start alone. */
int windhip_set_net(windhip_solver *s, windhip_net *net);
int windhip_set_guess_smoothing(windhip_solver *s, int steps);
/* MIXED_MG V-cycle: `pre` and `post` damped line-smoothing steps on the fine
 * level and `coarse_pre` (1..8), `coarse_post` (0..8) on the coarse levels
 * (defaults 2 on the fine level, 1 on the coarse ones) around each coarse correction; `depth` coarse levels used
 * (0, the default: down to the deepest level with at least 4 interior columns
 * each way; a level with one interior column is solved exactly, any other
 * last level is smoothed coarse_pre + coarse_post times); damping omega
 * (default 0.9). No counterpart in the reference. */
int windhip_set_multigrid(windhip_solver *s, int pre, int post, int coarse_pre, int coarse_post, int depth,
                          double omega);
/* MIXED_MG kernel forms (default fused_transfers = 1, onepass = 1):
 * fused_transfers 1, the one-launch descent (residual, restriction, first
 * coarse step) and return (prolongation, first post-step) between the levels
 * of one wavefront per column; 2, as 1, and the whole cycle from the first
 * such level down is one persistent launch whose steps are separated by
 * grid-wide barriers (bitwise the same values; slower at 512 x 512 x 40);
 * 0, separate kernels.
 * onepass != 0, smoothing steps whose residual feeds the line solve in
 * registers (one kernel instead of two) on the fine level and the larger
 * coarse levels of up to 44 rows per column, and the fine level's residual
 * and restriction in one kernel. Off, the separate kernels run; the one-pass
 * kernels compute the same values as their pairs. fused_transfers outside
 * 0..2: WINDHIP_ERR_ARG. No counterpart in the reference. */
int windhip_set_mg_fusion(windhip_solver *s, int fused_transfers, int onepass);
/* Coarse operators of the MIXED_MG hierarchy: levels from `first_level` on
 * (numbered as in windhip_mg_level) take the Galerkin product R A P of the
 * next finer level's operator with the cycle's transfers (R the mean over a
 * coarse column's 2 x 2 fine columns, P the bilinear prolongation): 27
 * coefficients per row, smoothed a wavefront per column. Levels before it
 * rediscretise the analytic operator at their dx from averaged metrics.
 * first_level 0: every level rediscretised; otherwise >= 2 (default 2: the
 * first coarse level keeps its cheap 15-point rows). Set before the first
 * MIXED_MG solve (a different value afterwards: WINDHIP_ERR_ARG). No
 * counterpart in the reference. */
int windhip_set_mg_galerkin(windhip_solver *s, int first_level);
/* Coarse level `level` (1 = first coarsening) of the MIXED_MG hierarchy built
 * by the last MIXED_MG solve: dims_out = (nx, ny, mz) of the level (interior
 * columns, rows per column); coef_host (may be null) receives its 6 x N FP32
 * coefficients B, C, H, I, L, M (the analytic operator rediscretised at the
 * level's dx from its averaged metrics), rows [k][j][i] of (nx+2) x (ny+2)
 * columns; on a Galerkin level coef_host must be null (WINDHIP_ERR_ARG
 * otherwise). Inspection only; no counterpart in the reference. */
int windhip_mg_level(windhip_solver *s, int level, int *dims_out, float *coef_host);
/* The 27 x N FP32 coefficients of Galerkin level `level` (numbered as in
 * windhip_mg_level): coefficient o = 9 (dk + 1) + 3 (dj + 1) + (di + 1), the
 * entry of row (i, k, j) for column (i + di, j + dj) at level k + dk, at
 * q27_host[o N + n], rows n in the layout of windhip_mg_level. q27_host null:
 * returns WINDHIP_OK if the level is a Galerkin level. A rediscretised level:
 * WINDHIP_ERR_ARG. Inspection only; no counterpart in the reference. */
int windhip_mg_stencil(windhip_solver *s, int level, float *q27_host);
/* One damped line-smoothing step of the V-cycle on coarse level `level`, out =
 * y + omega L^-1 (b - A y) (y null: out = omega L^-1 b), L the vertical
 * tridiagonal part of the level's operator; host FP32 arrays of the level's N
 * rows (layout of windhip_mg_level). fused = 1: the one-launch kernel the
 * V-cycle uses on small levels (a wavefront per column); 0: the residual and
 * line-solve kernels on the level's FP32 rows and factors; 2: the same kernels
 * on its bf16 rows and factors (levels of up to 44 rows per column; taller
 * levels: as 0); 3 and 4: the one-pass step (residual and line solve in one
 * kernel) on the bf16 (3, the form the V-cycle uses on the larger levels) or
 * FP32 (4) rows and factors, for y given and up to 44 rows per column
 * (otherwise as 2 and 0). A Galerkin level takes form 1 only (WINDHIP_ERR_ARG
 * otherwise). Overwrites the level's work vectors. Inspection only; no counterpart in the reference. */
int windhip_mg_smooth(windhip_solver *s, int level, int fused, const float *b, const float *y, float *out);
/* One V-cycle of the MIXED_MG hierarchy from coarse level `level` (as the
 * inner solve applies it below the fine level) for the rhs b (host FP32, the
 * level's N rows, layout of windhip_mg_level), from a zero guess; out
 * receives the level's correction. fused = 1: the levels of one wavefront per
 * column descend and return in one launch each (residual, restriction and
 * first step; prolongation and first post-step), as the solve runs them by
 * default; 2: as 1, with the cycle from the first such level down as one
 * persistent launch; 0: separate residual, restriction, smoothing and
 * prolongation kernels. Overwrites the hierarchy's work vectors. Inspection only; no
 * counterpart in the reference. */
int windhip_mg_cycle(windhip_solver *s, int level, int fused, const float *b, float *out);
/* The network input for the solver's current rhs (that of its last solve),
 * metrics and alpha, written to in_dev (1 x dims[0] x dims[1] x dims[2] x 8);
 * *scale turns the network output into lambda. dims_out receives the padded
 * volume (nz, ny, nx rounded up to multiples of 8); in_dev may be null to
 * query it. */
int windhip_learned_input(windhip_solver *s, float *in_dev, int *dims_out, double *scale);

/* Per-kernel device time accumulated by solves run with opts.profile: arrays of
 * WINDHIP_PROFILE_KINDS (calls, total ms) indexed by kernel kind: 0..4 FP64 lone
 * sweep (3rd and 4th sweeps of the retry ladder), last sweep + v = A p + sigma,
 * s-update, last sweep + t = A s + 5 dots, x/r/p-update; 5..9 the same in FP32;
 * 10 set-up, 11 FP64 residual of the mixed solve, 12 mixed-solve lambda update,
 * 13 true residual, 14 FGMRES inner set-up, 15 FGMRES FP64 matvec,
 * 16 the Gram-Schmidt orthogonalisation of one FGMRES step (its two passes
 * as one region); MIXED_LINE inner iteration: 17 line solve,
 * 18 residual in - A y, 19 v = A p_hat + sigma, 20 t = A s_hat + 5 dots (its
 * s- and x/r/p-updates count as 7 and 9); MIXED_MG adds 21 fine-level
 * restriction and prolongation, 22 the coarse levels (one region per V-cycle
 * from the fine level's restriction to its prolongation, every launch
 * between them included; `calls` counts those regions), 23 the
 * coarse levels' set-up (its fine-level line solves and residuals count as 17
 * and 18), 24 fine-level one-pass smoothing steps (residual and line solve in
 * one kernel), 25 the fine level's residual and restriction in one kernel.
 * Launches that a finished solve left behind (the rest of a batch of
 * iterations; they return at once) are not counted. reset != 0 clears the
 * counters after reading. */
#define WINDHIP_PROFILE_KINDS 26
int windhip_profile(windhip_solver *s, int reset, double *calls, double *ms);

/* Device memory held by the handle (bytes) and the process high-water mark. */
int windhip_memory(windhip_solver *s, double *bytes_handle, double *bytes_device_used);

#ifdef __cplusplus
}
#endif

#endif /* WINDHIP_H */
