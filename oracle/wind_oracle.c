/*
 * wind_oracle.c -- CPU oracle (test infrastructure only) for the mass-conserving
 * wind solve of HICAR's `wind_iterative` module.
 *
 * THIS FILE IS A CHECKER. Only the tests, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it. The product path (hicar_amd, libwindhip)
 * never links, loads or calls it.
 *
 * It restates, in plain C and FP64 throughout, the algorithm of
 *   src/physics/wind_iterative.F90  (module arrays, assembly, spmv, diagonal,
 *                                    preconditioner, vector ops, rhs, BiCGStab,
 *                                    retry ladder, correction, probing)
 *   src/physics/wind.F90            (calc_divergence: the operator D)
 * Every function cites the reference lines it follows.
 *
 * Arithmetic contract (shared with the HIP kernels so that both produce the same
 * bits): every expression is evaluated left to right exactly as the Fortran
 * source writes it, with no fused multiply-add (build with -ffp-contract=off),
 * x**2 evaluated as x*x, and every dot product summed in the canonical order of
 * canon_reduce() below (the reference leaves the order of its OpenACC / MPI
 * reductions unspecified, so this order is a free choice of the restatement).
 *
 * Layouts
 *   rows   (Krylov vectors, rhs, D_inv): index i + mx*(k + mz*j), i in 0..mx-1,
 *          k in 0..mz-1, j in 0..my-1, mx = nx+2, mz = nz+2, my = ny+2
 *          (wind_iterative.F90:359-370, 2279-2302 on one rank).
 *   fields (Fortran (i,k,j) column-major, 1-based, no halo on one rank):
 *          mass (nx, nz, ny), u (nx+1, nz, ny), v (nx, nz, ny+1), 2-D (nx, ny),
 *          dz_if (nx, nz+1, ny).
 *
 * Parity pins for every function are in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_NCOEF 15
/* Coefficient order A..O, offsets (i,k,j) as in wind_iterative.F90:889-904. */
enum { CA = 0, CB, CC, CD, CE, CF, CG, CH, CI, CJ, CK, CL, CM, CN, CO };

static const double breakdown_eps = 1.0e-30; /* wind_iterative.F90:61 */

typedef struct orc {
    int nx, ny, nz, mx, my, mz;
    double dx;
    long P, N;              /* plane size mx*my, rows N = P*mz */
    /* module arrays (init_module_vars, wind_iterative.F90:2274-2421) */
    double *dz_if;          /* (nx, nz+1, ny) */
    double *dzdx, *dzdy, *jaco, *jaco_w, *alpha, *adv_dz; /* (nx, nz, ny) */
    double *dzdx_surf, *dzdy_surf;                        /* (nx, ny) */
    double *dzdx_u, *jaco_u;                              /* (nx+1, nz, ny) */
    double *dzdy_v, *jaco_v;                              /* (nx, nz, ny+1) */
    double *adv_dz_col;                                   /* (nz) */
    double *coef[ORC_NCOEF];                              /* (nx, nz, ny) */
    int probed;             /* operator_probed */
    int structure_uploaded; /* first-call assembly done */
    int nsweeps;            /* precond_n_sweeps */
    int faithful;           /* FP32-faithful mode (orc_set_faithful) */
    /* Krylov state over rows */
    double *x, *r, *rh, *p, *ph, *v, *s, *sh, *t, *rhs, *dinv, *pres;
} orc_t;

/* ---------------------------------------------------------------- indexing */
#define M3(o, a, i, k, j) (a)[((i) - 1) + (long)(o)->nx * (((k) - 1) + (long)(o)->nz * ((j) - 1))]
#define DZ(o, i, k, j) (o)->dz_if[((i) - 1) + (long)(o)->nx * (((k) - 1) + (long)((o)->nz + 1) * ((j) - 1))]
#define U3(o, a, i, k, j) (a)[((i) - 1) + (long)((o)->nx + 1) * (((k) - 1) + (long)(o)->nz * ((j) - 1))]
#define V3(o, a, i, k, j) (a)[((i) - 1) + (long)(o)->nx * (((k) - 1) + (long)(o)->nz * ((j) - 1))]
#define S2(o, a, i, j) (a)[((i) - 1) + (long)(o)->nx * ((j) - 1)]
#define R(o, a, i, k, j) (a)[(i) + (long)(o)->mx * ((k) + (long)(o)->mz * (j))]
/* u-grid 2-D map factors (nx+1, ny), v-grid (nx, ny+1) */
#define U2(o, a, i, j) (a)[((i) - 1) + (long)((o)->nx + 1) * ((j) - 1)]
#define V2(o, a, i, j) (a)[((i) - 1) + (long)(o)->nx * ((j) - 1)]

static double *zalloc(long n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* ------------------------------------------------------ planted mistakes
 * Test hook: orc_plant(code) makes the oracle commit one plausible mistake,
 * so that the tests can show that their closed-form pins catch it. 0 (the
 * default) is the faithful restatement; nothing outside tests/ sets it.
 *   1  bottom BC row: sign of the x-slope terms flipped
 *   2  bottom BC row: alpha instead of alpha**2 in denom
 *   3  bottom BC row: slope terms over dx instead of 2 dx
 *   4  surface slopes: centred difference over dx instead of 2 dx
 *   5  surface slopes: dzdy_surf differenced along i instead of j
 *   6  bottom BC row: 1/dz_if taken at interface 2 instead of 1
 */
static int g_plant = 0;
void orc_plant(int code) { g_plant = code; }
static inline double f32r(double x) { return (double)(float)x; }

/* ------------------------------------------------------ canonical reduction
 * The order every dot product is summed in (shared with the HIP kernels).
 * Columns c = j*mx + i are grouped in tiles: tile (ti,tj) has 8 waves w of 64
 * lanes l at column i = 62 ti + l - 1, j = 6 tj + w - 1; lanes 1..62 of waves
 * 1..6 that fall on the grid are the tile's own (inner) columns, all other
 * lanes hold 0.
 *  1. per column: acc = 0; acc = acc + f(k) for k = 0..mz-1;
 *  2. per wave: xor butterfly v[l] = v[l] + v[l^off], off = 32,16,8,4,2,1,
 *     giving v[0]; per tile: waves summed in order from wave 0;
 *  3. tiles t = tj*nti + ti are folded by 512 lanes, lane q summing tiles
 *     q, q+512, ... in order from 0, then reduced like one tile.
 */
#define CANON_WAVES 8
#define CANON_TI 62
#define CANON_TJ (CANON_WAVES - 2)

static void butterfly64(double *v)
{
    double tmp[64];
    for (int off = 32; off >= 1; off >>= 1) {
        for (int l = 0; l < 64; ++l) tmp[l] = v[l] + v[l ^ off];
        memcpy(v, tmp, sizeof(tmp));
    }
}

static double block_of_waves(const double *vals /* CANON_WAVES*64 */)
{
    double s = 0.0;
    for (int w = 0; w < CANON_WAVES; ++w) {
        double v[64];
        memcpy(v, vals + 64 * w, sizeof(v));
        butterfly64(v);
        s = (w == 0) ? v[0] : s + v[0];
    }
    return s;
}

double orc_canon_reduce(const double *colsum, int mx, int my)
{
    const int nti = (mx + CANON_TI - 1) / CANON_TI, ntj = (my + CANON_TJ - 1) / CANON_TJ;
    const long nt = (long)nti * ntj;
    double *part = zalloc(nt);
    for (long t = 0; t < nt; ++t) {
        const int ti = (int)(t % nti), tj = (int)(t / nti);
        double vals[CANON_WAVES * 64];
        for (int w = 0; w < CANON_WAVES; ++w)
            for (int l = 0; l < 64; ++l) {
                const int i = CANON_TI * ti + l - 1, j = CANON_TJ * tj + w - 1;
                const int inner = l >= 1 && l <= CANON_TI && w >= 1 && w <= CANON_TJ && i < mx && j < my;
                vals[64 * w + l] = inner ? colsum[(long)j * mx + i] : 0.0;
            }
        part[t] = block_of_waves(vals);
    }
    double lanes[CANON_WAVES * 64];
    for (int q = 0; q < CANON_WAVES * 64; ++q) {
        double acc = 0.0;
        for (long t = q; t < nt; t += CANON_WAVES * 64) acc = acc + part[t];
        lanes[q] = acc;
    }
    free(part);
    return block_of_waves(lanes);
}

/* colsum of a*b over the rows of one column; k ascending from 0 */
static void colsums_dot(const orc_t *o, const double *a, const double *b, double *cs)
{
    long P = o->P;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < o->my; ++j) {
        for (int i = 0; i < o->mx; ++i) cs[(long)j * o->mx + i] = 0.0;
        for (int k = 0; k < o->mz; ++k)
            for (int i = 0; i < o->mx; ++i) {
                long c = (long)j * o->mx + i;
                cs[c] = cs[c] + R(o, a, i, k, j) * R(o, b, i, k, j);
            }
    }
    (void)P;
}

#ifdef ORC_BASELINE
/* CPU-baseline build: dot products as plain parallel sums (the reference's
 * local sums and MPI allreduce fix no order either); no bitwise contract. */
double orc_dot(const orc_t *o, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (long n = 0; n < o->N; ++n) s += a[n] * b[n];
    return s;
}
#else
double orc_dot(const orc_t *o, const double *a, const double *b)
{
    double *cs = zalloc(o->P);
    colsums_dot(o, a, b, cs);
    double s = orc_canon_reduce(cs, o->mx, o->my);
    free(cs);
    return s;
}
#endif

/* ------------------------------------------------------------ create/free */
orc_t *orc_create(int nx, int ny, int nz, double dx)
{
    if (nx < 3 || ny < 3 || nz < 2) return NULL;
    orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
    o->nx = nx; o->ny = ny; o->nz = nz;
    o->mx = nx + 2; o->my = ny + 2; o->mz = nz + 2;   /* wind_iterative.F90:2295-2297 */
    o->dx = dx;
    o->P = (long)o->mx * o->my;
    o->N = o->P * o->mz;
    long n3 = (long)nx * nz * ny;
    o->dz_if = zalloc((long)nx * (nz + 1) * ny);
    o->dzdx = zalloc(n3); o->dzdy = zalloc(n3); o->jaco = zalloc(n3);
    o->jaco_w = zalloc(n3); o->alpha = zalloc(n3); o->adv_dz = zalloc(n3);
    o->dzdx_surf = zalloc((long)nx * ny); o->dzdy_surf = zalloc((long)nx * ny);
    o->dzdx_u = zalloc((long)(nx + 1) * nz * ny); o->jaco_u = zalloc((long)(nx + 1) * nz * ny);
    o->dzdy_v = zalloc((long)nx * nz * (ny + 1)); o->jaco_v = zalloc((long)nx * nz * (ny + 1));
    o->adv_dz_col = zalloc(nz);
    for (int q = 0; q < ORC_NCOEF; ++q) o->coef[q] = zalloc(n3);
    double **vecs[] = {&o->x, &o->r, &o->rh, &o->p, &o->ph, &o->v, &o->s, &o->sh, &o->t, &o->rhs, &o->dinv, &o->pres};
    for (size_t q = 0; q < sizeof(vecs) / sizeof(vecs[0]); ++q) *vecs[q] = zalloc(o->N);
    for (long n = 0; n < o->N; ++n) o->dinv[n] = 1.0;  /* D_inv = 1 at allocation, :394 */
    o->nsweeps = 2;                                    /* BASE_PREC_SWEEPS, :76 */
    return o;
}

void orc_free(orc_t *o)
{
    if (!o) return;
    double *arrs[] = {o->dz_if, o->dzdx, o->dzdy, o->jaco, o->jaco_w, o->alpha, o->adv_dz,
                      o->dzdx_surf, o->dzdy_surf, o->dzdx_u, o->jaco_u, o->dzdy_v, o->jaco_v,
                      o->adv_dz_col, o->x, o->r, o->rh, o->p, o->ph, o->v, o->s, o->sh, o->t,
                      o->rhs, o->dinv, o->pres};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); ++q) free(arrs[q]);
    for (int q = 0; q < ORC_NCOEF; ++q) free(o->coef[q]);
    free(o);
}

long orc_rows(const orc_t *o) { return o->N; }

/* ---------------------------------------------------------- init_module_vars
 * wind_iterative.F90:2274-2421 on one rank (i_s = ids = 1, i_e = ide = nx,
 * k_s = 1, k_e = nz). `terrain` is neighbor_terrain over (nx, ny).
 */
void orc_set_metrics(orc_t *o, const double *dzdx, const double *dzdy, const double *jaco,
                     const double *jaco_u, const double *jaco_v, const double *jaco_w,
                     const double *dzdx_u, const double *dzdy_v, const double *adv_dz,
                     const double *terrain)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long n3 = (long)nx * nz * ny;
    memcpy(o->adv_dz, adv_dz, n3 * sizeof(double));
    memcpy(o->dzdx, dzdx, n3 * sizeof(double));
    memcpy(o->dzdy, dzdy, n3 * sizeof(double));
    memcpy(o->jaco, jaco, n3 * sizeof(double));
    memcpy(o->jaco_w, jaco_w, n3 * sizeof(double));
    memcpy(o->dzdx_u, dzdx_u, (long)(nx + 1) * nz * ny * sizeof(double));
    memcpy(o->jaco_u, jaco_u, (long)(nx + 1) * nz * ny * sizeof(double));
    memcpy(o->dzdy_v, dzdy_v, (long)nx * nz * (ny + 1) * sizeof(double));
    memcpy(o->jaco_v, jaco_v, (long)nx * nz * (ny + 1) * sizeof(double));
    if (o->faithful) {
        /* the module arrays are `real` (wind_iterative.F90:89-91) */
        double *arrs[] = {o->adv_dz, o->dzdx, o->dzdy, o->jaco, o->jaco_w};
        for (int q = 0; q < 5; ++q)
            for (long n = 0; n < n3; ++n) arrs[q][n] = f32r(arrs[q][n]);
        for (long n = 0; n < (long)(nx + 1) * nz * ny; ++n) {
            o->dzdx_u[n] = f32r(o->dzdx_u[n]); o->jaco_u[n] = f32r(o->jaco_u[n]);
        }
        for (long n = 0; n < (long)nx * nz * (ny + 1); ++n) {
            o->dzdy_v[n] = f32r(o->dzdy_v[n]); o->jaco_v[n] = f32r(o->jaco_v[n]);
        }
    }
    const double *az = o->adv_dz;
    /* dz_if: interface mean of advection_dz, :2350-2362 (in FP32 arithmetic in
     * the faithful mode: (a + b) / 2 with a, b `real`) */
    for (int j = 1; j <= ny; ++j)
        for (int k = 2; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                if (o->faithful)
                    DZ(o, i, k, j) = (double)(((float)M3(o, az, i, k, j) + (float)M3(o, az, i, k - 1, j)) / 2.0f);
                else
                    DZ(o, i, k, j) = (M3(o, az, i, k, j) + M3(o, az, i, k - 1, j)) / 2;
            }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            DZ(o, i, 1, j) = M3(o, az, i, 1, j);
            DZ(o, i, nz + 1, j) = M3(o, az, i, nz, j);
            S2(o, o->dzdx_surf, i, j) = M3(o, o->dzdx, i, 1, j);
            S2(o, o->dzdy_surf, i, j) = M3(o, o->dzdy, i, 1, j);
        }
    /* centred terrain slope away from the domain edge, :2398-2409: i in
     * i_s_bnd..i_e_bnd = 2..nx-1 for dzdx_surf (all j), j in 2..ny-1 for
     * dzdy_surf (all i); the edge columns keep dzdx(i,k_s,j) / dzdy(i,k_s,j)
     * from :2363-2364 */
    const double two_dx = g_plant == 4 ? o->dx : 2 * o->dx;
    for (int j = 1; j <= ny; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            if (o->faithful)
                S2(o, o->dzdx_surf, i, j) = (double)(((float)S2(o, terrain, i + 1, j) - (float)S2(o, terrain, i - 1, j))
                                                     / (float)two_dx);
            else
                S2(o, o->dzdx_surf, i, j) = (S2(o, terrain, i + 1, j) - S2(o, terrain, i - 1, j)) / two_dx;
        }
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 1; i <= nx; ++i) {
            double tp = S2(o, terrain, i, j + 1), tm = S2(o, terrain, i, j - 1);
            if (g_plant == 5) {
                if (i < 2 || i > nx - 1) continue;
                tp = S2(o, terrain, i + 1, j); tm = S2(o, terrain, i - 1, j);
            }
            if (o->faithful)
                S2(o, o->dzdy_surf, i, j) = (double)(((float)tp - (float)tm) / (float)two_dx);
            else
                S2(o, o->dzdy_surf, i, j) = (tp - tm) / two_dx;
        }
    for (int k = 1; k <= nz; ++k) o->adv_dz_col[k - 1] = M3(o, az, 1, k, 1); /* :2410-2413 */
}

/* calc_iter_winds :309-317 copies alpha over the solver range */
void orc_set_alpha(orc_t *o, const double *alpha)
{
    long n3 = (long)o->nx * o->nz * o->ny;
    memcpy(o->alpha, alpha, n3 * sizeof(double));
    if (o->faithful)
        for (long n = 0; n < n3; ++n) o->alpha[n] = f32r(o->alpha[n]);
}

/* ------------------------------------------------------------- assembly */
/* update_coefs_gpu, wind_iterative.F90:2555-2603: B, C (with X) and A per cell */
/* FP32-faithful update_coefs_gpu: every operand and intermediate `real`
 * (wind_iterative.F90:2556-2558), the same expression order as below. */
static void update_coefs_f32(orc_t *o)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    const float dx = (float)o->dx;
    double *B = o->coef[CB], *C = o->coef[CC], *A = o->coef[CA];
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                float adv_dz_k = (float)o->adv_dz_col[k - 1];
                float M_up, M_dwn;
                float dzdx = (float)M3(o, o->dzdx, i, k, j), dzdy = (float)M3(o, o->dzdy, i, k, j);
                float al = (float)M3(o, o->alpha, i, k, j);
                if (k == nz) {
                    M_up = (dzdx * dzdx + dzdy * dzdy) + al * al;
                } else {
                    float a1 = (float)o->adv_dz_col[k];
                    float sum_up = adv_dz_k + a1;
                    float jw = (float)M3(o, o->jaco_w, i, k, j);
                    float dzdx1 = (float)M3(o, o->dzdx, i, k + 1, j), dzdy1 = (float)M3(o, o->dzdy, i, k + 1, j);
                    float al1 = (float)M3(o, o->alpha, i, k + 1, j);
                    M_up = ((dzdx * (dzdx * a1 + dzdx1 * adv_dz_k)) / sum_up) / jw
                         + ((dzdy * (dzdy * a1 + dzdy1 * adv_dz_k)) / sum_up) / jw;
                    M_up = M_up + ((al * al * a1 + al1 * al1 * adv_dz_k) / sum_up) / jw;
                }
                if (k == 1) {
                    M_dwn = (dzdx * (float)S2(o, o->dzdx_surf, i, j) + dzdy * (float)S2(o, o->dzdy_surf, i, j)) + al * al;
                } else {
                    float am1 = (float)o->adv_dz_col[k - 2];
                    float sum_dwn = adv_dz_k + am1;
                    float jw = (float)M3(o, o->jaco_w, i, k - 1, j);
                    float dzdxm = (float)M3(o, o->dzdx, i, k - 1, j), dzdym = (float)M3(o, o->dzdy, i, k - 1, j);
                    float alm = (float)M3(o, o->alpha, i, k - 1, j);
                    M_dwn = ((dzdx * (dzdx * am1 + dzdxm * adv_dz_k)) / sum_dwn) / jw
                          + ((dzdy * (dzdy * am1 + dzdym * adv_dz_k)) / sum_dwn) / jw;
                    M_dwn = M_dwn + ((al * al * am1 + alm * alm * adv_dz_k) / sum_dwn) / jw;
                }
                float jc = (float)M3(o, o->jaco, i, k, j);
                float b = M_up / ((jc * adv_dz_k) * (float)DZ(o, 1, k + 1, 1));
                float c = M_dwn / ((jc * adv_dz_k) * (float)DZ(o, 1, k, 1));
                float md = (((float)DZ(o, i, k + 1, j) + (float)DZ(o, i, k, j)) * 2.0f) * dx;
                float X = -((((float)U3(o, o->dzdx_u, i + 1, k, j) / (float)U3(o, o->jaco_u, i + 1, k, j)
                              - (float)U3(o, o->dzdx_u, i, k, j) / (float)U3(o, o->jaco_u, i, k, j))
                             + ((float)V3(o, o->dzdy_v, i, k, j + 1) / (float)V3(o, o->jaco_v, i, k, j + 1)
                                - (float)V3(o, o->dzdy_v, i, k, j) / (float)V3(o, o->jaco_v, i, k, j))) / md);
                b = b + X;
                c = c - X;
                M3(o, B, i, k, j) = b;
                M3(o, C, i, k, j) = c;
                M3(o, A, i, k, j) = (-4.0f / (dx * dx) - b) - c;
            }
}

/* FP32-faithful initialize_coefs (:2424-2469): D..G and H..O in `real`. */
static void assemble_f32(orc_t *o)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    const float dx = (float)o->dx;
    const float inv = 1.0f / (dx * dx);
    long n3 = (long)nx * nz * ny;
    for (long n = 0; n < n3; ++n) {
        o->coef[CA][n] = 1.0;
        o->coef[CD][n] = inv; o->coef[CE][n] = inv; o->coef[CF][n] = inv; o->coef[CG][n] = inv;
    }
    for (int k = 1; k <= nz; ++k) {
        float md = (2.0f * dx) * ((float)DZ(o, 1, k + 1, 1) + (float)DZ(o, 1, k, 1));
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                float jc = (float)M3(o, o->jaco, i, k, j);
                float ax = (float)M3(o, o->dzdx, i, k, j) / jc;
                float ay = (float)M3(o, o->dzdy, i, k, j) / jc;
                float hx = ax + (float)U3(o, o->dzdx_u, i + 1, k, j) / (float)U3(o, o->jaco_u, i + 1, k, j);
                float ix = ax + (float)U3(o, o->dzdx_u, i, k, j) / (float)U3(o, o->jaco_u, i, k, j);
                float ly = ay + (float)V3(o, o->dzdy_v, i, k, j + 1) / (float)V3(o, o->jaco_v, i, k, j + 1);
                float my = ay + (float)V3(o, o->dzdy_v, i, k, j) / (float)V3(o, o->jaco_v, i, k, j);
                M3(o, o->coef[CH], i, k, j) = -(hx / md);
                M3(o, o->coef[CI], i, k, j) = ix / md;
                M3(o, o->coef[CL], i, k, j) = -(ly / md);
                M3(o, o->coef[CM], i, k, j) = my / md;
                M3(o, o->coef[CJ], i, k, j) = hx / md;
                M3(o, o->coef[CK], i, k, j) = -(ix / md);
                M3(o, o->coef[CN], i, k, j) = ly / md;
                M3(o, o->coef[CO], i, k, j) = -(my / md);
            }
    }
    update_coefs_f32(o);
    o->probed = 0;
}

void orc_update_coefs(orc_t *o)
{
    if (o->faithful) { update_coefs_f32(o); return; }
    int nx = o->nx, ny = o->ny, nz = o->nz;
    double dx = o->dx;
    double *B = o->coef[CB], *C = o->coef[CC], *A = o->coef[CA];
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                double adv_dz_k = o->adv_dz_col[k - 1];
                double M_up, M_dwn;
                double dzdx = M3(o, o->dzdx, i, k, j), dzdy = M3(o, o->dzdy, i, k, j);
                double al = M3(o, o->alpha, i, k, j);
                if (k == nz) {
                    M_up = (dzdx * dzdx + dzdy * dzdy) + al * al;
                } else {
                    double a1 = o->adv_dz_col[k];
                    double sum_up = adv_dz_k + a1;
                    double jw = M3(o, o->jaco_w, i, k, j);
                    double dzdx1 = M3(o, o->dzdx, i, k + 1, j), dzdy1 = M3(o, o->dzdy, i, k + 1, j);
                    double al1 = M3(o, o->alpha, i, k + 1, j);
                    M_up = ((dzdx * (dzdx * a1 + dzdx1 * adv_dz_k)) / sum_up) / jw
                         + ((dzdy * (dzdy * a1 + dzdy1 * adv_dz_k)) / sum_up) / jw;
                    M_up = M_up + ((al * al * a1 + al1 * al1 * adv_dz_k) / sum_up) / jw;
                }
                if (k == 1) {
                    M_dwn = (dzdx * S2(o, o->dzdx_surf, i, j) + dzdy * S2(o, o->dzdy_surf, i, j)) + al * al;
                } else {
                    double am1 = o->adv_dz_col[k - 2];
                    double sum_dwn = adv_dz_k + am1;
                    double jw = M3(o, o->jaco_w, i, k - 1, j);
                    double dzdxm = M3(o, o->dzdx, i, k - 1, j), dzdym = M3(o, o->dzdy, i, k - 1, j);
                    double alm = M3(o, o->alpha, i, k - 1, j);
                    M_dwn = ((dzdx * (dzdx * am1 + dzdxm * adv_dz_k)) / sum_dwn) / jw
                          + ((dzdy * (dzdy * am1 + dzdym * adv_dz_k)) / sum_dwn) / jw;
                    M_dwn = M_dwn + ((al * al * am1 + alm * alm * adv_dz_k) / sum_dwn) / jw;
                }
                double jc = M3(o, o->jaco, i, k, j);
                double b = M_up / ((jc * adv_dz_k) * DZ(o, 1, k + 1, 1));
                double c = M_dwn / ((jc * adv_dz_k) * DZ(o, 1, k, 1));
                double md = ((DZ(o, i, k + 1, j) + DZ(o, i, k, j)) * 2.0) * dx;
                double X = -(((U3(o, o->dzdx_u, i + 1, k, j) / U3(o, o->jaco_u, i + 1, k, j)
                               - U3(o, o->dzdx_u, i, k, j) / U3(o, o->jaco_u, i, k, j))
                              + (V3(o, o->dzdy_v, i, k, j + 1) / V3(o, o->jaco_v, i, k, j + 1)
                                 - V3(o, o->dzdy_v, i, k, j) / V3(o, o->jaco_v, i, k, j))) / md);
                b = b + X;
                c = c - X;
                M3(o, B, i, k, j) = b;
                M3(o, C, i, k, j) = c;
                M3(o, A, i, k, j) = (-4.0 / (dx * dx) - b) - c;
            }
}

/* initialize_coefs, wind_iterative.F90:2424-2469 (D..G, H..O), then B, C, A.
 * B/C/A use the per-cell form of update_coefs_gpu (:2555-2603) so that the first
 * call and every later alpha refresh produce the same operator. */
void orc_assemble(orc_t *o)
{
    if (o->faithful) { assemble_f32(o); return; }
    int nx = o->nx, ny = o->ny, nz = o->nz;
    double dx = o->dx;
    double inv = 1.0 / (dx * dx);
    long n3 = (long)nx * nz * ny;
    for (long n = 0; n < n3; ++n) {
        o->coef[CA][n] = 1.0;
        o->coef[CD][n] = inv; o->coef[CE][n] = inv; o->coef[CF][n] = inv; o->coef[CG][n] = inv;
    }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k) {
            double md = (2 * dx) * (DZ(o, 1, k + 1, 1) + DZ(o, 1, k, 1));
            for (int i = 1; i <= nx; ++i) {
                double ax = M3(o, o->dzdx, i, k, j) / M3(o, o->jaco, i, k, j);
                double ay = M3(o, o->dzdy, i, k, j) / M3(o, o->jaco, i, k, j);
                double hx = ax + U3(o, o->dzdx_u, i + 1, k, j) / U3(o, o->jaco_u, i + 1, k, j);
                double ix = ax + U3(o, o->dzdx_u, i, k, j) / U3(o, o->jaco_u, i, k, j);
                double ly = ay + V3(o, o->dzdy_v, i, k, j + 1) / V3(o, o->jaco_v, i, k, j + 1);
                double my = ay + V3(o, o->dzdy_v, i, k, j) / V3(o, o->jaco_v, i, k, j);
                M3(o, o->coef[CH], i, k, j) = -(hx / md);
                M3(o, o->coef[CI], i, k, j) = ix / md;
                M3(o, o->coef[CL], i, k, j) = -(ly / md);
                M3(o, o->coef[CM], i, k, j) = my / md;
                M3(o, o->coef[CJ], i, k, j) = hx / md;
                M3(o, o->coef[CK], i, k, j) = -(ix / md);
                M3(o, o->coef[CN], i, k, j) = ly / md;
                M3(o, o->coef[CO], i, k, j) = -(my / md);
            }
    }
    orc_update_coefs(o);
    o->probed = 0;
}

/* -------------------------------------------------------------- spmv
 * wind_iterative.F90:813-1027. Analytic branch: lateral identity rows, top
 * 2-point row, bottom 10-point terrain row, interior 15-point with D=E=F=G
 * replaced by the scalar 1/dx^2 and A derived from B and C. Probed branch
 * (:837-910): lateral and ghost-plane identity rows, 15 stored arrays.
 */
void orc_spmv(const orc_t *o, const double *x, double *y)
{
    int mx = o->mx, my = o->my, mz = o->mz;
    double dx = o->dx;
    double two_dx = 2.0 * dx;
    double inv_dx2 = 1.0 / (dx * dx);
    double neg4 = -4.0 * inv_dx2;
    /* lateral identity rows */
    for (int j = 0; j < my; ++j)
        for (int k = 0; k < mz; ++k) {
            R(o, y, 0, k, j) = R(o, x, 0, k, j);
            R(o, y, mx - 1, k, j) = R(o, x, mx - 1, k, j);
        }
    for (int k = 0; k < mz; ++k)
        for (int i = 0; i < mx; ++i) {
            R(o, y, i, k, 0) = R(o, x, i, k, 0);
            R(o, y, i, k, my - 1) = R(o, x, i, k, my - 1);
        }
    if (o->probed) {
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= my - 2; ++j) {
            for (int i = 1; i <= mx - 2; ++i) {
                R(o, y, i, 0, j) = R(o, x, i, 0, j);
                R(o, y, i, mz - 1, j) = R(o, x, i, mz - 1, j);
            }
            for (int k = 1; k <= mz - 2; ++k)
                for (int i = 1; i <= mx - 2; ++i) {
                    double *const *cf = o->coef;
#define CF(q) M3(o, cf[q], i, k, j)
                    R(o, y, i, k, j) = CF(CO) * R(o, x, i, k - 1, j - 1) + CF(CG) * R(o, x, i, k, j - 1)
                        + CF(CM) * R(o, x, i, k + 1, j - 1) + CF(CK) * R(o, x, i - 1, k - 1, j)
                        + CF(CC) * R(o, x, i, k - 1, j) + CF(CJ) * R(o, x, i + 1, k - 1, j)
                        + CF(CE) * R(o, x, i - 1, k, j) + CF(CA) * R(o, x, i, k, j)
                        + CF(CD) * R(o, x, i + 1, k, j) + CF(CI) * R(o, x, i - 1, k + 1, j)
                        + CF(CB) * R(o, x, i, k + 1, j) + CF(CH) * R(o, x, i + 1, k + 1, j)
                        + CF(CN) * R(o, x, i, k - 1, j + 1) + CF(CF) * R(o, x, i, k, j + 1)
                        + CF(CL) * R(o, x, i, k + 1, j + 1);
#undef CF
                }
        }
        return;
    }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= my - 2; ++j) {
        /* top BC, :950-960 */
        for (int i = 1; i <= mx - 2; ++i)
            R(o, y, i, mz - 1, j) = (R(o, x, i, mz - 1, j) - R(o, x, i, mz - 2, j)) / DZ(o, i, mz - 1, j);
        /* bottom BC, :962-988 */
        for (int i = 1; i <= mx - 2; ++i) {
            double dzdx_s = S2(o, o->dzdx_surf, i, j);
            double dzdy_s = S2(o, o->dzdy_surf, i, j);
            double dz_kp1 = DZ(o, i, g_plant == 6 ? 2 : 1, j);
            double al = M3(o, o->alpha, i, 1, j);
            double al2 = g_plant == 2 ? al : al * al;
            double denom = (2.0 * ((dzdx_s * dzdx_s + dzdy_s * dzdy_s) + al2)) / M3(o, o->jaco, i, 1, j);
            double sdx = g_plant == 3 ? dx : two_dx;
            double cy = dzdy_s / (denom * sdx);
            double cx = dzdx_s / (denom * sdx);
            if (g_plant == 1) cx = -cx;
            double cz = 1.0 / dz_kp1;
            R(o, y, i, 0, j) = cy * R(o, x, i, 0, j - 1) + cy * R(o, x, i, 1, j - 1)
                + cx * R(o, x, i - 1, 0, j) + (-cz) * R(o, x, i, 0, j)
                + (-cx) * R(o, x, i + 1, 0, j) + cx * R(o, x, i - 1, 1, j)
                + cz * R(o, x, i, 1, j) + (-cx) * R(o, x, i + 1, 1, j)
                + (-cy) * R(o, x, i, 0, j + 1) + (-cy) * R(o, x, i, 1, j + 1);
        }
        /* interior, :990-1025 */
        for (int k = 1; k <= mz - 2; ++k)
            for (int i = 1; i <= mx - 2; ++i) {
                double b_val = M3(o, o->coef[CB], i, k, j);
                double c_val = M3(o, o->coef[CC], i, k, j);
                double a_val = (neg4 - b_val) - c_val;
                R(o, y, i, k, j) = M3(o, o->coef[CO], i, k, j) * R(o, x, i, k - 1, j - 1)
                    + inv_dx2 * R(o, x, i, k, j - 1)
                    + M3(o, o->coef[CM], i, k, j) * R(o, x, i, k + 1, j - 1)
                    + M3(o, o->coef[CK], i, k, j) * R(o, x, i - 1, k - 1, j)
                    + c_val * R(o, x, i, k - 1, j)
                    + M3(o, o->coef[CJ], i, k, j) * R(o, x, i + 1, k - 1, j)
                    + inv_dx2 * R(o, x, i - 1, k, j)
                    + a_val * R(o, x, i, k, j)
                    + inv_dx2 * R(o, x, i + 1, k, j)
                    + M3(o, o->coef[CI], i, k, j) * R(o, x, i - 1, k + 1, j)
                    + b_val * R(o, x, i, k + 1, j)
                    + M3(o, o->coef[CH], i, k, j) * R(o, x, i + 1, k + 1, j)
                    + M3(o, o->coef[CN], i, k, j) * R(o, x, i, k - 1, j + 1)
                    + inv_dx2 * R(o, x, i, k, j + 1)
                    + M3(o, o->coef[CL], i, k, j) * R(o, x, i, k + 1, j + 1);
            }
    }
}

/* extract_diagonal, wind_iterative.F90:1041-1075 */
void orc_extract_diagonal(orc_t *o)
{
    int mx = o->mx, my = o->my, mz = o->mz;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < my; ++j)
        for (int k = 0; k < mz; ++k)
            for (int i = 0; i < mx; ++i) {
                double d;
                if (i <= 0 || j <= 0 || i >= mx - 1 || j >= my - 1) d = 1.0;
                else if (k >= mz - 1) d = o->probed ? 1.0 : 1.0 / DZ(o, i, k, j);
                else if (k <= 0) d = o->probed ? 1.0 : -1.0 / DZ(o, i, k + 1, j);
                else d = M3(o, o->coef[CA], i, k, j);
                if (fabs(d) < breakdown_eps) d = copysign(breakdown_eps, d);
                R(o, o->dinv, i, k, j) = 1.0 / d;
            }
}

/* compute_rhs_3d, wind_iterative.F90:1513-1530: b = -2 div on interior rows */
void orc_compute_rhs(orc_t *o, const double *div)
{
    int mx = o->mx, my = o->my, mz = o->mz;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < my; ++j)
        for (int k = 0; k < mz; ++k)
            for (int i = 0; i < mx; ++i) {
                if (i <= 0 || j <= 0 || i >= mx - 1 || j >= my - 1 || k <= 0 || k >= mz - 1)
                    R(o, o->rhs, i, k, j) = 0.0;
                else if (o->faithful) /* real(-2.0 * div, c_double) with div `real`, :1525 */
                    R(o, o->rhs, i, k, j) = (double)(-2.0f * (float)M3(o, div, i, k, j));
                else
                    R(o, o->rhs, i, k, j) = -2.0 * M3(o, div, i, k, j);
            }
}

/* apply_precond, wind_iterative.F90:1268-1315: y1 = D^-1 in; y += D^-1 (in - A y) */
void orc_apply_precond(orc_t *o, int nsweeps, const double *in, double *out)
{
    long N = o->N;
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) out[n] = o->dinv[n] * in[n];
    for (int sw = 2; sw <= nsweeps; ++sw) {
        orc_spmv(o, out, o->pres);
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; ++n) out[n] = out[n] + o->dinv[n] * (in[n] - o->pres[n]);
    }
}

/* ------------------------------------------------------------ BiCGStab
 * bicgstab_solve, wind_iterative.F90:555-791 (single rank: allreduces and halo
 * exchanges are identities). trace (optional): per iteration
 * {sigma, alpha, omega, rho_new, rnorm}.
 */
int orc_bicgstab(orc_t *o, int max_iters, double tol_rel, double tol_abs,
                 int *status_out, int *iters_out, double *res0_out, double *resf_out,
                 double *trace, int trace_cap)
{
    long N = o->N;
    double *x = o->x, *r = o->r, *rh = o->rh, *p = o->p, *ph = o->ph, *v = o->v;
    double *s = o->s, *sh = o->sh, *t = o->t, *b = o->rhs;
    *status_out = 1; *iters_out = 0; *res0_out = 0.0; *resf_out = 0.0;
    orc_spmv(o, x, t);
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) r[n] = 1.0 * b[n] + (-1.0) * t[n];
    memcpy(rh, r, N * sizeof(double));
    memcpy(p, r, N * sizeof(double));
    double rho = orc_dot(o, rh, r);
    double rr = orc_dot(o, r, r);
    double bb = orc_dot(o, b, b);
    double rnorm = sqrt(rr);
    *res0_out = rnorm; *resf_out = rnorm;
    double target = fmax(tol_abs, tol_rel * sqrt(bb));
    if (rnorm <= target) { *status_out = 0; *iters_out = 0; return 0; }
    double omega = 1.0, alpha = 1.0;
    double *cs = zalloc(o->P * 5);
    for (int it = 1; it <= max_iters; ++it) {
        orc_apply_precond(o, o->nsweeps, p, ph);
        orc_spmv(o, ph, v);
        double sigma = orc_dot(o, rh, v);
        if (trace && it <= trace_cap) trace[5 * (it - 1)] = sigma;
        if (fabs(sigma) < breakdown_eps) { *status_out = 2; *iters_out = it; break; }
        alpha = rho / sigma;
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; ++n) s[n] = 1.0 * r[n] + (-alpha) * v[n];
        orc_apply_precond(o, o->nsweeps, s, sh);
        orc_spmv(o, sh, t);
        /* vec_dots_fused :1454-1488 */
#ifdef ORC_BASELINE
        double ts = 0.0, tt = 0.0, rs = 0.0, rt = 0.0, ss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ts, tt, rs, rt, ss)
        for (long n = 0; n < N; ++n) {
            double tv = t[n], sv = s[n], rv = rh[n];
            ts += tv * sv; tt += tv * tv; rs += rv * sv; rt += rv * tv; ss += sv * sv;
        }
#else
        double *cts = cs, *ctt = cs + o->P, *crs = cs + 2 * o->P, *crt = cs + 3 * o->P, *css = cs + 4 * o->P;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < o->my; ++j) {
            for (int i = 0; i < o->mx; ++i) {
                long c = (long)j * o->mx + i;
                cts[c] = 0.0; ctt[c] = 0.0; crs[c] = 0.0; crt[c] = 0.0; css[c] = 0.0;
            }
            for (int k = 0; k < o->mz; ++k)
                for (int i = 0; i < o->mx; ++i) {
                    long c = (long)j * o->mx + i;
                    double tv = R(o, t, i, k, j), sv = R(o, s, i, k, j), rv = R(o, rh, i, k, j);
                    cts[c] = cts[c] + tv * sv;
                    ctt[c] = ctt[c] + tv * tv;
                    crs[c] = crs[c] + rv * sv;
                    crt[c] = crt[c] + rv * tv;
                    css[c] = css[c] + sv * sv;
                }
        }
        double ts = orc_canon_reduce(cts, o->mx, o->my), tt = orc_canon_reduce(ctt, o->mx, o->my);
        double rs = orc_canon_reduce(crs, o->mx, o->my), rt = orc_canon_reduce(crt, o->mx, o->my);
        double ss = orc_canon_reduce(css, o->mx, o->my);
#endif
        if (tt < breakdown_eps) { *status_out = 2; *iters_out = it; break; }
        omega = ts / tt;
        double rho_new = rs - omega * rt;
        double rn2 = (ss - (2.0 * omega) * ts) + (omega * omega) * tt;
        if (rn2 < 0.0) rn2 = 0.0;
        rnorm = sqrt(rn2);
        *resf_out = rnorm; *iters_out = it;
        if (trace && it <= trace_cap) {
            trace[5 * (it - 1) + 1] = alpha; trace[5 * (it - 1) + 2] = omega;
            trace[5 * (it - 1) + 3] = rho_new; trace[5 * (it - 1) + 4] = rnorm;
        }
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; ++n) x[n] = (x[n] + alpha * ph[n]) + omega * sh[n];
        if (rnorm <= target) { *status_out = 0; break; }
        if (fabs(omega) < breakdown_eps) { *status_out = 2; break; }
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; ++n) r[n] = 1.0 * s[n] + (-omega) * t[n];
        if (fabs(rho) < breakdown_eps) { *status_out = 2; break; }
        double beta = (rho_new / rho) * (alpha / omega);
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; ++n) p[n] = r[n] + beta * (p[n] - omega * v[n]);
        rho = rho_new;
    }
    free(cs);
    return 0;
}

/* calc_iter_winds solve part, wind_iterative.F90:295-528 (no correction):
 * alpha/div copy, first-call assembly or alpha-range refresh, diagonal, rhs,
 * BiCGStab from the current x (warm start), retry ladder. Returns 0, or 3 when
 * the reference would `stop` (diverged after the ladder, :506-528). */
int orc_solve(orc_t *o, const double *alpha, const double *div, int max_iters,
              double tol_rel, double tol_abs, int *status, int *iters, double *res0,
              double *resf, int *nsweeps_used)
{
    orc_set_alpha(o, alpha);
    if (!o->structure_uploaded) {
        orc_assemble(o);
        o->structure_uploaded = 1;
    } else {
        long n3 = (long)o->nx * o->nz * o->ny;
        double amin = HUGE_VAL, amax = -HUGE_VAL;
#pragma omp parallel for schedule(static) reduction(min : amin) reduction(max : amax)
        for (long n = 0; n < n3; ++n) { amin = fmin(amin, o->alpha[n]); amax = fmax(amax, o->alpha[n]); }
        if (amax > amin && !o->probed) orc_update_coefs(o);
    }
    orc_extract_diagonal(o);
    orc_compute_rhs(o, div);
    orc_bicgstab(o, max_iters, tol_rel, tol_abs, status, iters, res0, resf, NULL, 0);
    *nsweeps_used = o->nsweeps;
    int rc = 0;
    if (*status != 0 && *iters > 0 && *resf > 0.01 * *res0) {
        while (*status != 0 && *resf > 0.01 * *res0 && o->nsweeps < 4) {
            o->nsweeps += 1;
            memset(o->x, 0, o->N * sizeof(double));   /* vec_zero over the owned rows */
            orc_bicgstab(o, max_iters, tol_rel, tol_abs, status, iters, res0, resf, NULL, 0);
            *nsweeps_used = o->nsweeps;
        }
        o->nsweeps = 2;
        if (*status != 0 && *resf > *res0) rc = 3;
    }
    return rc;
}

void orc_set_nsweeps(orc_t *o, int n) { o->nsweeps = n; }
/* FP32-faithful mode: the oracle rounds where the reference stores or computes
 * in `real` (FP32): the module arrays, dx, dz_if and the surface slopes
 * (init_module_vars, wind_iterative.F90:86-91, 2350-2409), the coefficient
 * assembly (initialize_coefs / update_coefs_gpu :2424-2603), alpha, the rhs
 * product (:1525) and the correction with lambda rounded to FP32 (:1770-2062).
 * The Krylov solve stays FP64 as in the reference. Set it before set_metrics. */
void orc_set_faithful(orc_t *o, int on)
{
    o->faithful = on ? 1 : 0;
    if (o->faithful) o->dx = f32r(o->dx);
}
void orc_get_lambda(const orc_t *o, double *out) { memcpy(out, o->x, o->N * sizeof(double)); }
void orc_set_lambda(orc_t *o, const double *in) { memcpy(o->x, in, o->N * sizeof(double)); }
void orc_get_rows(const orc_t *o, int which, double *out)
{
    const double *src = which == 0 ? o->rhs : which == 1 ? o->dinv : which == 2 ? o->r : o->x;
    memcpy(out, src, o->N * sizeof(double));
}
void orc_get_coef(const orc_t *o, int q, double *out)
{
    memcpy(out, o->coef[q], (long)o->nx * o->nz * o->ny * sizeof(double));
}
void orc_set_coef(orc_t *o, int q, const double *in)
{
    memcpy(o->coef[q], in, (long)o->nx * o->nz * o->ny * sizeof(double));
}
void orc_set_probed(orc_t *o, int probed) { o->probed = probed; o->structure_uploaded = 1; }
void orc_get_planes(const orc_t *o, double *dz_if, double *dzdx_surf, double *dzdy_surf)
{
    memcpy(dz_if, o->dz_if, (long)o->nx * (o->nz + 1) * o->ny * sizeof(double));
    memcpy(dzdx_surf, o->dzdx_surf, (long)o->nx * o->ny * sizeof(double));
    memcpy(dzdy_surf, o->dzdy_surf, (long)o->nx * o->ny * sizeof(double));
}

/* ------------------------------------------------------------ correction
 * calc_updated_winds, wind_iterative.F90:1742-2071, single rank, FP64 lambda
 * (the reference rounds lambda to FP32 at :1770-1778; the FP64 plain form
 * keeps it in FP64). rho (nx,nz,ny) may be NULL (adv_den false or rho = 1);
 * map factors may be NULL (1.0).
 */
/* FP32-faithful calc_updated_winds: lambda rounded to `real` (lambda_3d,
 * :1770-1778) and every array and operation of :1804-2062 in FP32, as the
 * reference computes it; u, v, w are `real` fields, so they are rounded to
 * FP32 on entry and hold FP32 values on exit. */
static void correct_f32(const orc_t *o, double *u, double *v, double *w, const double *rho_in,
                        const double *mf_mx_u, const double *mf_my_v, int adv_den)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    const float dx = (float)o->dx;
    const int mx = o->mx, mz = o->mz, my = o->my;
    long n3 = (long)nx * nz * ny, nu = (long)(nx + 1) * nz * ny, nv = (long)nx * nz * (ny + 1);
    float *lam = (float *)calloc((size_t)o->N, sizeof(float));
    float *rho = (float *)calloc((size_t)n3, sizeof(float)), *rho_w = (float *)calloc((size_t)n3, sizeof(float));
    float *rho_u = (float *)calloc((size_t)nu, sizeof(float)), *rho_v = (float *)calloc((size_t)nv, sizeof(float));
    for (long n = 0; n < o->N; ++n) lam[n] = (float)o->x[n];
    for (long n = 0; n < n3; ++n) rho[n] = (adv_den && rho_in) ? (float)rho_in[n] : 1.0f;
#define LF(i, k, j) lam[(i) + (long)mx * ((k) + (long)mz * (j))]
#define DZF(k) ((float)DZ(o, 1, (k), 1))
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k) {
            for (int i = 2; i <= nx; ++i)
                U3(o, rho_u, i, k, j) = 0.5f * (M3(o, rho, i, k, j) + M3(o, rho, i - 1, k, j));
            U3(o, rho_u, nx + 1, k, j) = M3(o, rho, nx, k, j);
            U3(o, rho_u, 1, k, j) = M3(o, rho, 1, k, j);
        }
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
            for (int j = 2; j <= ny; ++j)
                V3(o, rho_v, i, k, j) = 0.5f * (M3(o, rho, i, k, j) + M3(o, rho, i, k, j - 1));
            V3(o, rho_v, i, k, 1) = M3(o, rho, i, k, 1);
            V3(o, rho_v, i, k, ny + 1) = M3(o, rho, i, k, ny);
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            for (int k = 1; k <= nz - 1; ++k) {
                float dzk = (float)M3(o, o->adv_dz, i, k, j), dzk1 = (float)M3(o, o->adv_dz, i, k + 1, j);
                M3(o, rho_w, i, k, j) = (M3(o, rho, i, k, j) * dzk1 + M3(o, rho, i, k + 1, j) * dzk) / (dzk + dzk1);
            }
            M3(o, rho_w, i, nz, j) = M3(o, rho, i, nz, j);
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx + 1; ++i) {
            float ut[nz + 2];
            for (int k = 0; k <= nz + 1; ++k) ut[k] = (LF(i, k, j) + LF(i - 1, k, j)) / 2.0f;
            for (int k = 1; k <= nz; ++k) {
                float dl;
                if (k == 1) dl = (ut[2] - ut[1]) / DZF(2);
                else if (k == nz) dl = (ut[nz] - ut[nz - 1]) / DZF(nz);
                else dl = (ut[k + 1] - ut[k - 1]) / (DZF(k + 1) + DZF(k));
                float mf = mf_mx_u ? (float)U2(o, mf_mx_u, i, j) : 1.0f;
                float grad = (LF(i, k, j) - LF(i - 1, k, j)) / dx
                           - ((float)U3(o, o->dzdx_u, i, k, j) * dl) / (float)U3(o, o->jaco_u, i, k, j);
                float uu = (float)U3(o, u, i, k, j);
                U3(o, u, i, k, j) = (double)(uu + ((0.5f * mf) * grad) / U3(o, rho_u, i, k, j));
            }
        }
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx; ++i) {
            float vt[nz + 2];
            for (int k = 0; k <= nz + 1; ++k) vt[k] = (LF(i, k, j) + LF(i, k, j - 1)) / 2.0f;
            for (int k = 1; k <= nz; ++k) {
                float dl;
                if (k == 1) dl = (vt[2] - vt[1]) / DZF(2);
                else if (k == nz) dl = (vt[nz] - vt[nz - 1]) / DZF(nz);
                else dl = (vt[k + 1] - vt[k - 1]) / (DZF(k + 1) + DZF(k));
                float mf = mf_my_v ? (float)V2(o, mf_my_v, i, j) : 1.0f;
                float grad = (LF(i, k, j) - LF(i, k, j - 1)) / dx
                           - ((float)V3(o, o->dzdy_v, i, k, j) * dl) / (float)V3(o, o->jaco_v, i, k, j);
                float vv = (float)V3(o, v, i, k, j);
                V3(o, v, i, k, j) = (double)(vv + ((0.5f * mf) * grad) / V3(o, rho_v, i, k, j));
            }
        }
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                float ww = (float)M3(o, w, i, k, j);
                if (k < nz) {
                    float al = (float)M3(o, o->alpha, i, k, j);
                    ww = ww + ((((0.5f * (al * al)) * (LF(i, k + 1, j) - LF(i, k, j))) / (float)DZ(o, i, k + 1, j))
                               / (float)M3(o, o->jaco_w, i, k, j)) / M3(o, rho_w, i, k, j);
                }
                M3(o, w, i, k, j) = (double)ww;
            }
#undef LF
#undef DZF
    (void)my;
    free(lam); free(rho); free(rho_w); free(rho_u); free(rho_v);
}

void orc_correct(const orc_t *o, double *u, double *v, double *w, const double *rho_in,
                 const double *mf_mx_u, const double *mf_my_v, int adv_den)
{
    if (o->faithful) { correct_f32(o, u, v, w, rho_in, mf_mx_u, mf_my_v, adv_den); return; }
    int nx = o->nx, ny = o->ny, nz = o->nz;
    double dx = o->dx;
    const double *lam = o->x;
    long n3 = (long)nx * nz * ny;
    double *rho = zalloc(n3), *rho_u = zalloc((long)(nx + 1) * nz * ny);
    double *rho_v = zalloc((long)nx * nz * (ny + 1)), *rho_w = zalloc(n3);
    for (long n = 0; n < n3; ++n) rho[n] = (adv_den && rho_in) ? rho_in[n] : 1.0;
    /* rho_u, :1822-1869 (both domain edges on one rank) */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k) {
            for (int i = 2; i <= nx; ++i)
                U3(o, rho_u, i, k, j) = 0.5 * (M3(o, rho, i, k, j) + M3(o, rho, i - 1, k, j));
            U3(o, rho_u, nx + 1, k, j) = M3(o, rho, nx, k, j);
            U3(o, rho_u, 1, k, j) = M3(o, rho, 1, k, j);
        }
    /* rho_v, :1871-1938 */
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
            for (int j = 2; j <= ny; ++j)
                V3(o, rho_v, i, k, j) = 0.5 * (M3(o, rho, i, k, j) + M3(o, rho, i, k, j - 1));
            V3(o, rho_v, i, k, 1) = M3(o, rho, i, k, 1);
            V3(o, rho_v, i, k, ny + 1) = M3(o, rho, i, k, ny);
        }
    /* rho_w, :1940-1955 */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            for (int k = 1; k <= nz - 1; ++k) {
                double dzk = M3(o, o->adv_dz, i, k, j), dzk1 = M3(o, o->adv_dz, i, k + 1, j);
                M3(o, rho_w, i, k, j) = (M3(o, rho, i, k, j) * dzk1 + M3(o, rho, i, k + 1, j) * dzk) / (dzk + dzk1);
            }
            M3(o, rho_w, i, nz, j) = M3(o, rho, i, nz, j);
        }
    /* u, :1957-2028 */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx + 1; ++i) {
            double ut[nz + 2];
            for (int k = 0; k <= nz + 1; ++k) ut[k] = (R(o, lam, i, k, j) + R(o, lam, i - 1, k, j)) / 2;
            for (int k = 1; k <= nz; ++k) {
                double dl;
                if (k == 1) dl = (ut[2] - ut[1]) / DZ(o, 1, 2, 1);
                else if (k == nz) dl = (ut[nz] - ut[nz - 1]) / DZ(o, 1, nz, 1);
                else dl = (ut[k + 1] - ut[k - 1]) / (DZ(o, 1, k + 1, 1) + DZ(o, 1, k, 1));
                double mf = mf_mx_u ? U2(o, mf_mx_u, i, j) : 1.0;
                double grad = (R(o, lam, i, k, j) - R(o, lam, i - 1, k, j)) / dx
                            - (U3(o, o->dzdx_u, i, k, j) * dl) / U3(o, o->jaco_u, i, k, j);
                U3(o, u, i, k, j) = U3(o, u, i, k, j) + ((0.5 * mf) * grad) / U3(o, rho_u, i, k, j);
            }
        }
    /* v, :1964-2040 */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx; ++i) {
            double vt[nz + 2];
            for (int k = 0; k <= nz + 1; ++k) vt[k] = (R(o, lam, i, k, j) + R(o, lam, i, k, j - 1)) / 2;
            for (int k = 1; k <= nz; ++k) {
                double dl;
                if (k == 1) dl = (vt[2] - vt[1]) / DZ(o, 1, 2, 1);
                else if (k == nz) dl = (vt[nz] - vt[nz - 1]) / DZ(o, 1, nz, 1);
                else dl = (vt[k + 1] - vt[k - 1]) / (DZ(o, 1, k + 1, 1) + DZ(o, 1, k, 1));
                double mf = mf_my_v ? V2(o, mf_my_v, i, j) : 1.0;
                double grad = (R(o, lam, i, k, j) - R(o, lam, i, k, j - 1)) / dx
                            - (V3(o, o->dzdy_v, i, k, j) * dl) / V3(o, o->jaco_v, i, k, j);
                V3(o, v, i, k, j) = V3(o, v, i, k, j) + ((0.5 * mf) * grad) / V3(o, rho_v, i, k, j);
            }
        }
    /* w (grid-relative, interface above cell k), :2052-2062 */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz - 1; ++k)
            for (int i = 1; i <= nx; ++i) {
                double al = M3(o, o->alpha, i, k, j);
                M3(o, w, i, k, j) = M3(o, w, i, k, j)
                    + ((((0.5 * (al * al)) * (R(o, lam, i, k + 1, j) - R(o, lam, i, k, j)))
                        / DZ(o, i, k + 1, j)) / M3(o, o->jaco_w, i, k, j)) / M3(o, rho_w, i, k, j);
            }
    free(rho); free(rho_u); free(rho_v); free(rho_w);
}

/* ------------------------------------------------------------ divergence
 * calc_divergence, wind.F90:250-523 (operator D), FP64, single rank
 * (ims = 1, ime = nx, ...). rho may be NULL (= 1). Map factors may be NULL.
 * mf_my_u (nx+1, ny), mf_mx_v (nx, ny+1), mf_mxy (nx, ny).
 */
void orc_divergence(const orc_t *o, const double *u, const double *v, const double *w,
                    const double *rho_in, const double *mf_my_u, const double *mf_mx_v,
                    const double *mf_mxy, int adv_den, int horz_only, int use_dqdt, double *div)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    double dx = o->dx;
    long n3 = (long)nx * nz * ny;
    double *rho = zalloc(n3);
    for (long n = 0; n < n3; ++n) rho[n] = rho_in ? rho_in[n] : 1.0;
    double *um = zalloc((long)(nx + 1) * nz * ny), *vm = zalloc((long)nx * nz * (ny + 1));
    double *wm = zalloc(n3), *rho_i = zalloc(n3);
    const double *jaco_u = o->jaco_u, *jaco_v = o->jaco_v, *jaco_w = o->jaco_w;
    if (adv_den) {
        /* :294-349 (the dqdt and data branches apply the same formulas) */
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k) {
                for (int i = 2; i <= nx; ++i)
                    U3(o, um, i, k, j) = ((U3(o, u, i, k, j) * U3(o, jaco_u, i, k, j))
                                          * (M3(o, rho, i - 1, k, j) + M3(o, rho, i, k, j))) / 2;
                U3(o, um, 1, k, j) = (U3(o, u, 1, k, j) * U3(o, jaco_u, 1, k, j))
                                     * (1.5 * M3(o, rho, 1, k, j) - 0.5 * M3(o, rho, 2, k, j));
                U3(o, um, nx + 1, k, j) = (U3(o, u, nx + 1, k, j) * U3(o, jaco_u, nx + 1, k, j))
                                          * (1.5 * M3(o, rho, nx, k, j) - 0.5 * M3(o, rho, nx - 1, k, j));
            }
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                for (int j = 2; j <= ny; ++j)
                    V3(o, vm, i, k, j) = ((V3(o, v, i, k, j) * V3(o, jaco_v, i, k, j))
                                          * (M3(o, rho, i, k, j - 1) + M3(o, rho, i, k, j))) / 2;
                V3(o, vm, i, k, 1) = (V3(o, v, i, k, 1) * V3(o, jaco_v, i, k, 1))
                                     * (1.5 * M3(o, rho, i, k, 1) - 0.5 * M3(o, rho, i, k, 2));
                V3(o, vm, i, k, ny + 1) = (V3(o, v, i, k, ny + 1) * V3(o, jaco_v, i, k, ny + 1))
                                          * (1.5 * M3(o, rho, i, k, ny) - 0.5 * M3(o, rho, i, k, ny - 1));
            }
        /* rho on w levels, :353-364 */
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz - 1; ++k)
                for (int i = 1; i <= nx; ++i) {
                    double dzk = M3(o, o->adv_dz, i, k, j), dzk1 = M3(o, o->adv_dz, i, k + 1, j);
                    M3(o, rho_i, i, k, j) = (M3(o, rho, i, k, j) * dzk1 + M3(o, rho, i, k + 1, j) * dzk) / (dzk + dzk1);
                }
    } else if (use_dqdt) {
        /* :366-383: the dqdt winds are used without metric factors */
        memcpy(um, u, (long)(nx + 1) * nz * ny * sizeof(double));
        memcpy(vm, v, (long)nx * nz * (ny + 1) * sizeof(double));
    } else {
        /* :384-400 */
        for (long n = 0; n < (long)(nx + 1) * nz * ny; ++n) um[n] = u[n] * jaco_u[n];
        for (long n = 0; n < (long)nx * nz * (ny + 1); ++n) vm[n] = v[n] * jaco_v[n];
    }
    /* horizontal part, :411-420 */
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                double m0 = mf_my_u ? U2(o, mf_my_u, i, j) : 1.0, m1 = mf_my_u ? U2(o, mf_my_u, i + 1, j) : 1.0;
                double m2 = mf_mx_v ? V2(o, mf_mx_v, i, j) : 1.0, m3 = mf_mx_v ? V2(o, mf_mx_v, i, j + 1) : 1.0;
                double mxy = mf_mxy ? S2(o, mf_mxy, i, j) : 1.0;
                M3(o, div, i, k, j) = (mxy * (((U3(o, um, i + 1, k, j) / m1 - U3(o, um, i, k, j) / m0)
                                               + V3(o, vm, i, k, j + 1) / m3) - V3(o, vm, i, k, j) / m2)) / dx;
            }
    if (!horz_only) {
        if (adv_den) {
            /* :424-458 */
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    for (int k = 1; k <= nz - 1; ++k)
                        M3(o, wm, i, k, j) = (M3(o, w, i, k, j) * M3(o, jaco_w, i, k, j)) * M3(o, rho_i, i, k, j);
                    M3(o, wm, i, nz, j) = (M3(o, w, i, nz, j) * M3(o, jaco_w, i, nz, j)) * M3(o, rho, i, nz, j);
                }
        } else {
            /* :460-480 */
            for (long n = 0; n < n3; ++n) wm[n] = w[n] * jaco_w[n];
        }
        /* :483-497 */
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k)
                for (int i = 1; i <= nx; ++i) {
                    double d = M3(o, div, i, k, j);
                    if (k == 1) d = d + M3(o, wm, i, k, j) / M3(o, o->adv_dz, i, k, j);
                    else d = d + (M3(o, wm, i, k, j) - M3(o, wm, i, k - 1, j)) / M3(o, o->adv_dz, i, k, j);
                    M3(o, div, i, k, j) = d / M3(o, o->jaco, i, k, j);
                }
    }
    free(rho); free(um); free(vm); free(wm); free(rho_i);
}

/* ------------------------------------------------------------ probing
 * probe_lambda_pattern / probe_record / probe_finalize,
 * wind_iterative.F90:1096-1248, orchestrated as in
 * calibrate_projection_operator, wind.F90:1010-1099: for each of the 27
 * colours, lambda = indicator lattice, u=v=w=0, apply G (correction), apply
 * D (divergence), record 2*D(G(lambda)) into the 15 arrays. Returns the
 * off-stencil leakage max |T|.
 */
double orc_probe(orc_t *o, const double *rho, const double *mf_mx_u, const double *mf_my_v,
                 const double *mf_my_u, const double *mf_mx_v, const double *mf_mxy, int adv_den)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long n3 = (long)nx * nz * ny;
    double *u = zalloc((long)(nx + 1) * nz * ny), *v = zalloc((long)nx * nz * (ny + 1));
    double *w = zalloc(n3), *T = zalloc(n3);
    double max_leak = 0.0;
    double *xsave = zalloc(o->N);
    memcpy(xsave, o->x, o->N * sizeof(double));
    for (int ca = 0; ca < 3; ++ca)
        for (int cb = 0; cb < 3; ++cb)
            for (int cc = 0; cc < 3; ++cc) {
                for (int j = 0; j < o->my; ++j)
                    for (int k = 0; k < o->mz; ++k)
                        for (int i = 0; i < o->mx; ++i) {
                            int on = i >= 1 && i <= o->mx - 2 && k >= 1 && k <= o->mz - 2 && j >= 1
                                  && j <= o->my - 2 && i % 3 == ca && k % 3 == cb && j % 3 == cc;
                            R(o, o->x, i, k, j) = on ? 1.0 : 0.0;
                        }
                memset(u, 0, (long)(nx + 1) * nz * ny * sizeof(double));
                memset(v, 0, (long)nx * nz * (ny + 1) * sizeof(double));
                memset(w, 0, n3 * sizeof(double));
                orc_correct(o, u, v, w, rho, mf_mx_u, mf_my_v, adv_den);
                orc_divergence(o, u, v, w, adv_den ? rho : NULL, mf_my_u, mf_mx_v, mf_mxy, adv_den, 0, 1, T);
                for (int j = 1; j <= ny; ++j)
                    for (int k = 1; k <= nz; ++k)
                        for (int i = 1; i <= nx; ++i) {
                            int di = (ca - i % 3 + 3) % 3; if (di == 2) di = -1;
                            int dk = (cb - k % 3 + 3) % 3; if (dk == 2) dk = -1;
                            int dj = (cc - j % 3 + 3) % 3; if (dj == 2) dj = -1;
                            double tv = 2.0 * M3(o, T, i, k, j);
                            int q = -1;
                            if (dk == 0) {
                                if (di == 0 && dj == 0) q = CA; else if (di == 1 && dj == 0) q = CD;
                                else if (di == -1 && dj == 0) q = CE; else if (di == 0 && dj == 1) q = CF;
                                else if (di == 0 && dj == -1) q = CG;
                            } else if (dk == 1) {
                                if (di == 0 && dj == 0) q = CB; else if (di == 1 && dj == 0) q = CH;
                                else if (di == -1 && dj == 0) q = CI; else if (di == 0 && dj == 1) q = CL;
                                else if (di == 0 && dj == -1) q = CM;
                            } else {
                                if (di == 0 && dj == 0) q = CC; else if (di == 1 && dj == 0) q = CJ;
                                else if (di == -1 && dj == 0) q = CK; else if (di == 0 && dj == 1) q = CN;
                                else if (di == 0 && dj == -1) q = CO;
                            }
                            if (q >= 0) M3(o, o->coef[q], i, k, j) = tv;
                            else max_leak = fmax(max_leak, fabs(tv));
                        }
            }
    o->probed = 1;
    o->structure_uploaded = 1;
    memset(o->x, 0, o->N * sizeof(double));   /* probe_finalize resets x_sol, :1241 */
    free(xsave); free(u); free(v); free(w); free(T);
    return max_leak;
}

/* ------------------------------------------------------------------ the caller's steps
 * balance_uvw (wind.F90:123-191): the horizontal divergence (calc_divergence
 * with horz_only, :250-523) of (u, v), density-weighted whatever adv_den is
 * (balance_uvw leaves advect_density at its default, true, :159/:164), then
 * calc_w (:193-234): w from the
 * bottom up, w(k) = (w(k-1) rho_i(k-1) jaco_w(k-1) - div(k) dz(k)) /
 * (jaco_w(k) rho_i(k)) with rho_i the dz-weighted interface density (rho at
 * the top), or without rho_i when not adv_den; dz = advection_dz. */
void orc_balance_uvw(const orc_t *o, const double *u, const double *v, double *w, const double *rho_in,
                     int adv_den, int use_dqdt)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long n3 = (long)nx * nz * ny;
    double *div = zalloc(n3);
    orc_divergence(o, u, v, w, rho_in, NULL, NULL, NULL, 1, 1, use_dqdt, div);
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            double wprev = 0.0, riprev = 0.0;
            for (int k = 1; k <= nz; ++k) {
                double rk = rho_in ? M3(o, rho_in, i, k, j) : 1.0;
                double dz = M3(o, o->adv_dz, i, k, j), jw = M3(o, o->jaco_w, i, k, j), dv = M3(o, div, i, k, j);
                double wk;
                if (adv_den) {
                    double ri = rk;
                    if (k < nz) {
                        double rk1 = rho_in ? M3(o, rho_in, i, k + 1, j) : 1.0;
                        double dz1 = M3(o, o->adv_dz, i, k + 1, j);
                        ri = (rk * dz1 + rk1 * dz) / (dz + dz1);
                    }
                    if (k == 1) wk = 0.0 - (dv * dz) / (jw * ri);
                    else wk = ((wprev * riprev) * M3(o, o->jaco_w, i, k - 1, j) - dv * dz) / (jw * ri);
                    riprev = ri;
                } else {
                    if (k == 1) wk = (0.0 - dv * dz) / jw;
                    else wk = (wprev * M3(o, o->jaco_w, i, k - 1, j) - dv * dz) / jw;
                }
                M3(o, w, i, k, j) = wk;
                wprev = wk;
            }
        }
    free(div);
}

/* calc_w_real (wind.F90:1175-1227): (u dzdx_u at i and i+1)/2 + (v dzdy_v at
 * j and j+1)/2 + (w_ext(k) + w_ext(k+1))/2, w_ext(1) = 0, w_ext(k+1) =
 * w_grid(k) jaco_w(k); a null w_grid is zero. */
static double w_real_at(const orc_t *o, const double *u, const double *v, const double *wg, int i, int k, int j)
{
    double uw = (U3(o, u, i, k, j) * U3(o, o->dzdx_u, i, k, j) + U3(o, u, i + 1, k, j) * U3(o, o->dzdx_u, i + 1, k, j)) * 0.5;
    double vw = (V3(o, v, i, k, j) * V3(o, o->dzdy_v, i, k, j) + V3(o, v, i, k, j + 1) * V3(o, o->dzdy_v, i, k, j + 1)) * 0.5;
    double lo = 0.0, hi = 0.0;
    if (wg) {
        if (k > 1) lo = M3(o, wg, i, k - 1, j) * M3(o, o->jaco_w, i, k - 1, j);
        hi = M3(o, wg, i, k, j) * M3(o, o->jaco_w, i, k, j);
    }
    return (uw + vw) + (lo + hi) * 0.5;
}

void orc_w_real(const orc_t *o, const double *u, const double *v, const double *w_grid, double *out)
{
    for (int j = 1; j <= o->ny; ++j)
        for (int k = 1; k <= o->nz; ++k)
            for (int i = 1; i <= o->nx; ++i) M3(o, out, i, k, j) = w_real_at(o, u, v, w_grid, i, k, j);
}

/* calc_idealized_wgrid (wind.F90:940-998): t = w_real - calc_w_real(u, v, 0),
 * restaggered to the level interfaces (advection_dz weights), divided by
 * jaco_w; the top level is 0. */
void orc_idealized_wgrid(const orc_t *o, const double *u, const double *v, const double *w_real, double *w)
{
    int nz = o->nz;
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            for (int k = 1; k < nz; ++k) {
                double t0 = M3(o, w_real, i, k, j) - w_real_at(o, u, v, NULL, i, k, j);
                double t1 = M3(o, w_real, i, k + 1, j) - w_real_at(o, u, v, NULL, i, k + 1, j);
                double a0 = M3(o, o->adv_dz, i, k, j), a1 = M3(o, o->adv_dz, i, k + 1, j);
                double st = (t0 * a1 + t1 * a0) / (a0 + a1);
                M3(o, w, i, k, j) = st / M3(o, o->jaco_w, i, k, j);
            }
            M3(o, w, i, nz, j) = 0.0;
        }
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
