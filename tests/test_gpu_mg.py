"""The multigrid inner preconditioner (MIXED_MG) through the C-ABI.

Its coarse operators are checked against the oracle's assembly at the
coarse size and spacing from metrics averaged by oracle/mg_oracle.py; its
solves against the oracle's FP64 lambda (1e-6 relative L2, the
mixed-precision bar).
"""
import numpy as np
import pytest

import mg_oracle as mo  # oracle/mg_oracle.py
from conftest import oracle_for, small_case

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _setup(name="fractal_alpha"):
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = small_case(name)
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    return c, o, div, s


def _level_vs_oracle(s, level, metrics, alpha, nz, dx):
    """Coefficients B, C, H, I, L, M of solver level `level` against the
    oracle's assembly at that level's size and dx from the averaged metrics."""
    from oracle import Oracle
    (nx, ny, mz), got = s.mg_level(level)
    assert mz == nz + 2
    o = Oracle(nx, ny, nz, dx)
    o.set_metrics(dict(metrics, terrain=metrics["terrain"]))
    o.set_alpha(alpha)
    o.assemble()
    for q, name in enumerate("BCHILM"):
        want = o.coef(name)                                   # (ny, nz, nx)
        g = got[q][1:-1, 1:-1, 1:-1].transpose(1, 0, 2)      # rows [k][j][i] -> (ny, nz, nx)
        scale = np.abs(want).max()
        assert np.allclose(g, want, rtol=2e-6, atol=2e-7 * scale), name
    o.close()
    return nx, ny


def test_coarse_levels_rediscretise_the_operator():
    """Level 1 and 2 hold the analytic operator assembled at 2 dx and 4 dx from
    the metrics and alpha averaged over 2 x 2 columns (odd edge: ny = 23)."""
    from hicar_amd import _native as nat
    c, o, div, s = _setup()
    s.set_mg_galerkin(0)  # every coarse level rediscretised
    info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-6, start=1, max_iters=2000)
    assert info["status"] == 0
    keys = mo.CELL + mo.U_FACE + mo.V_FACE + ("terrain",)
    m0 = {k: c.metrics[k] for k in keys}
    m1, a1 = mo.coarsen_metrics(m0, c.alpha)
    nx1, ny1 = _level_vs_oracle(s, 1, m1, a1, c.nz, 2 * c.dx)
    assert (nx1, ny1) == (mo.coarse_size(c.nx), mo.coarse_size(c.ny))
    m2, a2 = mo.coarsen_metrics(m1, a1)
    _level_vs_oracle(s, 2, m2, a2, c.nz, 4 * c.dx)


def test_mixed_mg_matches_oracle():
    from hicar_amd import _native as nat
    c, o, div, s = _setup()
    o.solve(c.alpha, div, max_iters=5000, tol_rel=1e-12)
    line = s.solve(c.alpha, div, precision=nat.MIXED_LINE, tol_rel=1e-8, start=1, max_iters=20000)
    s.set_lambda(np.zeros(s.rows_shape))
    info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=20000)
    assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
    assert _rel(s.get_lambda(), o.get_lambda()) <= 1e-6
    assert info["iters"] < line["iters"]
    # the same answer from a graph-replayed and an eagerly launched solve
    s.set_lambda(np.zeros(s.rows_shape))
    eager = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=20000, use_graph=0)
    assert eager["iters"] == info["iters"]


def test_mixed_mg_on_the_probed_operator():
    from hicar_amd import _native as nat
    c, o, div, s = _setup("hill")
    s.probe_operator(c.alpha, density=c.rho, adv_den=True)
    ref = s.solve(c.alpha, div, precision=nat.PLAIN_F64, tol_rel=1e-12, start=1, max_iters=20000)
    assert ref["status"] == 0
    lam = s.get_lambda()
    s.set_lambda(np.zeros(s.rows_shape))
    info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=20000)
    assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
    assert _rel(s.get_lambda(), lam) <= 1e-6


@pytest.mark.parametrize("name", ["fractal_alpha", "hill"])
def test_fused_coarse_smoothing_matches_the_two_kernel_step(name):
    """The one-launch smoothing step of the small levels (a wavefront per
    column, the Thomas recurrences as lane scans) against the residual + line
    solve kernels, on every coarse level, from a zero and a random iterate; the
    lateral identity rows give y + omega (b - y) exactly. The same two kernels
    on the level's bf16 rows and factors (the form of the larger levels) stay
    within bf16 rounding (2^-9 per coefficient) of the FP32 step."""
    from hicar_amd import _native as nat
    c, o, div, s = _setup(name)
    s.set_mg_galerkin(0)  # the kernel forms of the rediscretised levels
    s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-6, start=1, max_iters=2000)
    rng = np.random.default_rng(21)
    level = 1
    while True:
        try:
            (nx, ny, mz), _ = s.mg_level(level)
        except RuntimeError:
            break
        shape = (mz, ny + 2, nx + 2)
        b = rng.standard_normal(shape).astype(np.float32)
        y = rng.standard_normal(shape).astype(np.float32)
        for yy in (None, y):
            fused = s.mg_smooth(level, True, b, yy)
            split = s.mg_smooth(level, False, b, yy)
            assert np.all(np.isfinite(fused))
            assert _rel(fused, split) <= 1e-5, (level, yy is None)
            half = s.mg_smooth(level, 2, b, yy)
            assert 0 < _rel(half, split) <= 3e-2, (level, yy is None, _rel(half, split))
            if yy is not None:
                # the one-pass step (residual into the line solve in registers)
                # computes its two kernels' values exactly
                assert np.array_equal(s.mg_smooth(level, 3, b, yy), half), level
                assert np.array_equal(s.mg_smooth(level, 4, b, yy), split), level
            ring = np.zeros(shape, dtype=bool)
            ring[:, 0, :] = ring[:, -1, :] = ring[:, :, 0] = ring[:, :, -1] = True
            y0 = np.zeros(shape, np.float32) if yy is None else yy
            want = y0 + np.float32(0.9) * (b - y0)
            np.testing.assert_allclose(fused[ring], want[ring], rtol=1e-6, atol=1e-6)
        level += 1
    assert level > 2


@pytest.mark.parametrize("precision", ["MIXED_LINE", "MIXED_MG"])
def test_tall_columns_take_the_general_line_kernels(precision):
    """nz = 64 (66 rows per column): the line solves leave the register-resident
    kernels (at most 64 rows) for the batched march with its forward results in
    memory, and the small multigrid levels keep the two-kernel smoothing step;
    the answer is the oracle's as for short columns."""
    from hicar_amd import _native as nat
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = S.gaussian_hill_case(nx=14, ny=12, nz=64, dx=250.0, seed=2)
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    o.solve(c.alpha, div, max_iters=20000, tol_rel=1e-12)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    info = s.solve(c.alpha, div, precision=getattr(nat, precision), tol_rel=1e-8, start=1, max_iters=20000)
    assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
    assert _rel(s.get_lambda(), o.get_lambda()) <= 1e-6
    if precision == "MIXED_MG":
        (nx, ny, mz), _ = s.mg_level(1)
        assert mz == 66
        rng = np.random.default_rng(3)
        b = rng.standard_normal((mz, ny + 2, nx + 2)).astype(np.float32)
        y = rng.standard_normal((mz, ny + 2, nx + 2)).astype(np.float32)
        # no one-launch form above 64 rows: both requests take the two-kernel step
        assert np.array_equal(s.mg_smooth(1, True, b, y), s.mg_smooth(1, False, b, y))


def test_mixed_mg_with_f32_fields():
    """HICAR's `real` (FP32) caller fields: the multigrid solve of the promoted
    inputs reaches the oracle's lambda on the same FP32-rounded inputs."""
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    from oracle import Oracle
    _dev()
    c = small_case("fractal_alpha")
    m32 = {k: (np.asarray(x, dtype=np.float32) if isinstance(x, np.ndarray) else x) for k, x in c.metrics.items()}
    m64 = {k: (x.astype(np.float64) if isinstance(x, np.ndarray) else x) for k, x in m32.items()}
    s = WindSolver(c.nx, c.ny, c.nz, c.dx, dtype=np.float32)
    s.set_metrics(m32)
    o = Oracle(c.nx, c.ny, c.nz, c.dx)
    o.set_metrics(m64)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True).astype(np.float32)
    a32 = c.alpha.astype(np.float32)
    o.solve(a32.astype(np.float64), div.astype(np.float64), tol_rel=1e-12, max_iters=5000)
    info = s.solve(a32, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=20000, inner_tol=0.1)
    assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
    assert _rel(s.get_lambda(), o.get_lambda()) <= 1e-6


def test_fused_coarse_cycle_matches_the_separate_kernels():
    """The V-cycle below each coarse level with the one-launch descent
    (residual, restriction, first step) and return (prolongation, first
    post-step) on the levels of one wavefront per column equals the cycle of
    separate kernels: to FP32 rounding from the first such level down, and to
    the bf16 rounding of the separate residual (which reads the level's bf16
    rows) from a larger level. 200 x 180 fine columns: levels of 100 x 90
    (separate kernels), 50 x 45, 25 x 23, ... (one wavefront per column)."""
    from hicar_amd import _native as nat
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = S.fractal_case(nx=200, ny=180, nz=12, dx=50.0, relief=600.0, seed=8, lowest=10.0, top=6000.0)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    div = s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True)
    info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=2000, inner_tol=0.1)
    assert info["status"] == 0
    rng = np.random.default_rng(5)
    level, seen_wave = 1, False
    while True:
        try:
            (nx, ny, mz), _ = s.mg_level(level)
        except Exception:
            break
        b = rng.standard_normal((mz, ny + 2, nx + 2)).astype(np.float32)
        b[:, 0, :] = b[:, -1, :] = 0.0
        b[:, :, 0] = b[:, :, -1] = 0.0
        fused = s.mg_cycle(level, b, fused=True)
        sep = s.mg_cycle(level, b, fused=False)
        wave = (nx + 2) * (ny + 2) <= 4608
        seen_wave |= wave
        err = _rel(fused, sep)
        print(f"level {level} ({nx}x{ny}): fused vs separate {err:.2e}")
        assert np.all(np.isfinite(fused)) and np.abs(fused).max() > 0
        assert err <= (1e-5 if wave else 2e-2), (level, err)
        level += 1
    assert seen_wave and level > 6
    s.close()


def test_persistent_coarse_cycle_equals_the_step_launches():
    """The cycle from each coarse level down with the column-major levels as
    one persistent launch (grid-wide barriers between its steps) against the
    same cycle launched step by step: bitwise equal, since every column
    computes the same values in the same order; twice in a row (the barrier
    state returns to its start after each launch). Then MIXED_MG solves with
    and without it: the same iterations and a bitwise equal lambda."""
    from hicar_amd import _native as nat
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = S.fractal_case(nx=200, ny=180, nz=12, dx=50.0, relief=600.0, seed=8, lowest=10.0, top=6000.0)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_mg_galerkin(0)  # the persistent launch runs the rediscretised levels' steps
    s.set_metrics(c.metrics)
    div = s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True)
    runs = []
    for mode in (2, 1):
        s.set_mg_fusion(mode, True)
        s.set_lambda(np.zeros(s.rows_shape))
        info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=2000, inner_tol=0.1)
        assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
        runs.append((info, s.get_lambda()))
    (a, la), (b, lb) = runs
    print(f"persistent: {a['iters']} iterations, step launches: {b['iters']}")
    assert a["iters"] == b["iters"] and a["outer_iters"] == b["outer_iters"]
    assert np.array_equal(la, lb)
    s.set_mg_fusion(1, True)
    rng = np.random.default_rng(11)
    level, seen = 1, 0
    while True:
        try:
            (nx, ny, mz), _ = s.mg_level(level)
        except Exception:
            break
        b = rng.standard_normal((mz, ny + 2, nx + 2)).astype(np.float32)
        b[:, 0, :] = b[:, -1, :] = 0.0
        b[:, :, 0] = b[:, :, -1] = 0.0
        p1 = s.mg_cycle(level, b, fused=2)
        p2 = s.mg_cycle(level, b, fused=2)
        f = s.mg_cycle(level, b, fused=1)
        print(f"level {level} ({nx}x{ny}): persistent vs step launches max |diff| {np.abs(p1 - f).max():.3e}")
        assert np.all(np.isfinite(p1)) and np.abs(p1).max() > 0
        assert np.array_equal(p1, f), level
        assert np.array_equal(p1, p2), level
        seen += (nx + 2) * (ny + 2) <= 4608
        level += 1
    assert seen >= 3 and level > 6
    s.close()


@pytest.mark.parametrize("name", ["fractal_alpha", "hill"])
def test_onepass_smoothing_solve_equals_the_two_kernel_solve(name):
    """MIXED_MG with the one-pass smoothing steps (fine level and the larger
    coarse levels) against the same solve with the residual + line-solve
    kernels: the same iterations and a bitwise equal lambda, since the one-pass
    kernel rounds every value as the two kernels do."""
    from hicar_amd import _native as nat
    c, o, div, s = _setup(name)
    runs = []
    for onepass in (True, False):
        s.set_mg_fusion(1, onepass)
        s.set_lambda(np.zeros(s.rows_shape))
        s.profile(reset=True)
        info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=20000,
                       inner_tol=0.1, profile=1)
        prof = s.profile(reset=True)
        assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
        runs.append((info, s.get_lambda(), prof))
    (a, la, pa), (b, lb, pb) = runs
    assert pa.get("smooth_f32", (0, 0))[0] > 0 and pb.get("smooth_f32", (0, 0))[0] == 0
    assert a["iters"] == b["iters"] and a["outer_iters"] == b["outer_iters"]
    assert np.array_equal(la, lb)
    s.set_mg_fusion(1, True)


@pytest.mark.parametrize("inner_tol", [0.3, 0.01])
def test_fgmres_step_enqueued_ahead_keeps_the_solve(inner_tol):
    """MIXED_MG enqueues each direction's FGMRES step (FP64 matvec, Gram-Schmidt,
    Hessenberg column) right behind the first batch of its inner solve and
    repeats it when the inner solve had not ended there. Batches of 1 iteration
    (the default: the step is discarded whenever an inner solve needs more, as
    at inner_tol 0.01), of 8 and without graphs run the same iterations on the
    same values: equal iteration counts and a bitwise equal lambda."""
    from hicar_amd import _native as nat
    c, o, div, s = _setup("fractal_alpha")
    runs = []
    for ips, graph in ((1, 1), (8, 1), (1, 0)):
        s.set_lambda(np.zeros(s.rows_shape))
        info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=5000,
                       inner_tol=inner_tol, iters_per_sync=ips, use_graph=graph)
        assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
        runs.append((info, s.get_lambda()))
    (a, la), (b, lb), (d, ld) = runs
    if inner_tol < 0.1:
        assert a["iters"] > a["outer_iters"]  # inner solves of several iterations
    for info, lam in ((b, lb), (d, ld)):
        assert info["iters"] == a["iters"] and info["outer_iters"] == a["outer_iters"]
        assert np.array_equal(lam, la)
    s.close()


def test_profile_counts_only_the_launches_that_ran():
    """The profiling mode times every Krylov launch, but a launch enqueued
    after its solve has finished (the rest of a batch of iterations) returns at
    once: such launches are left out of the per-kind counts. Each BiCGStab
    iteration then counts one sigma and one omega matvec, and each V-cycle (two
    per iteration, zero start, fine V(2,2)) one residual + restriction, one
    coarse region and three one-pass smoothing steps; the plain FP64 solve one
    sigma sweep, one omega sweep and one x/r/p update per iteration."""
    from hicar_amd import _native as nat
    c, o, div, s = _setup("hill")
    s.set_lambda(np.zeros(s.rows_shape))
    s.profile(reset=True)
    info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=2000, inner_tol=0.3,
                   profile=1)
    prof = s.profile(reset=True)
    assert info["status"] == 0
    it = info["iters"]
    assert prof["mv_sigma_f32"][0] == it and prof["mv_omega_f32"][0] == it
    assert prof["resid_restrict_f32"][0] == 2 * it and prof["mg_coarse_f32"][0] == 2 * it
    assert prof["smooth_f32"][0] == 3 * 2 * it
    s.set_lambda(np.zeros(s.rows_shape))
    info = s.solve(c.alpha, div, precision=nat.PLAIN_F64, tol_rel=1e-8, start=1, max_iters=5000, profile=1,
                   iters_per_sync=64)
    prof = s.profile(reset=True)
    assert info["status"] == 0
    it = info["iters"]
    for kind in ("sweep_sigma_f64", "sweep_omega_f64", "update_xrp_f64", "update_s_f64"):
        assert prof[kind][0] == it, kind
    s.close()


# ----------------------------------------------------------- Galerkin levels
def _transfers(nxf, nyf, nxc, nyc, mz):
    """R (mean over a coarse column's 2 x 2 fine columns) and P (cell-centred
    bilinear, 9/16 3/16 3/16 1/16; ring columns carry nothing) as sparse
    matrices on rows [k][j][i], written from the definitions."""
    import scipy.sparse as sp
    mxf, myf, mxc, myc = nxf + 2, nyf + 2, nxc + 2, nyc + 2
    fidx = lambda i, k, j: (k * myf + j) * mxf + i  # noqa: E731
    cidx = lambda i, k, j: (k * myc + j) * mxc + i  # noqa: E731
    rr, rc, rv, pr, pc, pv = [], [], [], [], [], []
    for J in range(1, nyc + 1):
        for I in range(1, nxc + 1):
            fi = [i for i in (2 * I - 1, 2 * I) if i <= nxf]
            fj = [j for j in (2 * J - 1, 2 * J) if j <= nyf]
            for k in range(mz):
                for j in fj:
                    for i in fi:
                        rr.append(cidx(I, k, J)); rc.append(fidx(i, k, j)); rv.append(1.0 / (len(fi) * len(fj)))
    for j in range(1, nyf + 1):
        for i in range(1, nxf + 1):
            I, J = (i + 1) // 2, (j + 1) // 2
            In, Jn = (I - 1 if i % 2 else I + 1), (J - 1 if j % 2 else J + 1)
            for k in range(mz):
                for (a, b), w in (((I, J), 0.5625), ((In, J), 0.1875), ((I, Jn), 0.1875), ((In, Jn), 0.0625)):
                    if 1 <= a <= nxc and 1 <= b <= nyc:
                        pr.append(fidx(i, k, j)); pc.append(cidx(a, k, b)); pv.append(w)
    R = sp.csr_matrix((rv, (rr, rc)), shape=(mxc * myc * mz, mxf * myf * mz))
    P = sp.csr_matrix((pv, (pr, pc)), shape=(mxf * myf * mz, mxc * myc * mz))
    return R, P


def _stencil_to_matrix(q, nx, ny, mz):
    """The sparse matrix of 27 coefficient arrays [27, mz, ny+2, nx+2] (rows [k][j][i])."""
    import scipy.sparse as sp
    mx, my = nx + 2, ny + 2
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    rows, cols, vals = [], [], []
    for o in range(27):
        di, dj, dk = o % 3 - 1, (o // 3) % 3 - 1, o // 9 - 1
        v = q[o]
        ok = (v != 0) & (i + di >= 0) & (i + di < mx) & (j + dj >= 0) & (j + dj < my) & (k + dk >= 0) & (k + dk < mz)
        rows.append(((k * my + j) * mx + i)[ok])
        cols.append((((k + dk) * my + (j + dj)) * mx + (i + di))[ok])
        vals.append(v[ok].astype(np.float64))
    n = mx * my * mz
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def _galerkin(Af, nxf, nyf, nxc, nyc, mz):
    """R A P with the coarse ring rows identity and the ring columns dropped."""
    import scipy.sparse as sp
    R, P = _transfers(nxf, nyf, nxc, nyc, mz)
    A = (R @ Af @ P).tocsr()
    mx, my = nxc + 2, nyc + 2
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    ring = ((i == 0) | (i == mx - 1) | (j == 0) | (j == my - 1)).ravel()
    D = sp.diags((~ring).astype(float))
    return (D @ A @ D + sp.diags(ring.astype(float))).tocsr()


def _to_kji(A_oracle, nx, ny, mz):
    """A matrix on the oracle's rows i + mx (k + mz j) moved to rows [k][j][i]."""
    mx, my = nx + 2, ny + 2
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    perm = (i + mx * (k + mz * j)).ravel()  # [k][j][i] row -> oracle row
    return A_oracle[perm][:, perm].tocsr()


def _rowwise_close(got, want, rtol):
    """|got - want| <= rtol * max |want| of the row, entry by entry."""
    d = (got - want).tocsr()
    scale = abs(want).max(axis=1).toarray().ravel()
    err = abs(d).max(axis=1).toarray().ravel()
    return float(np.max(err / np.maximum(scale, 1e-300)))


def test_galerkin_levels_are_the_product_R_A_P():
    """From the second coarse level on, each level's operator is R A P of the
    finer level's: level 2 against the product of the oracle's level-1
    operator (assembled at 2 dx from the averaged metrics: the rediscretised
    first level), level 3 against the product of level 2's own coefficients,
    to the FP32 rounding of the stored coefficients; 27 offsets per row,
    identity rows on the ring. The transfers are written here from their
    definitions (sparse matrices), independently of the kernels."""
    from hicar_amd import _native as nat
    from oracle import Oracle
    from test_oracle import stencil_matrix
    c, o, div, s = _setup()
    info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-6, start=1, max_iters=2000)
    assert info["status"] == 0
    assert s.mg_stencil(1) is None and s.mg_stencil(2) is not None
    keys = mo.CELL + mo.U_FACE + mo.V_FACE + ("terrain",)
    m1, a1 = mo.coarsen_metrics({k: c.metrics[k] for k in keys}, c.alpha)
    (nx1, ny1, mz), _ = s.mg_level(1)
    o1 = Oracle(nx1, ny1, c.nz, 2 * c.dx)
    o1.set_metrics(dict(m1, terrain=m1["terrain"]))
    o1.set_alpha(a1)
    o1.assemble()

    class _Case:
        pass
    cc = _Case(); cc.alpha = a1; cc.metrics = m1
    A1 = _to_kji(stencil_matrix(o1, cc), nx1, ny1, mz)
    o1.close()
    (nx2, ny2, _), _ = s.mg_level(2)
    q2 = s.mg_stencil(2)
    assert q2.shape == (27, mz, ny2 + 2, nx2 + 2)
    A2 = _stencil_to_matrix(q2, nx2, ny2, mz)
    err2 = _rowwise_close(A2, _galerkin(A1, nx1, ny1, nx2, ny2, mz), 0)
    print(f"level 2 vs R A1 P: {err2:.2e}")
    assert err2 <= 2e-5
    (nx3, ny3, _), _ = s.mg_level(3)
    A3 = _stencil_to_matrix(s.mg_stencil(3), nx3, ny3, mz)
    err3 = _rowwise_close(A3, _galerkin(A2, nx2, ny2, nx3, ny3, mz), 0)
    print(f"level 3 vs R A2 P: {err3:.2e}")
    assert err3 <= 2e-5
    # interior rows below the ghost planes carry all 27 offsets
    assert np.count_nonzero(q2[:, mz // 2, 2, 2]) == 27


def test_galerkin_smoothing_step():
    """The wavefront smoothing step on a Galerkin level, out = y + 0.9 M^-1
    (b - A y) with A the level's 27-point operator and M its vertical
    tridiagonal part, against the same step by sparse solves in FP64."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from hicar_amd import _native as nat
    c, o, div, s = _setup("fractal")
    s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-6, start=1, max_iters=2000)
    rng = np.random.default_rng(4)
    for level in (2, 3):
        (nx, ny, mz), _ = s.mg_level(level)
        q = s.mg_stencil(level)
        A = _stencil_to_matrix(q, nx, ny, mz)
        Ac = A.tocoo()
        vert = np.abs(Ac.row - Ac.col) % ((nx + 2) * (ny + 2)) == 0
        M = sp.csr_matrix((Ac.data[vert], (Ac.row[vert], Ac.col[vert])), shape=A.shape)
        shape = (mz, ny + 2, nx + 2)
        b = rng.standard_normal(shape).astype(np.float32)
        y = rng.standard_normal(shape).astype(np.float32)
        got = s.mg_smooth(level, True, b, y)
        r = b.ravel().astype(np.float64) - A @ y.ravel().astype(np.float64)
        want = y.ravel() + 0.9 * spla.spsolve(M.tocsc(), r)
        err = _rel(got.ravel(), want)
        print(f"level {level} ({nx}x{ny}): smoothing step vs FP64 sparse {err:.2e}")
        assert err <= 1e-4, (level, err)
        with pytest.raises(RuntimeError):
            s.mg_smooth(level, False, b, y)  # the 15-point kernel forms do not apply


def test_galerkin_levels_take_fewer_iterations():
    """Steep fractal terrain (1500 m over 100 x 90 columns at 50 m): the
    hierarchy with Galerkin operators from its second coarse level reaches the
    same tolerance in fewer inner iterations than the rediscretised one, and
    both reach the answer (the same lambda to the solve's accuracy)."""
    from hicar_amd import _native as nat
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = S.fractal_case(nx=100, ny=90, nz=16, dx=50.0, relief=1500.0, seed=5, lowest=10.0, top=8000.0)
    res = {}
    for gal in (2, 0):
        s = WindSolver(c.nx, c.ny, c.nz, c.dx)
        s.set_mg_galerkin(gal)
        s.set_metrics(c.metrics)
        div = s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True)
        info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=1, max_iters=5000, inner_tol=0.1)
        assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
        res[gal] = (info["iters"], s.get_lambda())
        s.close()
    print(f"inner iterations: Galerkin {res[2][0]}, rediscretised {res[0][0]}")
    assert res[2][0] < res[0][0]
    assert _rel(res[2][1], res[0][1]) <= 1e-6
