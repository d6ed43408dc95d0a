"""Parity of the HIP solver (through the C-ABI) with the CPU oracle.

The FP64 path is designed to be bit-identical to the oracle (same expression
order, no FMA, same reduction tree), so most comparisons are exact; the bar
from the north star (<= 1e-10 relative L2) is asserted as well where it is
the contract. The mixed FP32/FP64 path is held to 1e-6 relative L2 in lambda
and to the 1e-8 true residual.
"""
import numpy as np
import pytest

from conftest import oracle_for, small_case

pytestmark = pytest.mark.gpu


def _solver(case, dtype=np.float64):
    from hicar_amd.wind_iterative import WindSolver
    s = WindSolver(case.nx, case.ny, case.nz, case.dx, dtype=dtype)
    s.set_metrics(case.metrics)
    return s


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("name", ["hill", "fractal", "fractal_alpha"])
def test_operator_bitwise(name, rng):
    c = small_case(name)
    s = _solver(c)
    o = oracle_for(c)
    s.set_alpha(c.alpha)
    o.set_alpha(c.alpha)
    o.assemble()
    o.extract_diagonal()
    for q in "BCHILM":
        g = s.get_rows(q)[1:-1, 1:-1, 1:-1]
        assert np.array_equal(g, o.coef(q)), f"coefficient {q} differs"
    assert np.array_equal(s.get_rows("dinv"), o.rows("dinv"))
    for _ in range(3):
        x = rng.standard_normal(s.rows_shape)
        assert np.array_equal(s.apply_operator(x), o.spmv(x))


@pytest.mark.parametrize("name", ["hill", "fractal", "fractal_alpha"])
@pytest.mark.parametrize("graph", [0, 1])
def test_solve_bitwise(name, graph):
    c = small_case(name)
    s = _solver(c)
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    ref = o.solve(c.alpha, div, max_iters=3000, tol_rel=1e-8)
    got = s.solve(c.alpha, div, max_iters=3000, tol_rel=1e-8, use_graph=graph, iters_per_sync=7)
    assert got["status"] == ref["status"] == 0
    assert got["iters"] == ref["iters"]
    assert got["res0"] == ref["res0"] and got["res_final"] == ref["res_final"]
    lam_g, lam_o = s.get_lambda(), o.get_lambda()
    assert np.array_equal(lam_g, lam_o)
    assert _rel(lam_g, lam_o) <= 1e-10
    # true residual reported by the device
    rhs = o.rows("rhs")
    tr = np.linalg.norm(rhs - o.spmv(lam_o)) / np.linalg.norm(rhs)
    assert abs(got["true_rel_res"] - tr) <= 1e-6 * tr + 1e-16


def test_warm_start_and_alpha_refresh():
    """Three calls in a row: warm start from the previous lambda, alpha refresh
    only when alpha varies (wind_iterative.F90:410-442)."""
    c = small_case("fractal_alpha")
    s = _solver(c)
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    alphas = [c.alpha, np.ones_like(c.alpha) * 0.7, np.clip(c.alpha * 1.1, 0, 1)]
    for a in alphas:
        ref = o.solve(a, div * 1.3, max_iters=3000, tol_rel=1e-9)
        got = s.solve(a, div * 1.3, max_iters=3000, tol_rel=1e-9)
        assert got["iters"] == ref["iters"]
        assert np.array_equal(s.get_lambda(), o.get_lambda())


@pytest.mark.parametrize("adv_den", [False, True])
def test_correction_and_divergence_bitwise(adv_den):
    c = small_case("hill")
    rho = 1.0 + 0.05 * np.random.default_rng(5).standard_normal(c.rho.shape)
    s = _solver(c)
    o = oracle_for(c)
    div_o = o.divergence(c.u, c.v, c.w, rho=rho if adv_den else None, adv_den=adv_den, use_dqdt=True)
    div_g = s.divergence(c.u, c.v, c.w, density=rho if adv_den else None, adv_den=adv_den, use_dqdt=True)
    assert np.array_equal(div_g, div_o)
    o.solve(c.alpha, div_o, max_iters=3000, tol_rel=1e-8)
    s.solve(c.alpha, div_o, max_iters=3000, tol_rel=1e-8)
    uo, vo, wo = o.correct(c.u, c.v, c.w, rho=rho, adv_den=adv_den)
    ug, vg, wg = c.u.copy(), c.v.copy(), c.w.copy()
    s.apply_correction(ug, vg, wg, density=rho, adv_den=adv_den)
    assert np.array_equal(ug, uo) and np.array_equal(vg, vo) and np.array_equal(wg, wo)


def test_config0_calc_iter_winds():
    """64x64x20 Gaussian hill, zero start, rel. tolerance 1e-8: lambda, u, v, w
    and the final divergence within 1e-10 relative L2 of the oracle."""
    c = small_case("config0")
    s = _solver(c)
    o = oracle_for(c)
    div = s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True)
    div_o = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    assert np.array_equal(div, div_o)
    ug, vg, wg = c.u.copy(), c.v.copy(), c.w.copy()
    info = s.calc_iter_winds(c.alpha, div, ug, vg, wg, density=c.rho, adv_den=True, tol_rel=1e-8,
                             max_iters=5000, start=1)
    ref = o.solve(c.alpha, div_o, max_iters=5000, tol_rel=1e-8)
    uo, vo, wo = o.correct(c.u, c.v, c.w, rho=c.rho, adv_den=True)
    assert info["iters"] == ref["iters"]
    for a, b in ((s.get_lambda(), o.get_lambda()), (ug, uo), (vg, vo), (wg, wo)):
        assert _rel(a, b) <= 1e-10
    d1g = s.divergence(ug, vg, wg, density=c.rho, adv_den=True, use_dqdt=True)
    d1o = o.divergence(uo, vo, wo, rho=c.rho, adv_den=True, use_dqdt=True)
    assert _rel(d1g, d1o) <= 1e-10
    assert info["true_rel_res"] <= 1.5e-8


def test_reference_unit_behaviour():
    """tests/unit/test_wind_iterative.F90:51-209: flat 250 m plane, 20 levels,
    alpha = 1, seeded cos/sin winds, adv_den false: finite and max|div| falls."""
    from hicar_amd import synthetic as S
    c = S.flat_plane_case(nx=40, ny=36, nz=20)
    s = _solver(c)
    div0 = s.divergence(c.u, c.v, c.w, adv_den=False, use_dqdt=True)
    u, v, w = c.u.copy(), c.v.copy(), c.w.copy()
    s.calc_iter_winds(c.alpha, div0, u, v, w, adv_den=False)
    div1 = s.divergence(u, v, w, adv_den=False, use_dqdt=True)
    assert np.all(np.isfinite(div1))
    assert np.abs(div1).max() < np.abs(div0).max()


def test_probed_operator():
    c = small_case("hill")
    s = _solver(c)
    o = oracle_for(c)
    o.set_alpha(c.alpha)
    leak_o = o.probe(rho=c.rho, adv_den=True)
    leak_g = s.probe_operator(c.alpha, density=c.rho, adv_den=True)
    assert leak_o == 0.0 and leak_g == 0.0
    x = np.random.default_rng(2).standard_normal(s.rows_shape)
    s.set_alpha(c.alpha)
    o.extract_diagonal()
    assert np.array_equal(s.apply_operator(x), o.spmv(x))
    assert np.array_equal(s.get_rows("dinv"), o.rows("dinv"))
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    ref = o.solve(c.alpha, div, max_iters=3000, tol_rel=1e-11)
    got = s.solve(c.alpha, div, max_iters=3000, tol_rel=1e-11)
    assert got["iters"] == ref["iters"]
    assert np.array_equal(s.get_lambda(), o.get_lambda())
    # A = 2 D(G(.)): after the probed solve the corrected winds are divergence free
    u, v, w = c.u.copy(), c.v.copy(), c.w.copy()
    s.apply_correction(u, v, w, density=c.rho, adv_den=True)
    d1 = s.divergence(u, v, w, density=c.rho, adv_den=True, use_dqdt=True)
    assert np.linalg.norm(d1) <= 1e-9 * np.linalg.norm(div)


def test_f32_fields_match_promoted_inputs():
    """HICAR passes `real` (FP32) fields; the solve promotes them exactly, so the
    result equals the oracle's on the FP32-rounded inputs."""
    c = small_case("hill")
    from hicar_amd.wind_iterative import WindSolver
    s = WindSolver(c.nx, c.ny, c.nz, c.dx, dtype=np.float32)
    m32 = {k: (np.asarray(v, dtype=np.float32) if isinstance(v, np.ndarray) else v) for k, v in c.metrics.items()}
    s.set_metrics(m32)
    m64 = {k: (np.asarray(v, dtype=np.float32).astype(np.float64) if isinstance(v, np.ndarray) else v)
           for k, v in c.metrics.items()}
    from oracle import Oracle
    o = Oracle(c.nx, c.ny, c.nz, c.dx)
    o.set_metrics(m64)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True).astype(np.float32)
    a32 = c.alpha.astype(np.float32)
    ref = o.solve(a32.astype(np.float64), div.astype(np.float64), tol_rel=1e-8, max_iters=3000)
    got = s.solve(a32, div, tol_rel=1e-8, max_iters=3000)
    assert got["iters"] == ref["iters"]
    assert np.array_equal(s.get_lambda(), o.get_lambda())


def test_mixed_precision_small():
    c = small_case("fractal_alpha")
    s = _solver(c)
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    o.solve(c.alpha, div, max_iters=5000, tol_rel=1e-12)
    info = s.solve(c.alpha, div, precision=1, tol_rel=1e-8, start=1, max_iters=20000)
    assert info["status"] == 0
    assert info["true_rel_res"] <= 1e-8
    assert _rel(s.get_lambda(), o.get_lambda()) <= 1e-6


def test_error_paths():
    from hicar_amd import _native as nat
    c = small_case("hill")
    s = _solver(c)
    bad = c.alpha.copy()
    div = np.zeros_like(c.alpha)
    div[3, 2, 1] = np.nan
    with pytest.raises(nat.WindHipError) as e:
        s.solve(c.alpha, div)
    assert e.value.code == nat.WINDHIP_ERR_NAN
    # zero rhs converges at iteration 0 (wind_iterative.F90:635-640)
    info = s.solve(bad, np.zeros_like(c.alpha), start=1)
    assert info["status"] == 0 and info["iters"] == 0
    # iteration cap -> status 1 with the retry ladder disabled
    o = oracle_for(c)
    d = o.divergence(c.u, c.v, c.w)
    info = s.solve(c.alpha, d, start=1, max_iters=3, retry=0, tol_rel=1e-12)
    assert info["status"] == 1 and info["iters"] == 3


def test_map_factors_bitwise_and_cleared():
    """Map factors (calc_divergence's mapfac_mxy, mapfac_my_u, mapfac_mx_v;
    calc_updated_winds' mapfac_mx_u, mapfac_my_v; both in probing): the
    divergence, the correction and the probed operator and its solve are
    bit-identical to the oracle's; a later set_metrics without map factors
    means factors of 1 again."""
    c = small_case("hill")
    rng = np.random.default_rng(11)
    f = lambda shp: 1.0 + 0.1 * rng.uniform(-1.0, 1.0, shp)  # noqa: E731
    mf = dict(mf_mx_u=f((c.ny, c.nx + 1)), mf_my_u=f((c.ny, c.nx + 1)), mf_my_v=f((c.ny + 1, c.nx)),
              mf_mx_v=f((c.ny + 1, c.nx)), mf_mxy=f((c.ny, c.nx)))
    m = dict(c.metrics)
    m.update({"mapfac_" + k[3:]: a for k, a in mf.items()})
    from hicar_amd.wind_iterative import WindSolver
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(m)
    o = oracle_for(c)
    dkw = dict(mf_my_u=mf["mf_my_u"], mf_mx_v=mf["mf_mx_v"], mf_mxy=mf["mf_mxy"])
    div_o = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True, **dkw)
    assert np.array_equal(s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True), div_o)
    o.solve(c.alpha, div_o, max_iters=3000, tol_rel=1e-8)
    s.solve(c.alpha, div_o, max_iters=3000, tol_rel=1e-8)
    assert np.array_equal(s.get_lambda(), o.get_lambda())
    uo, vo, wo = o.correct(c.u, c.v, c.w, rho=c.rho, adv_den=True, mf_mx_u=mf["mf_mx_u"], mf_my_v=mf["mf_my_v"])
    ug, vg, wg = c.u.copy(), c.v.copy(), c.w.copy()
    s.apply_correction(ug, vg, wg, density=c.rho, adv_den=True)
    assert np.array_equal(ug, uo) and np.array_equal(vg, vo) and np.array_equal(wg, wo)
    # probing with map factors: the same operator and solve
    o.set_alpha(c.alpha)
    assert o.probe(rho=c.rho, adv_den=True, **mf) == 0.0
    assert s.probe_operator(c.alpha, density=c.rho, adv_den=True) == 0.0
    x = np.random.default_rng(12).standard_normal(s.rows_shape)
    s.set_alpha(c.alpha)
    assert np.array_equal(s.apply_operator(x), o.spmv(x))
    ref = o.solve(c.alpha, div_o, max_iters=3000, tol_rel=1e-11)
    got = s.solve(c.alpha, div_o, max_iters=3000, tol_rel=1e-11)
    assert got["iters"] == ref["iters"] and np.array_equal(s.get_lambda(), o.get_lambda())
    ug, vg, wg = c.u.copy(), c.v.copy(), c.w.copy()
    s.apply_correction(ug, vg, wg, density=c.rho, adv_den=True)
    d1 = s.divergence(ug, vg, wg, density=c.rho, adv_den=True, use_dqdt=True)
    assert np.linalg.norm(d1) <= 1e-9 * np.linalg.norm(div_o)
    # metrics again without map factors: factors of 1
    s.set_metrics(c.metrics)
    assert np.array_equal(s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True),
                          o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True))


@pytest.mark.parametrize("adv_den", [False, True])
@pytest.mark.parametrize("horz_only", [False, True])
@pytest.mark.parametrize("use_dqdt", [False, True])
def test_every_divergence_branch_bitwise(adv_den, horz_only, use_dqdt):
    """calc_divergence (wind.F90:250-523) in all eight branch combinations,
    with density, bit for bit."""
    c = small_case("fractal_alpha")
    rho = 1.0 + 0.05 * np.random.default_rng(9).standard_normal(c.rho.shape)
    s = _solver(c)
    o = oracle_for(c)
    want = o.divergence(c.u, c.v, c.w, rho=rho, adv_den=adv_den, horz_only=horz_only, use_dqdt=use_dqdt)
    got = s.divergence(c.u, c.v, c.w, density=rho, adv_den=adv_den, horz_only=horz_only, use_dqdt=use_dqdt)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("max_iters", [5, 20, 40])
def test_retry_ladder_bitwise(max_iters):
    """At the iteration cap without a 100x residual drop the solve retries from
    lambda = 0 with 3, then 4 preconditioner sweeps (wind_iterative.F90:474-504);
    a residual still above the initial one after the ladder is the reference's
    fatal divergence (:506-528), an error here. Cap 5: divergence; cap 20: the
    whole ladder, status 1; cap 40: status 1 without the ladder. Status, sweep
    count, iterations, residual and lambda as the oracle's."""
    from hicar_amd import _native as nat
    c = small_case("hill")
    s = _solver(c)
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    ref = o.solve(c.alpha, div, max_iters=max_iters, tol_rel=1e-12)
    assert ref["status"] == 1
    if ref["res_final"] > ref["res0"]:
        with pytest.raises(nat.WindHipError) as e:
            s.solve(c.alpha, div, max_iters=max_iters, tol_rel=1e-12, start=1)
        assert e.value.code == nat.WINDHIP_ERR_DIVERGED
        return
    got = s.solve(c.alpha, div, max_iters=max_iters, tol_rel=1e-12, start=1)
    assert (got["status"], got["nsweeps"], got["iters"]) == (ref["status"], ref["nsweeps"], ref["iters"])
    assert got["res0"] == ref["res0"] and got["res_final"] == ref["res_final"]
    assert np.array_equal(s.get_lambda(), o.get_lambda())
