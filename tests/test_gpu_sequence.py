"""A short solve sequence with turning forcing (the shape of the 1024x1024x60
sequence benchmark, at a size the tests can afford): every start converges to
the same answer, and starting from the previous step's lambda saves work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_sequence_starts_agree_and_warm_start_saves_iterations():
    from hicar_amd import _native as nat
    from hicar_amd import learned
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    case = S.fractal_case(nx=48, ny=40, nz=12, dx=50.0, relief=800.0, seed=5, lowest=10.0, top=6000.0)
    m = case.metrics
    s = WindSolver(case.nx, case.ny, case.nz, case.dx)
    s.set_metrics(m)
    net = learned.Net(learned.init_params(3))
    steps = 5
    finals, iters = {}, {}
    for name, start in (("zero", nat.START_ZERO), ("warm", nat.START_WARM), ("learned", nat.START_LEARNED)):
        learned.attach(s, net if start == nat.START_LEARNED else None)
        s.set_lambda(np.zeros(s.rows_shape))
        total = 0
        for step in range(steps):
            u, v, w = S.power_law_winds(m, 8.0 * (1.0 + 0.02 * step), 25.0, 1.0, np.random.default_rng(9),
                                        direction_deg=float(step))
            div = s.divergence(u, v, w, adv_den=True, use_dqdt=True)
            info = s.solve(case.alpha, div, precision=nat.MIXED_MG, start=start, tol_rel=1e-10, max_iters=20000,
                           inner_tol=0.1)
            assert info["status"] == 0 and info["true_rel_res"] <= 1e-10
            if step > 0:
                total += info["iters"]
        finals[name] = s.get_lambda()
        iters[name] = total
    ref = finals["zero"]
    for name in ("warm", "learned"):
        assert np.linalg.norm(finals[name] - ref) <= 1e-6 * np.linalg.norm(ref)
    assert iters["warm"] < iters["zero"]
