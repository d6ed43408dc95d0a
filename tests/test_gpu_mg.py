"""The multigrid inner preconditioner (MIXED_MG) through the C-ABI.

Its coarse operators are checked against the Galerkin product R A P formed
with the oracle's FP64 spmv and the aggregation transfers of
oracle/mg_oracle.py; its solves against the oracle's FP64 lambda (1e-6
relative L2, the mixed-precision bar).
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


def _interior(e):
    e = e.copy()
    e[0] = e[-1] = 0.0
    e[:, :, 0] = e[:, :, -1] = 0.0
    return e


def test_coarse_levels_are_galerkin_products():
    """With hscale = 1 level 1 is R A P of the fine operator and level 2 is
    R A_1 P, i.e. the 4 x 4 aggregation of A (odd edges included: ny = 23)."""
    from hicar_amd import _native as nat
    c, o, div, s = _setup()
    o.set_alpha(c.alpha)
    o.assemble()
    s.set_multigrid(hscale=1.0)
    info = s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-6, start=1, max_iters=2000)
    assert info["status"] == 0
    rng = np.random.default_rng(11)
    (nx1, ny1, mz), a1 = s.mg_level(1)
    assert (nx1, ny1, mz) == (mo.coarse_size(c.nx), mo.coarse_size(c.ny), c.nz + 2)
    e1 = _interior(rng.standard_normal((ny1 + 2, mz, nx1 + 2)))
    want = mo.restrict(o.spmv(mo.prolong(e1, c.nx, c.ny)), c.nx, c.ny)
    got = mo.apply15(mo.solver_level_to_rows(a1), e1)
    inner = (slice(1, -1), slice(None), slice(1, -1))
    assert _rel(got[inner], want[inner]) <= 1e-5
    (nx2, ny2, _), a2 = s.mg_level(2)
    e2 = _interior(rng.standard_normal((ny2 + 2, mz, nx2 + 2)))
    pe = mo.prolong(mo.prolong(e2, nx1, ny1), c.nx, c.ny)
    want2 = mo.restrict(mo.restrict(o.spmv(pe), c.nx, c.ny), nx1, ny1)
    got2 = mo.apply15(mo.solver_level_to_rows(a2), e2)
    assert _rel(got2[inner], want2[inner]) <= 1e-5


def test_hscale_scales_only_the_same_level_couplings():
    from hicar_amd import _native as nat
    c, o, div, s = _setup()
    levels = {}
    for hs in (1.0, 0.7):
        s.set_multigrid(hscale=hs)
        s.set_lambda(np.zeros(s.rows_shape))
        s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-6, start=1, max_iters=2000)
        levels[hs] = s.mg_level(1)[1].astype(np.float64)
    a, b = levels[1.0], levels[0.7]
    mid = (slice(1, -1), slice(1, -1), slice(1, -1))  # interior rows, k = 1 .. mz-2
    for q in range(1, 15):
        if q in (3, 4, 5, 6):  # D, E, F, G
            assert np.allclose(b[q][mid], 0.7 * a[q][mid], rtol=1e-6, atol=0)
        else:
            assert np.array_equal(b[q][mid], a[q][mid])
    for k in (0, -1):  # ghost-plane rows keep the exact product
        assert np.array_equal(b[:, k], a[:, k])
    # every non-lateral row still sums to zero (the row sum is entry 0)
    scale = np.abs(b[1:]).max()
    assert np.abs(b[0][:, 1:-1, 1:-1]).max() <= 1e-6 * scale


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
