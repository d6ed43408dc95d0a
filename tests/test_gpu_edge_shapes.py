"""Grids at and near the smallest size the solve accepts (nx, ny >= 3,
nz >= 2; both the library and the oracle refuse smaller) and sizes that
straddle the 62-column stencil tiles, on hand-built metrics (small random
slopes, Jacobians near 1, stretched levels): the plain FP64 solve is
bit-identical to the oracle, the multigrid mixed solve within 1e-6 in lambda,
and a grid below the minimum is refused with an argument error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(3, 3, 2), (3, 4, 2), (4, 3, 3), (5, 3, 2), (3, 70, 2), (70, 3, 2), (3, 3, 3),
          (60, 5, 2), (61, 4, 3), (62, 7, 2), (63, 3, 4), (125, 6, 2)]


def metrics(nx, ny, nz, seed):
    dz = 12.0 * 1.35 ** np.arange(nz)
    rng = np.random.default_rng(seed)
    m = {}
    for k, shp in (("dzdx", (ny, nz, nx)), ("dzdy", (ny, nz, nx)), ("dzdx_u", (ny, nz, nx + 1)),
                   ("dzdy_v", (ny + 1, nz, nx))):
        m[k] = 0.05 * rng.uniform(-1.0, 1.0, shp)
    for k, shp in (("jacobian", (ny, nz, nx)), ("jacobian_w", (ny, nz, nx)), ("jacobian_u", (ny, nz, nx + 1)),
                   ("jacobian_v", (ny + 1, nz, nx))):
        m[k] = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, shp)
    m["advection_dz"] = np.broadcast_to(dz[None, :, None], (ny, nz, nx)).copy()
    m["terrain"] = 20.0 * rng.uniform(0.0, 1.0, (ny, nx))
    return m


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_small_and_tile_straddling_grids(shape):
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    from oracle import Oracle
    nx, ny, nz = shape
    m = metrics(nx, ny, nz, seed=nx * 1000 + ny * 10 + nz)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((ny, nz, nx + 1))
    v = rng.standard_normal((ny + 1, nz, nx))
    w = np.zeros((ny, nz, nx))
    rho = 1.0 + 0.05 * rng.random((ny, nz, nx))
    alpha = 0.5 + 0.5 * rng.random((ny, nz, nx))
    o = Oracle(nx, ny, nz, 100.0)
    o.set_metrics(m)
    s = WindSolver(nx, ny, nz, 100.0)
    s.set_metrics(m)
    div_o = o.divergence(u, v, w, rho=rho, adv_den=True)
    div_g = s.divergence(u, v, w, density=rho, adv_den=True, use_dqdt=True)
    assert np.array_equal(div_g, div_o)
    ref = o.solve(alpha, div_o, tol_rel=1e-9, max_iters=5000)
    got = s.solve(alpha, div_o, tol_rel=1e-9, max_iters=5000, start=nat.START_ZERO)
    assert got["status"] == ref["status"] == 0 and got["iters"] == ref["iters"]
    assert np.array_equal(s.get_lambda(), o.get_lambda())
    o.solve(alpha, div_o, tol_rel=1e-12, max_iters=20000)
    s.set_lambda(np.zeros(s.rows_shape))
    mg = s.solve(alpha, div_o, precision=nat.MIXED_MG, tol_rel=1e-8, start=nat.START_ZERO, max_iters=20000,
                 inner_tol=0.1)
    assert mg["status"] == 0 and mg["true_rel_res"] <= 1e-8
    assert _rel(s.get_lambda(), o.get_lambda()) <= 1e-6
    s.close()


@pytest.mark.parametrize("shape", [(2, 5, 3), (5, 2, 3), (5, 5, 1)])
def test_grids_below_the_minimum_are_refused(shape):
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    with pytest.raises(nat.WindHipError):
        WindSolver(*shape, 100.0)
