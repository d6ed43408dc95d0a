"""Pins of the oracle's terrain boundary against closed forms that do not come
from the oracle: the surface slopes of init_module_vars
(wind_iterative.F90:2363-2364, 2398-2409) on a tilted-plane terrain, and the
bottom 10-point BC row of spmv (:962-988) applied to a manufactured lambda that
is linear in x, y and the level index. Each pin is shown to fail on a planted
mistake (orc_plant, set on the CPU only).

Closed form of the bottom row. With lambda = a x + b y + c k (x = i dx,
y = j dx, k the row level), the ten terms of row (i, 0, j) telescope:
    y = -(a sx + b sy) jaco / (sx^2 + sy^2 + alpha^2) + c / dz_if(i,1,j)
with sx, sy the surface slopes, jaco and alpha at level 1 and
dz_if(i,1,j) = advection_dz(i,1,j): the discrete no-normal-flow condition
d(lambda)/dz = (sx dlambda/dx + sy dlambda/dy) jaco / (sx^2 + sy^2 + alpha^2).
The top row (:950-960) gives c / advection_dz(i,nz,j).

It also pins the FP32-faithful mode (orc_set_faithful) against the pinned FP64
oracle: every coefficient within FP32 rounding of the FP64 one and exactly
FP32-representable, and the solve + correction distance between the two stated.
"""
import numpy as np
import pytest

import oracle as orc

SX, SY = 0.31, -0.17


def tilted_case(nx=11, ny=9, nz=5, dx=250.0, seed=5):
    """Hand-built metrics on the terrain h = SX x + SY y. The 3-D slope
    metrics are random (they matter only at the edge columns, where the
    surface slope is dzdx(i, k_s, j)); levels are stretched."""
    rng = np.random.default_rng(seed)
    dz = 20.0 * 1.4 ** np.arange(nz)
    m = {}
    for k, shp in (("dzdx", (ny, nz, nx)), ("dzdy", (ny, nz, nx)), ("dzdx_u", (ny, nz, nx + 1)),
                   ("dzdy_v", (ny + 1, nz, nx))):
        m[k] = 0.2 * rng.uniform(-1.0, 1.0, shp)
    for k, shp in (("jacobian", (ny, nz, nx)), ("jacobian_w", (ny, nz, nx)), ("jacobian_u", (ny, nz, nx + 1)),
                   ("jacobian_v", (ny + 1, nz, nx))):
        m[k] = 1.0 + 0.2 * rng.uniform(-1.0, 1.0, shp)
    m["advection_dz"] = np.broadcast_to(dz[None, :, None], (ny, nz, nx)).copy()
    x = dx * np.arange(1, nx + 1)
    y = dx * np.arange(1, ny + 1)
    m["terrain"] = 900.0 + SX * x[None, :] + SY * y[:, None]
    alpha = rng.uniform(0.3, 0.95, (ny, nz, nx))
    return m, alpha, dx


def _oracle(m, alpha, dx):
    ny, nz, nx = m["dzdx"].shape
    o = orc.Oracle(nx, ny, nz, dx)
    o.set_metrics(m)
    o.set_alpha(alpha)
    o.assemble()
    return o


@pytest.fixture(autouse=True)
def _no_plant():
    orc.plant(0)
    yield
    orc.plant(0)


def expected_slopes(m):
    ny, nz, nx = m["dzdx"].shape
    sx = np.full((ny, nx), SX)
    sy = np.full((ny, nx), SY)
    sx[:, 0], sx[:, -1] = m["dzdx"][:, 0, 0], m["dzdx"][:, 0, -1]     # i = 1, nx: dzdx(i, k_s, j)
    sy[0, :], sy[-1, :] = m["dzdy"][0, 0, :], m["dzdy"][-1, 0, :]     # j = 1, ny: dzdy(i, k_s, j)
    return sx, sy


def _slope_error(o, m):
    _, sx, sy = o.planes()
    ex, ey = expected_slopes(m)
    return max(np.abs(sx - ex).max(), np.abs(sy - ey).max())


def test_surface_slopes_of_a_tilted_plane():
    m, alpha, dx = tilted_case()
    assert _slope_error(_oracle(m, alpha, dx), m) <= 1e-12


@pytest.mark.parametrize("code", [4, 5])
def test_surface_slope_pin_catches_planted_mistakes(code):
    m, alpha, dx = tilted_case()
    orc.plant(code)
    assert _slope_error(_oracle(m, alpha, dx), m) > 0.1


def manufactured(nx, ny, nz, dx, a, b, c):
    """lambda = a x + b y + c k over every solver row (lateral rows included)."""
    i = np.arange(nx + 2)
    j = np.arange(ny + 2)
    k = np.arange(nz + 2)
    return (a * dx * i[None, None, :] + b * dx * j[:, None, None] + c * k[None, :, None]) * np.ones((ny + 2, nz + 2, nx + 2))


def bottom_closed_form(m, alpha, dx, a, b, c):
    sx, sy = expected_slopes(m)
    jaco = m["jacobian"][:, 0, :]
    al = alpha[:, 0, :]
    dz1 = m["advection_dz"][:, 0, :]
    return -(a * sx + b * sy) * jaco / (sx ** 2 + sy ** 2 + al ** 2) + c / dz1


def _bottom_error(o, m, alpha, dx, abc):
    ny, nz, nx = m["dzdx"].shape
    lam = manufactured(nx, ny, nz, dx, *abc)
    y = o.spmv(lam)
    got = y[1:-1, 0, 1:-1]
    want = bottom_closed_form(m, alpha, dx, *abc)
    return np.abs(got - want).max() / np.abs(want).max()


ABC = [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.7, -0.4, 0.3)]


@pytest.mark.parametrize("abc", ABC)
def test_bottom_row_closed_form(abc):
    m, alpha, dx = tilted_case()
    o = _oracle(m, alpha, dx)
    assert _bottom_error(o, m, alpha, dx, abc) <= 1e-12
    # the top row: (lambda(mz-1) - lambda(mz-2)) / dz_if(nz+1) = c / advection_dz(nz)
    ny, nz, nx = m["dzdx"].shape
    y = o.spmv(manufactured(nx, ny, nz, dx, *abc))
    np.testing.assert_allclose(y[1:-1, -1, 1:-1], abc[2] / m["advection_dz"][:, -1, :], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("code,abc", [(1, (1.0, 0.0, 0.0)), (2, (0.7, -0.4, 0.3)), (3, (0.0, 1.0, 0.0)),
                                      (6, (0.0, 0.0, 1.0)), (4, (1.0, 0.0, 0.0)), (5, (0.0, 1.0, 0.0))])
def test_bottom_row_pin_catches_planted_mistakes(code, abc):
    """A flipped x-term sign, alpha for alpha^2 in denom, slope terms over dx
    instead of 2 dx, the wrong interface for 1/dz_if, and the two slope
    mistakes each move the bottom row away from its closed form."""
    m, alpha, dx = tilted_case()
    orc.plant(code)
    o = _oracle(m, alpha, dx)
    assert _bottom_error(o, m, alpha, dx, abc) > 1e-3


# --------------------------------------------------------------- FP32-faithful mode
def _is_f32(a):
    return np.array_equal(a, a.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("name", ["hill", "fractal_alpha"])
def test_faithful_coefficients_within_fp32_rounding(name):
    from conftest import small_case
    c = small_case(name)
    o64 = orc.Oracle(c.nx, c.ny, c.nz, c.dx)
    o32 = orc.Oracle(c.nx, c.ny, c.nz, c.dx, faithful=True)
    for o in (o64, o32):
        o.set_metrics(c.metrics)
        o.set_alpha(c.alpha)
        o.assemble()
    dzi64, sx64, sy64 = o64.planes()
    dzi32, sx32, sy32 = o32.planes()
    for a32, a64 in ((dzi32, dzi64), (sx32, sx64), (sy32, sy64)):
        assert _is_f32(a32)
        assert np.abs(a32 - a64).max() <= 4e-7 * np.abs(a64).max()
    differ = 0
    for q in "ABCDEFGHIJKLMNO":
        c32, c64 = o32.coef(q), o64.coef(q)
        assert _is_f32(c32), q
        scale = np.abs(c64).max() if q in "BCA" else np.abs(c64).max()
        # B and C carry a cancelling slope term X; A = -4/dx^2 - B - C
        assert np.abs(c32 - c64).max() <= 2e-6 * max(scale, 1e-300), q
        differ += int(np.any(c32 != c64))
    assert differ >= 8  # it does round: most arrays change


def test_faithful_distance_config0():
    """The distance between this build's FP64 path (bit-identical to the FP64
    oracle) and the reference's FP32 rounding, on the 64x64x20 Gaussian-hill
    case at 1e-8. The solve itself runs in FP64 in both; the difference comes
    from the FP32 coefficients, the FP32 lambda and the FP32 correction."""
    from conftest import small_case
    from faithful import faithful_distance
    d = faithful_distance(small_case("config0"))
    print("faithful distance 64x64x20:", d)
    assert 0.0 < d["lam"] < 1e-7
    assert 0.0 < d["u"] < 1e-6 and 0.0 < d["v"] < 1e-6 and 0.0 < d["w"] < 1e-5
    assert 0.0 < d["du"] < 1e-5 and 0.0 < d["div"] < 1e-3
