"""Pins of the oracle's map factors, CPU only: calc_divergence
(wind.F90:423-437: the cell sum times mapfac_mxy, each face flux divided by
its transverse factor) and calc_updated_winds (wind_iterative.F90:2022-2038:
the u and v increments times mapfac_mx_u and mapfac_my_v).

- a discrete divergence theorem: summed over a level, div dx / mapfac_mxy
  telescopes to the boundary fluxes u / mapfac_my_u and v / mapfac_mx_v;
- scaling all three divergence factors by one constant changes nothing;
- the correction's u and v increments are exactly the factor-free ones
  times mapfac_mx_u and mapfac_my_v, and w does not see map factors;
- with map factors, probing still builds A = 2 D(G(.)) exactly.
"""
import numpy as np

from conftest import oracle_for, small_case
from hicar_amd import synthetic as S


def _factors(nx, ny, seed):
    rng = np.random.default_rng(seed)
    f = lambda shp: 1.0 + 0.1 * rng.uniform(-1.0, 1.0, shp)  # noqa: E731
    return dict(mf_mx_u=f((ny, nx + 1)), mf_my_u=f((ny, nx + 1)), mf_my_v=f((ny + 1, nx)),
                mf_mx_v=f((ny + 1, nx)), mf_mxy=f((ny, nx)))


def test_divergence_theorem_with_map_factors():
    c = S.flat_plane_case(nx=9, ny=8, nz=4, dx=200.0)
    o = oracle_for(c)
    f = _factors(c.nx, c.ny, 1)
    rng = np.random.default_rng(2)
    u, v = rng.standard_normal(c.u.shape), rng.standard_normal(c.v.shape)
    w = np.zeros(c.w.shape)
    div = o.divergence(u, v, w, adv_den=False, horz_only=True, use_dqdt=False, mf_my_u=f["mf_my_u"],
                       mf_mx_v=f["mf_mx_v"], mf_mxy=f["mf_mxy"])
    # flat terrain: jacobian_u = jacobian_v = 1, so the face fluxes are u and v
    assert np.all(c.metrics["jacobian_u"] == 1.0) and np.all(c.metrics["jacobian_v"] == 1.0)
    inside = (div * 200.0 / f["mf_mxy"][:, None, :]).sum(axis=(0, 2))
    my, mx = f["mf_my_u"][:, None, :], f["mf_mx_v"][:, None, :]
    boundary = ((u[:, :, -1] / my[:, :, -1] - u[:, :, 0] / my[:, :, 0]).sum(axis=0)
                + (v[-1] / mx[-1] - v[0] / mx[0]).sum(axis=1))
    np.testing.assert_allclose(inside, boundary, rtol=1e-12, atol=1e-12)


def test_divergence_factors_cancel_under_common_scaling():
    c = small_case("hill")
    o = oracle_for(c)
    f = _factors(c.nx, c.ny, 3)
    rng = np.random.default_rng(4)
    u, v, w = rng.standard_normal(c.u.shape), rng.standard_normal(c.v.shape), rng.standard_normal(c.w.shape)
    kw = dict(rho=c.rho, adv_den=True, use_dqdt=True)
    d1 = o.divergence(u, v, w, mf_my_u=f["mf_my_u"], mf_mx_v=f["mf_mx_v"], mf_mxy=f["mf_mxy"], **kw)
    d2 = o.divergence(u, v, w, mf_my_u=1.7 * f["mf_my_u"], mf_mx_v=1.7 * f["mf_mx_v"], mf_mxy=1.7 * f["mf_mxy"],
                      **kw)
    np.testing.assert_allclose(d2, d1, rtol=1e-13, atol=1e-13 * np.abs(d1).max())
    ones = o.divergence(u, v, w, mf_my_u=np.ones_like(f["mf_my_u"]), mf_mx_v=np.ones_like(f["mf_mx_v"]),
                        mf_mxy=np.ones_like(f["mf_mxy"]), **kw)
    assert np.array_equal(ones, o.divergence(u, v, w, **kw))
    assert not np.allclose(d1, ones)


def test_correction_increments_scale_with_map_factors():
    c = small_case("fractal")
    o = oracle_for(c)
    o.set_alpha(c.alpha)
    f = _factors(c.nx, c.ny, 5)
    rng = np.random.default_rng(6)
    lam = np.zeros(o.rows_shape)
    lam[1:-1, 1:-1, 1:-1] = rng.standard_normal(lam[1:-1, 1:-1, 1:-1].shape)
    o.set_lambda(lam)
    z = (np.zeros(c.u.shape), np.zeros(c.v.shape), np.zeros(c.w.shape))
    for adv_den in (False, True):
        u0, v0, w0 = o.correct(*z, rho=c.rho, adv_den=adv_den)
        u1, v1, w1 = o.correct(*z, rho=c.rho, adv_den=adv_den, mf_mx_u=f["mf_mx_u"], mf_my_v=f["mf_my_v"])
        np.testing.assert_allclose(u1, f["mf_mx_u"][:, None, :] * u0, rtol=1e-14, atol=0)
        np.testing.assert_allclose(v1, f["mf_my_v"][:, None, :] * v0, rtol=1e-14, atol=0)
        assert np.array_equal(w1, w0)
        assert np.abs(u0).max() > 0 and np.abs(v0).max() > 0


def test_probing_with_map_factors_is_the_composition():
    c = small_case("hill")
    o = oracle_for(c)
    o.set_alpha(c.alpha)
    f = _factors(c.nx, c.ny, 7)
    assert o.probe(rho=c.rho, adv_den=True, **f) == 0.0
    rng = np.random.default_rng(8)
    x = np.zeros(o.rows_shape)
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal(x[1:-1, 1:-1, 1:-1].shape)
    y = o.spmv(x)
    o.set_lambda(x)
    u, v, w = o.correct(np.zeros_like(c.u), np.zeros_like(c.v), np.zeros_like(c.w), rho=c.rho, adv_den=True,
                        mf_mx_u=f["mf_mx_u"], mf_my_v=f["mf_my_v"])
    dg = o.divergence(u, v, w, rho=c.rho, adv_den=True, use_dqdt=True, mf_my_u=f["mf_my_u"], mf_mx_v=f["mf_mx_v"],
                      mf_mxy=f["mf_mxy"])
    np.testing.assert_allclose(y[1:-1, 1:-1, 1:-1], 2.0 * dg, rtol=1e-11, atol=1e-12 * np.abs(2 * dg).max())
