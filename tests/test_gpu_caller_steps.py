"""Parity of the caller's steps around the solve on the GPU (balance_uvw +
calc_w, calc_w_real, calc_idealized_wgrid; wind.F90:123-234, 940-998,
1175-1227) with the CPU oracle, through the C-ABI.

FP64 fields: bit-identical (same expression order, no FMA). FP32 fields (the
reference's `real`): the kernels compute in FP64 and round where the reference
stores an array, so they agree with the oracle on the FP32-rounded inputs to
FP32 rounding, held here to 2e-6 relative L2 (w of balance_uvw is a running
column sum of nz rounded terms).
"""
import numpy as np
import pytest

from conftest import oracle_for, small_case

pytestmark = pytest.mark.gpu


def _solver(c, dtype=np.float64, device_fields=False, metrics=None):
    from hicar_amd.wind_iterative import WindSolver
    s = WindSolver(c.nx, c.ny, c.nz, c.dx, dtype=dtype, device_fields=device_fields)
    s.set_metrics(c.metrics if metrics is None else metrics)
    return s


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _winds(c, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(c.u.shape), rng.standard_normal(c.v.shape), rng.standard_normal(c.w.shape),
            1.0 + 0.1 * rng.random(c.w.shape))


@pytest.mark.parametrize("name", ["hill", "fractal_alpha"])
@pytest.mark.parametrize("adv_den", [True, False])
@pytest.mark.parametrize("with_rho", [True, False])
def test_balance_uvw_bitwise(name, adv_den, with_rho):
    c = small_case(name)
    s, o = _solver(c), oracle_for(c)
    u, v, w, rho = _winds(c, 11)
    rho = rho if with_rho else None
    want = o.balance_uvw(u, v, w, rho=rho, adv_den=adv_den)
    wg = w.copy()
    got = s.balance_uvw(u, v, wg, density=rho, adv_den=adv_den)
    assert got is wg
    assert np.array_equal(wg, want)
    # the balanced winds have zero full divergence on the device too
    if adv_den or not with_rho:
        div = s.divergence(u, v, wg, density=rho, adv_den=adv_den, horz_only=False, use_dqdt=False)
        scale = np.abs(s.divergence(u, v, wg, density=rho, adv_den=adv_den, horz_only=True, use_dqdt=False)).max()
        assert np.abs(div).max() <= 1e-12 * scale


@pytest.mark.parametrize("name", ["hill", "fractal"])
def test_w_real_and_idealized_wgrid_bitwise(name):
    c = small_case(name)
    s, o = _solver(c), oracle_for(c)
    u, v, w, _ = _winds(c, 12)
    for wgrid in (w, None):
        assert np.array_equal(s.calc_w_real(u, v, wgrid), o.w_real(u, v, wgrid))
    wr = o.w_real(u, v, w)
    got = s.idealized_wgrid(u, v, wr)
    assert np.array_equal(got, o.idealized_wgrid(u, v, wr))
    assert np.all(got[:, -1] == 0.0)


def test_f32_fields_within_rounding():
    c = small_case("fractal_alpha")
    m32 = {k: (np.asarray(x, dtype=np.float32) if isinstance(x, np.ndarray) else x) for k, x in c.metrics.items()}
    m64 = {k: (x.astype(np.float64) if isinstance(x, np.ndarray) else x) for k, x in m32.items()}
    s = _solver(c, dtype=np.float32, metrics=m32)
    from oracle import Oracle
    o = Oracle(c.nx, c.ny, c.nz, c.dx)
    o.set_metrics(m64)
    u, v, w, rho = (a.astype(np.float32) for a in _winds(c, 13))
    w32 = w.copy()
    s.balance_uvw(u, v, w32, density=rho, adv_den=True)
    assert w32.dtype == np.float32
    assert _rel(w32, o.balance_uvw(u, v, w, rho=rho, adv_den=True)) <= 2e-6
    assert _rel(s.calc_w_real(u, v, w), o.w_real(u, v, w)) <= 2e-6
    wr = o.w_real(u, v, w).astype(np.float32)
    assert _rel(s.idealized_wgrid(u, v, wr), o.idealized_wgrid(u, v, wr)) <= 2e-6


def test_device_fields_match_host_fields():
    import torch
    c = small_case("hill")
    u, v, w, rho = _winds(c, 14)
    host = _solver(c)
    want_w = host.balance_uvw(u, v, w.copy(), density=rho, adv_den=True)
    want_r = host.calc_w_real(u, v, w)
    want_i = host.idealized_wgrid(u, v, want_r)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")
    keys = ("dzdx", "dzdy", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w", "dzdx_u", "dzdy_v",
            "advection_dz", "terrain")
    dev = _solver(c, device_fields=True, metrics={k: t(c.metrics[k]) for k in keys})
    du, dv, dw, dr, dw0 = t(u), t(v), t(w), t(rho), t(w)
    wr, wi = torch.empty_like(dw), torch.empty_like(dw)
    torch.cuda.synchronize()
    dev.balance_uvw(du, dv, dw, density=dr, adv_den=True)
    dev.calc_w_real(du, dv, dw0, out=wr)
    dev.idealized_wgrid(du, dv, wr, out=wi)
    torch.cuda.synchronize()
    assert np.array_equal(dw.cpu().numpy(), want_w)
    assert np.array_equal(wr.cpu().numpy(), want_r)
    assert np.array_equal(wi.cpu().numpy(), want_i)
