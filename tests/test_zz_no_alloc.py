"""windhip_balance_uvw_noalloc, the allocation-free variant of
windhip_balance_uvw (its FP64 divergence scratch is a static device array):
through the C-ABI, on two calls in a row (the second on a smaller grid), its w
equals windhip_balance_uvw's bit for bit, the call adds no device memory to the
handle, and it refuses host fields and grids larger than its scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _solver(nx, ny, nz, seed, dtype):
    import torch
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver
    c = S.fractal_case(nx=nx, ny=ny, nz=nz, dx=50.0, relief=600.0, seed=seed, lowest=10.0, top=6000.0)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    D = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda").to(tdt)  # noqa: E731
    s = WindSolver(nx, ny, nz, c.dx, dtype=dtype, device_fields=True)
    s.set_metrics({k: D(v) for k, v in c.metrics.items()
                   if k in ("dzdx", "dzdy", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w", "dzdx_u",
                            "dzdy_v", "advection_dz", "terrain")})
    rho = D(1.0 + 0.05 * np.random.default_rng(seed).random((ny, nz, nx)))
    return s, D(c.u), D(c.v), rho


@pytest.mark.parametrize("adv_den,use_dqdt", [(1, 1), (0, 0)])
def test_noalloc_equals_balance_uvw_on_two_calls(adv_den, use_dqdt):
    _need_gpu()
    import torch
    from hicar_amd import _native as nat
    lib = nat.load()
    # the second call runs on a smaller grid (and FP32 fields) after the first
    for (nx, ny, nz), dtype in (((130, 70, 24), np.float64), ((37, 29, 9), np.float32)):
        s, u, v, rho = _solver(nx, ny, nz, seed=nx, dtype=dtype)
        before = s.memory()[0]
        w_new = torch.full(s.mass_shape, 7.0, dtype=u.dtype, device="cuda")
        rc = lib.windhip_balance_uvw_noalloc(s.h, C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()),
                                             C.c_void_p(w_new.data_ptr()), C.c_void_p(rho.data_ptr()),
                                             adv_den, use_dqdt)
        nat.check(rc)
        torch.cuda.synchronize()
        assert s.memory()[0] == before  # nothing allocated by the call
        w_ref = torch.zeros(s.mass_shape, dtype=u.dtype, device="cuda")
        nat.check(lib.windhip_balance_uvw(s.h, C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()),
                                          C.c_void_p(w_ref.data_ptr()), C.c_void_p(rho.data_ptr()),
                                          adv_den, use_dqdt))
        torch.cuda.synchronize()
        assert s.memory()[0] > before  # the allocating entry point made its scratch
        print(f"{nx}x{ny}x{nz} {np.dtype(dtype).name}: max|w| {float(w_ref.abs().max()):.3e}", flush=True)
        assert bool(torch.isfinite(w_ref).all()) and float(w_ref.abs().max()) > 0
        assert torch.equal(w_new, w_ref)
        s.close()


def test_noalloc_refuses_host_fields_and_oversized_grids():
    _need_gpu()
    from hicar_amd import _native as nat
    lib = nat.load()
    dummy = C.c_void_p(1)
    # host fields: refused before any device work
    d = nat.Desc(nx=8, ny=8, nz=4, dx=100.0, ims=1, ime=8, kms=1, kme=4, jms=1, jme=8,
                 field_dtype=nat.WINDHIP_F64, field_mem=nat.WINDHIP_HOST)
    h = C.c_void_p()
    nat.check(lib.windhip_create(C.byref(d), C.byref(h)))
    assert lib.windhip_balance_uvw_noalloc(h, dummy, dummy, dummy, None, 1, 1) == nat.WINDHIP_ERR_ARG
    assert b"device memory" in lib.windhip_last_error()
    lib.windhip_destroy(h)
    # a mass grid of 4096 x 4 x 4096 memory elements (> 16,777,216): refused by size
    d = nat.Desc(nx=8, ny=8, nz=4, dx=100.0, ims=1, ime=4096, kms=1, kme=4, jms=1, jme=4096,
                 field_dtype=nat.WINDHIP_F64, field_mem=nat.WINDHIP_DEVICE)
    h = C.c_void_p()
    nat.check(lib.windhip_create(C.byref(d), C.byref(h)))
    assert lib.windhip_balance_uvw_noalloc(h, dummy, dummy, dummy, None, 1, 1) == nat.WINDHIP_ERR_ARG
    assert b"static scratch" in lib.windhip_last_error()
    lib.windhip_destroy(h)
