"""Handle life cycle on the GPU: solver handles that ran every solve path
(plain FP64, mixed, line, multigrid with the learned guess on its second
stream, probing, the caller's steps) give all their device memory back when
destroyed."""
import numpy as np
import pytest

from conftest import small_case

pytestmark = pytest.mark.gpu


def _used():
    from hicar_amd import _native as nat
    import ctypes as C
    from hicar_amd.wind_iterative import WindSolver
    s = WindSolver(4, 4, 3, 100.0)
    h, used = C.c_double(), C.c_double()
    nat.check(s.lib.windhip_memory(s.h, C.byref(h), C.byref(used)))
    s.close()
    return used.value


def test_destroyed_handles_free_their_memory():
    import torch
    from hicar_amd import learned
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = small_case("fractal_alpha")
    net = learned.Net(learned.init_params(3))

    def cycle():
        s = WindSolver(c.nx, c.ny, c.nz, c.dx)
        s.set_metrics(c.metrics)
        div = s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True)
        for prec in (nat.PLAIN_F64, nat.MIXED, nat.MIXED_LINE, nat.MIXED_MG):
            s.solve(c.alpha, div, precision=prec, tol_rel=1e-8, start=nat.START_ZERO, max_iters=20000)
        learned.attach(s, net)
        s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=nat.START_LEARNED, max_iters=20000)
        learned.attach(s, None)
        s.balance_uvw(c.u, c.v, c.w.copy(), density=c.rho)
        s.probe_operator(c.alpha, density=c.rho)
        s.solve(c.alpha, div, precision=nat.MIXED_MG, tol_rel=1e-8, start=nat.START_ZERO, max_iters=20000)
        s.close()

    cycle()  # first use: runtime and code-object set-up
    base = _used()
    for _ in range(5):
        cycle()
    grown = _used() - base
    print(f"device memory after 5 more create/solve/destroy cycles: {grown / 2**20:.1f} MiB above the baseline")
    assert grown < 16 * 2**20
