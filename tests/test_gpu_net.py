"""GPU tests of the learned-guess network (csrc/windhip_net.hip) against the
CPU restatement (oracle/net_oracle.py, bf16 rounding at the same points):
forward pass, loss, every parameter gradient, the Adam step; the network
input assembled by the solver; the start selection of the learned start."""
import numpy as np
import pytest

import net_oracle as no  # oracle/net_oracle.py
from conftest import oracle_for, small_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _input(rng, B=1, D=8, H=16, W=16):
    x = np.zeros((B, D, H, W, 8), dtype=np.float32)
    x[..., :6] = rng.standard_normal((B, D, H, W, 6)).astype(np.float32)
    return x


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_forward_matches_oracle():
    from hicar_amd import learned
    dev = _dev()
    rng = np.random.default_rng(0)
    p = learned.init_params(11)
    p[:] += (0.02 * rng.standard_normal(p.size)).astype(np.float32)
    net = learned.Net(p)
    x = _input(rng)
    out = torch.empty((1, 8, 16, 16), dtype=torch.float32, device=dev)
    net.forward(torch.from_numpy(x).to(dev), out)
    ref = no.Net(p, bf16=True).forward(x.astype(np.float64))
    assert _rel(out.cpu().numpy(), ref) <= 1e-2


def test_loss_and_gradients_match_oracle():
    from hicar_amd import learned
    dev = _dev()
    rng = np.random.default_rng(1)
    p = learned.init_params(12)
    p[:] += (0.02 * rng.standard_normal(p.size)).astype(np.float32)
    net = learned.Net(p)
    x = _input(rng, B=2)
    tgt = rng.standard_normal((2, 8, 16, 16)).astype(np.float32)
    loss, g = net.loss_grad(torch.from_numpy(x).to(dev), torch.from_numpy(tgt).to(dev), want_grad=True)
    lref, gref = no.Net(p, bf16=True).loss_grad(x.astype(np.float64), tgt.astype(np.float64))
    assert abs(loss - lref) <= 1e-3 * lref
    cos = float(g @ gref / np.linalg.norm(g) / np.linalg.norm(gref))
    assert cos >= 0.999
    # every layer's weight and bias gradient
    off = 0
    for name, (ci, co) in zip(no.NAMES, no.CONVS):
        nw = co * 27 * ci
        assert _rel(g[off:off + nw], gref[off:off + nw]) <= 3e-2, name
        assert _rel(g[off + nw:off + nw + co], gref[off + nw:off + nw + co]) <= 3e-2, name + " bias"
        off += nw + co
    assert _rel(g[off:], gref[off:]) <= 1e-2


def test_adam_step_matches_formula():
    from hicar_amd import learned
    dev = _dev()
    rng = np.random.default_rng(2)
    p = learned.init_params(13)
    net = learned.Net(p)
    x = torch.from_numpy(_input(rng)).to(dev)
    tgt = torch.from_numpy(rng.standard_normal((1, 8, 16, 16)).astype(np.float32)).to(dev)
    m = np.zeros(p.size, np.float64)
    v = np.zeros(p.size, np.float64)
    pr = p.astype(np.float64)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    for step in (1, 2):
        _, g = net.loss_grad(x, tgt, want_grad=True)
        net.adam_step(lr, b1, b2, eps)
        g = g.astype(np.float64)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        pr = pr - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
        got = net.get_params()
        assert np.max(np.abs(got - pr)) <= 1e-5 * max(1.0, np.max(np.abs(pr)))


def test_training_reduces_loss():
    from hicar_amd import learned
    dev = _dev()
    rng = np.random.default_rng(3)
    net = learned.Net(learned.init_params(14))
    xs = [torch.from_numpy(_input(rng)).to(dev) for _ in range(4)]
    # a smooth target the network can represent: a linear combination of the inputs
    ts = [(0.5 * x[..., 0] - 0.25 * x[..., 3]).contiguous() for x in xs]
    hist = learned.train(net, xs, ts, epochs=40, batch=2, lr=2e-3)
    assert hist[-1] < 0.5 * hist[0]


def test_learned_input_matches_fields():
    """The solver's network input against the same quantities restated on the CPU."""
    from hicar_amd import learned
    from hicar_amd.wind_iterative import WindSolver
    dev = _dev()
    c = small_case("fractal")
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    s.solve(c.alpha, div, max_iters=5, start=1)
    inp, scale = learned.learned_input(s, "cuda")
    got = inp.cpu().numpy()[0]
    D, H, W, _ = got.shape
    assert (D, H, W) == ((c.nz + 7) // 8 * 8, (c.ny + 7) // 8 * 8, (c.nx + 7) // 8 * 8)
    tr = lambda a: np.transpose(a, (1, 0, 2))  # (ny, nz, nx) -> (nz, ny, nx)  # noqa: E731
    b = -2.0 * div
    rms = np.sqrt(np.mean(b * b))
    sl = (slice(0, c.nz), slice(0, c.ny), slice(0, c.nx))
    assert np.allclose(got[sl + (0,)], tr(b / rms), rtol=1e-5, atol=1e-6)
    assert np.allclose(got[sl + (1,)], tr(c.metrics["dzdx"]), rtol=1e-6, atol=1e-7)
    assert np.allclose(got[sl + (2,)], tr(c.metrics["dzdy"]), rtol=1e-6, atol=1e-7)
    assert np.allclose(got[sl + (3,)], tr(c.metrics["jacobian"]), rtol=1e-6)
    assert np.allclose(got[sl + (4,)], tr(c.alpha), rtol=1e-6)
    dz = c.metrics["advection_dz"]
    h = (np.cumsum(dz, axis=1) - 0.5 * dz) / dz.sum(axis=1, keepdims=True)
    assert np.allclose(got[sl + (5,)], tr(h), rtol=1e-5)
    assert np.all(got[c.nz:] == 0) and np.all(got[:, c.ny:] == 0) and np.all(got[..., 6:] == 0)
    assert np.isclose(scale, rms * c.dx * c.dx, rtol=1e-12)


def test_learned_start_selection():
    """A random network guesses badly: the guard keeps the plain start, and the
    solve is bit-identical to the plain warm start; with the solution itself as
    the plain start, nothing beats it either."""
    from hicar_amd import learned
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = small_case("hill")
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    ref = s.solve(c.alpha, div, start=nat.START_ZERO, tol_rel=1e-8, max_iters=3000)
    lam_ref = s.get_lambda()
    net = learned.Net(learned.init_params(21))
    learned.attach(s, net)
    s.set_lambda(np.zeros_like(lam_ref))
    info = s.solve(c.alpha, div, start=nat.START_LEARNED, tol_rel=1e-8, max_iters=3000)
    assert info["guess_res"] > 0 and info["plain_res"] > 0
    if info["start_used"] == nat.START_WARM:
        assert info["guess_res"] >= info["plain_res"]
        assert info["iters"] == ref["iters"]
        assert np.array_equal(s.get_lambda(), lam_ref)
    else:
        assert info["start_used"] == nat.START_LEARNED and info["guess_res"] < info["plain_res"]
        assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
    # plain start = the converged solution: its residual is already at tolerance
    info2 = s.solve(c.alpha, div, start=nat.START_LEARNED, tol_rel=1e-8, max_iters=3000)
    assert info2["start_used"] == nat.START_WARM and info2["iters"] == 0
    learned.attach(s, None)
    with pytest.raises(nat.WindHipError):
        s.solve(c.alpha, div, start=nat.START_LEARNED)


def test_line_relaxation_matches_tridiagonal_solve():
    """One line-relaxation step from zero (a zero-parameter network, one
    smoothing step, no iterations) returns L^-1 b, L the tridiagonal column part
    of A, here solved directly from the independently assembled sparse matrix."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from test_oracle import stencil_matrix
    from hicar_amd import learned
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = small_case("fractal_alpha")
    o = oracle_for(c)
    o.set_alpha(c.alpha)
    o.assemble()
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    o.compute_rhs(div)
    b = o.rows("rhs").ravel()
    A = stencil_matrix(o, c).tocoo()
    mx, mz = c.nx + 2, c.nz + 2
    col_r, col_c = A.row % mx + mx * (A.row // (mx * mz)), A.col % mx + mx * (A.col // (mx * mz))
    keep = (col_r == col_c) & (np.abs(A.row // mx % mz - A.col // mx % mz) <= 1)
    T = sp.csr_matrix((A.data[keep], (A.row[keep], A.col[keep])), shape=A.shape)
    ref = spl.spsolve(T.tocsc(), b).reshape(o.rows_shape)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    zero = learned.Net(np.zeros(learned.Net().n, dtype=np.float32))
    learned.attach(s, zero, smooth=1)
    s.set_lambda(np.zeros(s.rows_shape))
    info = s.solve(c.alpha, div, start=nat.START_LEARNED, max_iters=0, retry=0)
    assert info["start_used"] == nat.START_LEARNED and info["guess_res"] < info["plain_res"]
    assert _rel(s.get_lambda(), ref) <= 1e-11


def test_mixed_line_matches_oracle():
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = small_case("fractal_alpha")
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    o.solve(c.alpha, div, max_iters=5000, tol_rel=1e-12)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    mixed = s.solve(c.alpha, div, precision=nat.MIXED, tol_rel=1e-8, start=1, max_iters=20000)
    s.set_lambda(np.zeros(s.rows_shape))
    info = s.solve(c.alpha, div, precision=nat.MIXED_LINE, tol_rel=1e-8, start=1, max_iters=20000)
    assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
    assert _rel(s.get_lambda(), o.get_lambda()) <= 1e-6
    assert info["iters"] < mixed["iters"]


def test_mixed_learned_start_keeps_the_answer():
    """In the mixed solves the learned guess (computed on a second stream
    during the solver's set-up) starts the FGMRES when its residual is below
    the plain start's: a zero-parameter network (its guess is the relaxed
    zero start) always qualifies, a random one may not; either way the
    converged answer stays within tolerance."""
    from hicar_amd import learned
    from hicar_amd import _native as nat
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    c = small_case("fractal_alpha")
    o = oracle_for(c)
    div = o.divergence(c.u, c.v, c.w, rho=c.rho, adv_den=True, use_dqdt=True)
    o.solve(c.alpha, div, max_iters=5000, tol_rel=1e-12)
    s = WindSolver(c.nx, c.ny, c.nz, c.dx)
    s.set_metrics(c.metrics)
    nets = {"zero": learned.Net(np.zeros(learned.init_params(0).size, dtype=np.float32)),
            "random": learned.Net(learned.init_params(5))}
    for name, net in nets.items():
        learned.attach(s, net)
        for prec in (nat.MIXED, nat.MIXED_LINE, nat.MIXED_MG):
            s.set_lambda(np.zeros(s.rows_shape))
            info = s.solve(c.alpha, div, precision=prec, start=nat.START_LEARNED, tol_rel=1e-8, max_iters=20000)
            assert info["guess_res"] > 0 and info["plain_res"] > 0
            used = info["guess_res"] < info["plain_res"]
            assert info["start_used"] == (nat.START_LEARNED if used else nat.START_WARM)
            if name == "zero":
                assert used
            assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
            assert _rel(s.get_lambda(), o.get_lambda()) <= 1e-6

def test_trained_network_gain_on_held_out_terrain():
    """What the trained network (hicar_amd/data/lnet_params.npy) gains over the
    method without it, on 128x128x40 fractal patches drawn like the training
    set's (relief 500-3000 m, wind 2-20 m/s, alpha constant or varying) from
    seeds the training never saw. The method without it is the same start
    pipeline with a zero-parameter network: the relaxed zero start. Against
    the converged lambda the trained guess has the smaller error and the
    smaller residual. The iteration counts of plain FP64 solves from each start
    are reported, not asserted: a relaxed start has already removed the
    fast-converging error, and BiCGStab's count to 1e-8 moves by tens of
    iterations either way (see the printed table)."""
    from hicar_amd import learned
    from hicar_amd import _native as nat
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver
    _dev()
    trained = learned.Net.load()
    relax_only = learned.Net(np.zeros(learned.init_params(0).size, dtype=np.float32))
    rows = []
    for seed in (50001, 50002, 50003):
        rng = np.random.default_rng(seed)
        c = S.fractal_case(128, 128, 40, dx=50.0, relief=float(rng.uniform(500.0, 3000.0)),
                           speed=float(rng.uniform(2.0, 20.0)), seed=seed, varying_alpha=bool(rng.integers(0, 2)),
                           direction_deg=float(rng.uniform(0.0, 360.0)))
        s = WindSolver(c.nx, c.ny, c.nz, c.dx)
        s.set_metrics(c.metrics)
        div = s.divergence(c.u, c.v, c.w, density=c.rho, adv_den=True, use_dqdt=True)
        zero = s.solve(c.alpha, div, start=nat.START_ZERO, tol_rel=1e-8, max_iters=20000)
        s.solve(c.alpha, div, start=nat.START_ZERO, tol_rel=1e-11, max_iters=20000)
        lam_star = s.get_lambda()
        row = dict(seed=seed, zero=zero["iters"])
        for name, net in (("relax", relax_only), ("trained", trained)):
            learned.attach(s, net)
            s.set_lambda(np.zeros(s.rows_shape))
            g = s.solve(c.alpha, div, start=nat.START_LEARNED, tol_rel=1e-8, max_iters=0)
            assert g["iters"] == 0 and g["start_used"] == nat.START_LEARNED
            row[name + "_err"] = _rel(s.get_lambda(), lam_star)
            row[name + "_res"] = g["guess_res"] / g["plain_res"]
            s.set_lambda(np.zeros(s.rows_shape))
            info = s.solve(c.alpha, div, start=nat.START_LEARNED, tol_rel=1e-8, max_iters=20000)
            assert info["status"] == 0 and info["true_rel_res"] <= 1e-8
            row[name] = info["iters"]
        learned.attach(s, None)
        rows.append(row)
        s.close()
    for r in rows:
        print(f"seed {r['seed']}: lambda error of the guess: relaxed zero start {r['relax_err']:.3f}, "
              f"trained {r['trained_err']:.3f}; residual / |b|: {r['relax_res']:.3f}, {r['trained_res']:.3f}; "
              f"iterations: zero start {r['zero']}, relaxed {r['relax']}, trained {r['trained']}")
    for r in rows:
        assert r["trained_err"] < r["relax_err"] < 1.0, r
        assert r["trained_res"] < r["relax_res"] < 1.0, r
