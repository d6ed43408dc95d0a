"""Pins of the CPU restatement of the learned-guess network (oracle/net_oracle.py):
finite differences of its float64 gradient, bf16 rounding known answers, the
convolution tap convention against a direct loop, adjoint pairs."""
import numpy as np

import net_oracle as no  # oracle/net_oracle.py (oracle/ is on sys.path)
from hicar_amd import learned


def _tiny_input(rng, D=8, H=8, W=8):
    x = np.zeros((1, D, H, W, 8))
    x[..., :6] = rng.standard_normal((1, D, H, W, 6))
    return x


def test_param_layout_matches_host():
    assert no.param_count() == sum(co * 27 * ci + co for ci, co in learned.CONVS) + 17
    assert no.CONVS == learned.CONVS
    p = learned.init_params(3)
    assert p.size == no.param_count()
    assert np.array_equal(no.init_params(3), p)


def test_round_bf16_known_answers():
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing in [1, 2)
    vals = np.array([one, one + ulp / 2, one + ulp / 2 + 2 ** -20, one + 1.5 * ulp, -3.0, 0.0, 2.0 ** -130])
    got = no.round_bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                 # tie -> even
    assert got[2] == one + ulp           # above the tie
    assert got[3] == one + 2 * ulp       # tie -> even (odd below)
    assert got[4] == -3.0 and got[5] == 0.0
    assert got[6] > 0                    # subnormal stays representable


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    p = no.init_params(1)
    net = no.Net(p, bf16=False)
    w, b = net.P["c0a"]
    x = rng.standard_normal((1, 3, 4, 5, 8))
    y = net._cols(x).reshape(-1, 27 * 8) @ w.reshape(16, -1).T + b
    y = y.reshape(1, 3, 4, 5, 16)
    ref = np.zeros_like(y)
    for z in range(3):
        for yy in range(4):
            for xx in range(5):
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            zz, y2, x2 = z + dz, yy + dy, xx + dx
                            if 0 <= zz < 3 and 0 <= y2 < 4 and 0 <= x2 < 5:
                                t = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)
                                ref[0, z, yy, xx] += w[:, t, :] @ x[0, zz, y2, x2]
    ref += b
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_pool_up_adjoint():
    rng = np.random.default_rng(1)
    net = no.Net(no.init_params(0), bf16=False)
    a = rng.standard_normal((1, 2, 3, 4, 5))
    bb = rng.standard_normal((1, 4, 6, 8, 5))
    assert np.isclose(np.sum(net.up(a) * bb), np.sum(a * net.up_back(bb)))
    # pool_back with zero skip is the adjoint of pool
    c = rng.standard_normal((1, 4, 6, 8, 5))
    assert np.isclose(np.sum(net.pool(c) * a), np.sum(c * net.pool_back(a, 0.0)))


def test_gradient_finite_differences():
    """float64 backward of the whole network against central differences."""
    rng = np.random.default_rng(2)
    p = no.init_params(5).astype(np.float64)
    p[:] += 0.01 * rng.standard_normal(p.size)  # nonzero biases
    x = _tiny_input(rng)
    tgt = rng.standard_normal((1, 8, 8, 8))
    net = no.Net(p, bf16=False)
    loss, g = net.loss_grad(x, tgt)
    assert np.all(np.isfinite(g))
    # components across every layer: first weight, a middle weight, the bias of each conv; the head
    idx, off = [], 0
    for ci, co in no.CONVS:
        nw = co * 27 * ci
        idx += [off, off + nw // 2 + 3, off + nw + co // 2]
        off += nw + co
    idx += [off, off + 7, off + 16]
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        lp = no.Net(pp, bf16=False).loss_grad(x, tgt)[0]
        lm = no.Net(pm, bf16=False).loss_grad(x, tgt)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[q]) <= 1e-5 * max(1.0, abs(g[q])) + 1e-7, (q, fd, g[q])


def test_bf16_mode_close_to_exact():
    rng = np.random.default_rng(3)
    p = no.init_params(7)
    x = _tiny_input(rng)
    tgt = rng.standard_normal((1, 8, 8, 8))
    l64, g64 = no.Net(p, bf16=False).loss_grad(x, tgt)
    l16, g16 = no.Net(p, bf16=True).loss_grad(x, tgt)
    assert abs(l16 - l64) <= 2e-2 * l64
    cos = g16 @ g64 / np.linalg.norm(g16) / np.linalg.norm(g64)
    assert cos > 0.99
