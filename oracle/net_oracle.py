"""CPU restatement of the learned-guess network (hicar_amd/csrc/windhip_net.hip).

Test infrastructure only: imported by ``tests/`` -- never by the product.
The network has no counterpart in the reference; this numpy version states the
same architecture, parameter layout and rounding points so that the GPU
forward pass and every parameter gradient can be checked against it.

Two modes:
  ``bf16=True``  rounds to bf16 exactly where the GPU stores bf16 (the input,
                 every convolution / pooling / upsampling output, every
                 activation gradient); products are exact and sums are carried
                 in float64, so the GPU differs only by its FP32 summation
                 order.
  ``bf16=False`` float64 throughout; its gradients are pinned by central finite
                 differences of the loss (tests/test_net_oracle.py).

Arrays are channels-last (B, D, H, W, C).
"""
from __future__ import annotations

import numpy as np

# (ci, co) of the 3x3x3 convolutions in parameter order
CONVS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 64),
         (128, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]
NAMES = ["c0a", "c0b", "c1a", "c1b", "c2a", "c2b", "c3a", "c3b", "d2a", "d2b", "d1a", "d1b", "d0a", "d0b"]


def param_count() -> int:
    return sum(co * 27 * ci + co for ci, co in CONVS) + 17


def unpack(p: np.ndarray) -> dict:
    """Views of the flat parameter array: name -> (W (co, 27, ci), b (co,))."""
    out, off = {}, 0
    for name, (ci, co) in zip(NAMES, CONVS):
        w = p[off:off + co * 27 * ci].reshape(co, 27, ci)
        off += co * 27 * ci
        out[name] = (w, p[off:off + co])
        off += co
    out["head"] = (p[off:off + 16], p[off + 16:off + 17])
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal convolution weights, zero biases, small head."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in CONVS:
        parts.append(rng.standard_normal(co * 27 * ci) * np.sqrt(2.0 / (27 * ci)))
        parts.append(np.zeros(co))
    parts.append(rng.standard_normal(16) * np.sqrt(1.0 / 16))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


def round_bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest bf16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


class Net:
    def __init__(self, params: np.ndarray, bf16: bool = True):
        self.p = np.asarray(params, dtype=np.float64)
        self.bf16 = bf16
        self.P = unpack(self.p)

    def r(self, x):
        return round_bf16(x) if self.bf16 else x

    def w(self, name):
        w, b = self.P[name]
        return (self.r(w) if self.bf16 else w), b

    # ---- primitives
    @staticmethod
    def _cols(x):
        B, D, H, W, C = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
        cols = np.empty((B, D, H, W, 27, C))
        for t in range(27):
            a, b_, c = t // 9, (t // 3) % 3, t % 3
            cols[..., t, :] = xp[:, a:a + D, b_:b_ + H, c:c + W, :]
        return cols

    def conv(self, name, x):
        w, b = self.w(name)
        co = w.shape[0]
        cols = self._cols(x)
        y = cols.reshape(-1, w.shape[1] * w.shape[2]) @ w.reshape(co, -1).T + b
        return self.r(np.maximum(y, 0.0)).reshape(x.shape[:4] + (co,))

    def conv_back(self, name, g, yout, x, need_dx=True):
        """g: gradient w.r.t. the ReLU output yout; returns dW, db, dx."""
        w, _ = self.w(name)
        co, _, ci = w.shape
        gm = np.where(yout > 0, g, 0.0)
        cols = self._cols(x).reshape(-1, 27 * ci)
        g2 = gm.reshape(-1, co)
        dW = (g2.T @ cols).reshape(co, 27, ci)
        db = g2.sum(axis=0)
        dx = None
        if need_dx:
            dcols = (g2 @ w.reshape(co, -1)).reshape(x.shape[:4] + (27, ci))
            B, D, H, W, _ = x.shape
            dxp = np.zeros((B, D + 2, H + 2, W + 2, ci))
            for t in range(27):
                a, b_, c = t // 9, (t // 3) % 3, t % 3
                dxp[:, a:a + D, b_:b_ + H, c:c + W, :] += dcols[..., t, :]
            dx = self.r(dxp[:, 1:-1, 1:-1, 1:-1, :])
        return dW, db, dx

    def pool(self, x):
        B, D, H, W, C = x.shape
        y = x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))
        return self.r(y)

    @staticmethod
    def up(x):
        return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)

    def up_back(self, g):
        B, D, H, W, C = g.shape
        return self.r(g.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6)))

    def pool_back(self, g, skip):
        return self.r(self.up(g) * 0.125 + skip)

    # ---- network
    def forward(self, inp):
        """inp (B, D, H, W, 8) -> output (B, D, H, W); keeps the activations."""
        a = {}
        x0 = self.r(inp.astype(np.float64))
        a["x0"] = x0
        a["a0"] = self.conv("c0a", x0)
        a["s0"] = self.conv("c0b", a["a0"])
        a["p1"] = self.pool(a["s0"])
        a["a1"] = self.conv("c1a", a["p1"])
        a["s1"] = self.conv("c1b", a["a1"])
        a["p2"] = self.pool(a["s1"])
        a["a2"] = self.conv("c2a", a["p2"])
        a["s2"] = self.conv("c2b", a["a2"])
        a["p3"] = self.pool(a["s2"])
        a["a3"] = self.conv("c3a", a["p3"])
        a["b3"] = self.conv("c3b", a["a3"])
        a["cat2"] = np.concatenate([self.up(a["b3"]), a["s2"]], axis=-1)
        a["e2a"] = self.conv("d2a", a["cat2"])
        a["e2b"] = self.conv("d2b", a["e2a"])
        a["cat1"] = np.concatenate([self.up(a["e2b"]), a["s1"]], axis=-1)
        a["e1a"] = self.conv("d1a", a["cat1"])
        a["e1b"] = self.conv("d1b", a["e1a"])
        a["cat0"] = np.concatenate([self.up(a["e1b"]), a["s0"]], axis=-1)
        a["e0a"] = self.conv("d0a", a["cat0"])
        a["e0b"] = self.conv("d0b", a["e0a"])
        hw, hb = self.P["head"]
        out = a["e0b"] @ hw + hb[0]
        self.a = a
        return out

    def loss_grad(self, inp, target):
        """MSE loss and the flat gradient of every parameter."""
        out = self.forward(inp)
        a = self.a
        M = out.size
        e = out - target
        loss = float(np.sum(e * e) / M)
        dout = 2.0 * e / M
        if self.bf16:
            dout = dout.astype(np.float32).astype(np.float64)
        grads = {}
        hw, _ = self.P["head"]
        grads["head"] = ((a["e0b"] * dout[..., None]).reshape(-1, 16).sum(axis=0), np.array([dout.sum()]))
        g = self.r(dout[..., None] * hw)
        gW, gb, g = self.conv_back("d0b", g, a["e0b"], a["e0a"]); grads["d0b"] = (gW, gb)
        gW, gb, gcat0 = self.conv_back("d0a", g, a["e0a"], a["cat0"]); grads["d0a"] = (gW, gb)
        g = self.up_back(gcat0[..., :32])
        gW, gb, g = self.conv_back("d1b", g, a["e1b"], a["e1a"]); grads["d1b"] = (gW, gb)
        gW, gb, gcat1 = self.conv_back("d1a", g, a["e1a"], a["cat1"]); grads["d1a"] = (gW, gb)
        g = self.up_back(gcat1[..., :64])
        gW, gb, g = self.conv_back("d2b", g, a["e2b"], a["e2a"]); grads["d2b"] = (gW, gb)
        gW, gb, gcat2 = self.conv_back("d2a", g, a["e2a"], a["cat2"]); grads["d2a"] = (gW, gb)
        g = self.up_back(gcat2[..., :64])
        gW, gb, g = self.conv_back("c3b", g, a["b3"], a["a3"]); grads["c3b"] = (gW, gb)
        gW, gb, g = self.conv_back("c3a", g, a["a3"], a["p3"]); grads["c3a"] = (gW, gb)
        g = self.pool_back(g, gcat2[..., 64:])
        gW, gb, g = self.conv_back("c2b", g, a["s2"], a["a2"]); grads["c2b"] = (gW, gb)
        gW, gb, g = self.conv_back("c2a", g, a["a2"], a["p2"]); grads["c2a"] = (gW, gb)
        g = self.pool_back(g, gcat1[..., 64:])
        gW, gb, g = self.conv_back("c1b", g, a["s1"], a["a1"]); grads["c1b"] = (gW, gb)
        gW, gb, g = self.conv_back("c1a", g, a["a1"], a["p1"]); grads["c1a"] = (gW, gb)
        g = self.pool_back(g, gcat0[..., 32:])
        gW, gb, g = self.conv_back("c0b", g, a["s0"], a["a0"]); grads["c0b"] = (gW, gb)
        gW, gb, _ = self.conv_back("c0a", g, a["a0"], a["x0"], need_dx=False); grads["c0a"] = (gW, gb)
        flat = []
        for name in NAMES:
            flat += [grads[name][0].ravel(), grads[name][1].ravel()]
        flat += [grads["head"][0], grads["head"][1]]
        return loss, np.concatenate(flat)
