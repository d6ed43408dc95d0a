"""Learned initial guess for the wind solve (no counterpart in the reference).

Host side of the network in ``csrc/windhip_net.hip``: parameter handling, the
training pairs (network input assembled by the solver from its own fields,
target = the plain FP64 solution lambda), the Adam training loop, and
attaching a trained network to a :class:`~hicar_amd.wind_iterative.WindSolver`
so that ``solve(..., start=3)`` starts from the network's guess whenever its
FP64 residual is below the plain start's.

All arithmetic runs in the HIP kernels; torch only provides device buffers.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _native as nat

_DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "lnet_params.npy")

# (ci, co) of the 3x3x3 convolutions in parameter order (windhip.h)
CONVS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 64),
         (128, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal convolution weights, zero biases, small head weights."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in CONVS:
        parts.append(rng.standard_normal(co * 27 * ci) * np.sqrt(2.0 / (27 * ci)))
        parts.append(np.zeros(co))
    parts.append(rng.standard_normal(16) * np.sqrt(1.0 / 16))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


def conv_flops(D: int, H: int, W: int, B: int = 1) -> float:
    """Multiply-add flops (2 per MAC) of one forward pass."""
    vox = [B * D * H * W / 8 ** lev for lev in range(4)]
    levels = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0]
    f = sum(2.0 * 27 * ci * co * vox[lv] for (ci, co), lv in zip(CONVS, levels))
    return f + 2.0 * 16 * vox[0]


class Net:
    """A windhip_net handle."""

    def __init__(self, params: np.ndarray | None = None):
        self.lib = nat.load()
        h = C.c_void_p()
        self._check(self.lib.windhip_net_create(C.byref(h)), None)
        self.h = h
        self.n = int(self.lib.windhip_net_param_count())
        self.set_params(init_params() if params is None else params)

    def _check(self, rc, h):
        if rc != nat.WINDHIP_OK:
            msg = self.lib.windhip_net_last_error(h) if h else b"windhip_net_create failed"
            raise nat.WindHipError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.windhip_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n:
            raise ValueError(f"expected {self.n} parameters, got {p.size}")
        self._check(self.lib.windhip_net_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), self.h)

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n, dtype=np.float32)
        self._check(self.lib.windhip_net_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), self.h)
        return p

    def save(self, path: str) -> None:
        np.save(path, self.get_params())

    @classmethod
    def load(cls, path: str = _DEFAULT_WEIGHTS) -> "Net":
        return cls(np.load(path, allow_pickle=False))

    def forward(self, inp, out) -> None:
        """inp: device tensor (B, D, H, W, 8) float32; out: (B, D, H, W) float32."""
        B, D, H, W, _ = inp.shape
        self._check(self.lib.windhip_net_forward(self.h, inp.data_ptr(), B, D, H, W, out.data_ptr()), self.h)

    def loss_grad(self, inp, target, want_grad: bool = False):
        B, D, H, W, _ = inp.shape
        loss = C.c_double()
        g = np.empty(self.n, dtype=np.float32) if want_grad else None
        gp = g.ctypes.data_as(C.POINTER(C.c_float)) if want_grad else None
        self._check(self.lib.windhip_net_loss_grad(self.h, inp.data_ptr(), target.data_ptr(), B, D, H, W,
                                                   C.byref(loss), gp), self.h)
        return (loss.value, g) if want_grad else loss.value

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self._check(self.lib.windhip_net_adam_step(self.h, lr, beta1, beta2, eps), self.h)

    def time_forward(self, D: int, H: int, W: int, B: int = 1, reps: int = 5) -> float:
        ms = C.c_double()
        self._check(self.lib.windhip_net_time_forward(self.h, B, D, H, W, reps, C.byref(ms)), self.h)
        return ms.value


def attach(solver, net: Net | None, smooth: int | None = None) -> None:
    """Use `net` for solver.solve(..., start=nat.START_LEARNED); `smooth`
    sets the line-relaxation steps applied to its guess (default 4)."""
    nat.check(solver.lib.windhip_set_net(solver.h, net.h if net is not None else None))
    if smooth is not None:
        nat.check(solver.lib.windhip_set_guess_smoothing(solver.h, int(smooth)))
    solver._net = net  # keep it alive while attached


def learned_input(solver, device: str = "cuda"):
    """(input tensor (1, D, H, W, 8), scale) for the solver's last rhs."""
    import torch

    dims = (C.c_int * 3)()
    nat.check(solver.lib.windhip_learned_input(solver.h, None, dims, None))
    inp = torch.empty((1, dims[0], dims[1], dims[2], 8), dtype=torch.float32, device=device)
    scale = C.c_double()
    nat.check(solver.lib.windhip_learned_input(solver.h, inp.data_ptr(), dims, C.byref(scale)))
    return inp, scale.value


def lambda_target(solver, inp_shape, scale: float):
    """The solver's lambda on the cells, as the network target lambda / scale
    on the padded volume (zero outside the grid)."""
    lam = solver.get_lambda()  # (ny+2, nz+2, nx+2) rows
    _, D, H, W, _ = inp_shape
    t = np.zeros((1, D, H, W), dtype=np.float32)
    inner = lam[1:-1, 1:-1, 1:-1] / scale  # (ny, nz, nx)
    t[0, :solver.nz, :solver.ny, :solver.nx] = np.transpose(inner, (1, 0, 2))
    return t


def load_case(solver, case, device: str = "cuda"):
    """Upload a synthetic case's metrics into `solver` (device fields); returns
    (alpha, div) device tensors, div = D(u, v, w) of the case's winds."""
    import torch

    D_ = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)  # noqa: E731
    solver.set_metrics({k: D_(v) for k, v in case.metrics.items()
                        if k in ("dzdx", "dzdy", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w", "dzdx_u",
                                 "dzdy_v", "advection_dz", "terrain")})
    div = torch.empty(solver.mass_shape, dtype=torch.float64, device=device)
    solver.divergence(D_(case.u), D_(case.v), D_(case.w), density=D_(case.rho), adv_den=True, use_dqdt=True,
                      out=div)
    return D_(case.alpha), div


def make_pair(case, solver, device: str = "cuda", tol_rel: float = 1e-8):
    """One training pair from a synthetic case: the plain FP64 solve of its
    divergence, the network input assembled by the solver, the scaled target."""
    import torch

    alpha, div = load_case(solver, case, device)
    info = solver.solve(alpha, div, precision=0, start=nat.START_ZERO, tol_rel=tol_rel, max_iters=20000)
    inp, scale = learned_input(solver, device)
    tgt = torch.from_numpy(lambda_target(solver, tuple(inp.shape), scale)).to(device)
    return inp, tgt, info


def evaluate(solver, alpha, div, net: Net, tol_rel: float = 1e-8, precisions=(0, 1), smooth: int = 4) -> dict:
    """Iterations and time of the same system solved from three starts, for
    each precision (0 plain FP64, 1 mixed): zero; the zero start after `smooth`
    line-relaxation steps (a network with zero parameters: what the relaxation
    alone gives); the network's guess after the same steps."""
    zero_net = Net(np.zeros(net.n, dtype=np.float32))
    out = {}
    for prec in precisions:
        for name, start, nn in (("zero", nat.START_ZERO, None), ("relax", nat.START_LEARNED, zero_net),
                                ("learned", nat.START_LEARNED, net)):
            attach(solver, nn, smooth)
            solver.set_lambda(np.zeros(solver.rows_shape))
            info = solver.solve(alpha, div, precision=prec, start=start, tol_rel=tol_rel, max_iters=50000)
            out[f"{'plain' if prec == 0 else 'mixed'}_{name}"] = info
    attach(solver, None)
    zero_net.close()
    return out


def train(net: Net, inputs, targets, epochs: int, batch: int = 8, lr: float = 1e-3, seed: int = 0,
          log=None, time_budget_s: float | None = None) -> list:
    """Adam over (input, target) device tensors of equal shape; returns the
    mean loss per epoch."""
    import torch

    rng = np.random.default_rng(seed)
    n = len(inputs)
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        order = rng.permutation(n)
        tot, cnt = 0.0, 0
        for q in range(0, n - batch + 1, batch):
            idx = order[q:q + batch]
            xb = torch.cat([inputs[i] for i in idx], dim=0).contiguous()
            yb = torch.cat([targets[i] for i in idx], dim=0).contiguous()
            tot += net.loss_grad(xb, yb)
            net.adam_step(lr)
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e} ({time.time() - t0:.0f}s)")
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            break
    return hist


# ------------------------------------------------------------------ training CLI
def _patch(args):
    """A seeded training case (runs in a worker process; numpy only)."""
    from . import synthetic

    seed, n, nz = args
    rng = np.random.default_rng(seed)
    return synthetic.fractal_case(n, n, nz, dx=50.0, relief=float(rng.uniform(500.0, 3000.0)),
                                  speed=float(rng.uniform(2.0, 20.0)), seed=seed,
                                  varying_alpha=bool(rng.integers(0, 2)),
                                  direction_deg=float(rng.uniform(0.0, 360.0)))


def main(argv=None) -> int:
    """python -m hicar_amd.learned --patches N --out params.npy

    Generates seeded 128x128x40 fractal-terrain patches (relief 500-3000 m,
    wind 2-20 m/s, half of them with a varying stability weight), solves each
    with the plain FP64 GPU solver to rel. residual 1e-8, and trains the
    network with Adam (lr 1e-3, batch 8) on (input, lambda) pairs."""
    import argparse
    import multiprocessing as mp
    import sys

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--patches", type=int, default=2000)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--time-budget", type=float, default=None, help="stop training after this many seconds")
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    log = lambda m: print(f"[learned] {m}", flush=True)  # noqa: E731
    # worker processes are forked before this process touches the GPU
    pool = mp.get_context("fork").Pool(args.workers)
    import torch

    from .wind_iterative import WindSolver

    t0 = time.time()
    inputs, targets, iters = [], [], []
    jobs = [(args.seed + q, args.size, args.nz) for q in range(args.patches)]
    solver = WindSolver(args.size, args.size, args.nz, 50.0, device_fields=True)
    for q, case in enumerate(pool.imap(_patch, jobs, chunksize=2)):
        inp, tgt, info = make_pair(case, solver, "cuda")
        if info["status"] != 0:
            log(f"patch {q}: plain solve status {info['status']}, skipped")
            continue
        inputs.append(inp)
        targets.append(tgt)
        iters.append(info["iters"])
        if (q + 1) % 100 == 0:
            log(f"{q + 1}/{args.patches} pairs ({time.time() - t0:.0f}s, mean {np.mean(iters):.0f} FP64 iterations)")
    pool.close()
    solver.close()
    log(f"{len(inputs)} pairs in {time.time() - t0:.0f}s")
    net = Net(np.load(args.init, allow_pickle=False) if args.init else None)
    hist = train(net, inputs, targets, args.epochs, args.batch, args.lr, seed=args.seed, log=log,
                 time_budget_s=args.time_budget)
    net.save(args.out)
    log(f"saved {args.out}; loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    del inputs, targets
    torch.cuda.empty_cache()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
