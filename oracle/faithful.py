"""Distance between the FP64 path and the reference's FP32 rounding.

Test infrastructure (like the rest of oracle/): the FP64 oracle is
bit-identical to the GPU's plain FP64 solve, and the oracle's FP32-faithful
mode rounds where HICAR's wind_iterative computes in `real` (module arrays,
coefficient assembly, the rhs product, lambda before the correction and the
correction itself: src/physics/wind_iterative.F90:86-91, 1525, 1770-2062).
Both start from the same FP32 inputs, as HICAR's fields are `real`.

    python oracle/faithful.py --config 0 1 2 --out profiles/faithful_distance.json

solves each configuration with both oracles to the stop and to 1e-12 and
writes the relative L2 distances of lambda, the corrected u, v, w, their
increments and the divergence after the correction.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, _HERE)
sys.path.insert(0, os.path.dirname(_HERE))

import oracle as orc  # noqa: E402


def faithful_distance(c, tol_rel=1e-8, max_iters=20000):
    """Relative L2 distances (FP64 path vs FP32-faithful restatement) of
    lambda, the corrected u, v, w, the u increment and the divergence after
    the correction, with the iteration counts of both solves."""
    f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)  # noqa: E731
    m = {k: (f32(v) if isinstance(v, np.ndarray) else v) for k, v in c.metrics.items()}
    alpha, rho = f32(c.alpha), f32(c.rho)
    u0, v0, w0 = f32(c.u), f32(c.v), f32(c.w)
    res = {}
    for mode in (False, True):
        o = orc.Oracle(c.nx, c.ny, c.nz, c.dx, faithful=mode)
        o.set_metrics(m)
        div = f32(o.divergence(u0, v0, w0, rho=rho, adv_den=True))
        r = o.solve(alpha, div, tol_rel=tol_rel, max_iters=max_iters)
        if r["status"] != 0:
            raise RuntimeError(f"oracle solve (faithful={mode}) did not converge: {r}")
        u, v, w = o.correct(u0, v0, w0, rho=rho, adv_den=True)
        d = o.divergence(u, v, w, rho=rho, adv_den=True)
        res[mode] = dict(lam=o.get_lambda(), u=u, v=v, w=w, div=d, iters=r["iters"])
        o.close()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    out = {k: rel(res[False][k], res[True][k]) for k in ("lam", "u", "v", "w", "div")}
    out["du"] = rel(res[False]["u"] - u0, res[True]["u"] - u0)
    out["dv"] = rel(res[False]["v"] - v0, res[True]["v"] - v0)
    out["iters_fp64"], out["iters_faithful"] = res[False]["iters"], res[True]["iters"]
    return out


def _case(n):
    from hicar_amd import synthetic as S
    if n == 0:
        return S.gaussian_hill_case(), "64x64x20 Gaussian hill, alpha = 1"
    if n == 1:
        return S.fractal_case(), "512x512x40 fractal terrain, alpha = 1"
    if n == 2:
        return S.fractal_case(varying_alpha=True), "512x512x40 fractal terrain, varying alpha"
    raise ValueError(n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, nargs="+", default=[0])
    ap.add_argument("--tol", type=float, nargs="+", default=[1e-8, 1e-12])
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    result = {"what": "relative L2 distance, FP64 path vs the oracle's FP32-faithful mode, same FP32 inputs",
              "threads": orc.num_threads(), "cases": {}}
    if a.out and os.path.exists(a.out):
        with open(a.out) as f:
            result["cases"].update(json.load(f).get("cases", {}))
    for n in a.config:
        c, label = _case(n)
        for tol in a.tol:
            t0 = time.time()
            d = faithful_distance(c, tol_rel=tol, max_iters=40000)
            d["label"], d["tol_rel"], d["seconds"] = label, tol, round(time.time() - t0, 1)
            key = f"config{n}_tol{tol:g}"
            result["cases"][key] = d
            print(key, json.dumps(d), flush=True)
            if a.out:
                with open(a.out, "w") as f:
                    json.dump(result, f, indent=1)


if __name__ == "__main__":
    main()
