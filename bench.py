"""Benchmark of the wind solve on one MI355X.

Workload: a 512x512x40 wind solve to relative residual 1e-8, dx = dy = 50 m,
seeded fractal terrain (k^-2 spectrum, 0-2500 m relief), 40 stretched levels
10 m -> 12 km, power-law winds 8 m/s at 10 m with 25 deg direction shear plus
N(0, 1 m/s) noise, alpha = 1, rho = 1. Inputs (metrics, alpha, divergence) are
resident in HBM before the timed region. One step = one complete solve
(operator refresh, diagonal, rhs, start selection, Krylov iterations, true
residual) from a zero lambda.

Arms (each solves the same system to rel. residual 1e-8; --mixed-tol sets
another stop for the mixed-precision arms):
  plain          the reference's BiCGStab in FP64 (2 Jacobi-Richardson sweeps)
  mixed          FP64 FGMRES around that BiCGStab in FP32
  line           FP64 FGMRES around an FP32 BiCGStab preconditioned by two
                 vertical line-relaxation steps
  mg             FP64 FGMRES around an FP32 BiCGStab preconditioned by one
                 multigrid V-cycle (horizontal semi-coarsening, coarse levels
                 rediscretised from averaged metrics, vertical line
                 smoothing; the fine level reads bf16 coefficient rows)
  mg_learned     `mg` started from the learned guess when its residual is
                 below the zero start's (the headline `value`: learned guess +
                 FP32/FP64 refinement)
  line_learned, mixed_learned   `line`, `mixed` with the learned guess
Errors (relative L2) of every arm are reported for lambda, the corrected u,
v, w and the divergence after the correction, against the plain FP64 solve
at the same tolerance and against a tight FP64 reference (the plain FP64
solve to rel. residual 1e-11, untimed). Per-kernel
device times come from hipEvents recorded around every hot launch on the
solver's stream (one extra profiled solve per arm). Every arm reports the
solve alone (`ms`, the `value`) and the solve followed by the u, v, w
correction (`ms_calc_iter_winds`, what calc_iter_winds does). The CPU baseline
is the reference algorithm restated in C (the oracle's source, built here with
-O3 -march=native -fopenmp -ffast-math) timed on this host on a bounded sample
and scaled to a full solve; --cpu-study adds the build without -ffast-math,
1-thread runs and one full solve. `faithful_distance` carries the relative L2
distance of the FP64 path from HICAR's own FP32 rounding (lambda, u, v, w,
divergence; written on the host by oracle/faithful.py into
profiles/faithful_distance.json).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --sequence S   (the 1024x1024x60 solve sequence instead)

--sequence S: S consecutive solves on a 1024x1024x60 grid (dx = 50 m, 3000 m
fractal relief, 60 levels 10 m -> 12 km) whose forcing wind turns 1 deg and is
rescaled by a seeded U[0.95, 1.05] factor each step (the mean power-law flow is
rotated on the device; its N(0, 1 m/s) noise stays fixed), each step solved
by the learned + FP32/FP64 multigrid solve to the mixed tolerance from three
starts: the network guess (used when its residual is below the previous
lambda's), the previous-step lambda, and zero. Prints one JSON line with
ms and iterations per start and the peak device memory in use.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms per wind solve to rel. residual 1e-8 at 512²×40: learned+FP32 vs plain FP64"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
WEIGHTS = os.path.join(ROOT, "hicar_amd", "data", "lnet_params.npy")
TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic.json")
# the FP64 path against the reference's FP32 rounding (written by oracle/faithful.py on the host)
FAITHFUL = os.path.join(ROOT, "profiles", "faithful_distance.json")


def kernel_bytes(kind: str, nx: int, ny: int, nz: int, coef_bytes: int = 0) -> float:
    """Algorithmic HBM bytes of one launch: every array the kernel must touch,
    once. Coefficients (B, C, H, I, L, M) are read on interior cells, BC planes
    (4 per column) once, vectors over all rows. coef_bytes: bytes per
    coefficient (or line factor) when they are stored narrower than the vectors
    (2: the bf16 rows of the multigrid inner solve)."""
    cells = nx * ny * nz
    P = (nx + 2) * (ny + 2)
    N = P * (nz + 2)
    t = 8 if kind.endswith("f64") or kind in ("ir_resid", "true_resid", "fg_matvec") else 4
    coef = 6 * (coef_bytes or t) * cells + 4 * t * P
    table = {
        "sweep": 3 * N * t + coef,        # y (stencil), in, coefficients -> out
        "sweep_sigma": 5 * N * t + coef,  # p_hat1 (stencil), p, r_hat, coefficients -> p_hat, v
        "sweep_omega": 5 * N * t + coef,  # s_hat1 (stencil), s, r_hat, coefficients -> s_hat, t
        "update_s": 5 * N * t,            # r, v, D^-1 -> s, s_hat
        "update_xrp": 12 * N * t,         # x, p_hat, s_hat, s, t, p, v, D^-1 -> x, r, p, p_hat
        "init": 7 * N * t + coef,         # x (stencil), b, D^-1, coefficients -> r, r_hat, p, p_hat
        "line_solve": 5.5 * N * t,        # in, lo, 1/pivot, cp -> out (+ y on every second call)
        "resid": 3 * N * t + coef,        # y (stencil), in, coefficients -> r
        "mv_sigma": 3 * N * t + coef,     # p_hat (stencil), r_hat, coefficients -> v
        "mv_omega": 4 * N * t + coef,     # s_hat (stencil), s, r_hat, coefficients -> t
    }
    if kind == "line_solve_f32" and coef_bytes:  # in, out, y on every second call in FP32; 3 factor rows narrower
        return N * (2.5 * 4 + 3 * coef_bytes)
    if kind == "ir_resid":    # lambda (stencil), b, FP64 coefficients -> v0
        return 3 * N * 8 + 6 * 8 * cells + 4 * 8 * P
    if kind == "ir_accumulate":  # lambda += sum of y_q z_q: lambda in/out + the FP32 directions (counted once)
        return N * 16 + N * 4
    if kind == "true_resid":
        return 2 * N * 8 + 6 * 8 * cells + 4 * 8 * P
    if kind == "fg_setup":    # v (FP64 in/out) -> x32, r32, rh32, p32, ph32 and D32^-1 (the Jacobi inner solve)
        return N * (16 + 4 * 6)
    if kind == "fg_matvec":   # z (FP64 widened), v0, coefficients -> w
        return 3 * N * 8 + 6 * 8 * cells + 4 * 8 * P
    if kind == "smooth_f32":  # y (stencil), in, coefficients, 3 factor rows -> out (the kernel re-reads y's
        fb = coef_bytes or 4  # centre for the update, mostly from cache: not counted)
        return 3 * N * 4 + 6 * fb * cells + 4 * 4 * P + 3 * fb * N
    if kind == "resid_restrict_f32":  # y (stencil), in, coefficients -> the coarse rhs (a quarter of the rows)
        fb = coef_bytes or 4
        return 2 * N * 4 + 6 * fb * cells + 4 * 4 * P + N
    if kind == "update_s_line_f32":  # r, v -> s
        return 3 * N * 4
    if kind == "update_xrp_line_f32":  # x, p_hat, s_hat, s, t, p, v -> x, r, p
        return 10 * N * 4
    if kind == "mg_transfer_f32":  # restriction (r -> b_1) and prolongation (y += P e_1) alternate
        return 0.5 * (1.25 + 2.25) * N * 4
    if kind in ("mg_coarse_f32", "mg_setup", "fg_gram_schmidt"):  # no single size per call
        return None
    base = kind.rsplit("_", 1)[0]
    return float(table[base])


def gram_schmidt_bytes(nx: int, ny: int, nz: int, steps: int, restart: int = 32) -> float:
    """Average algorithmic bytes of the Gram-Schmidt region of one FGMRES step
    over `steps` consecutive steps (cycles of `restart` steps): step j reads w
    and v_1..v_j for the inner products (none at j = 0; h_0 comes with the
    matvec), then w and v_0..v_j, writing w, for the update and ||w||^2."""
    N = (nx + 2) * (ny + 2) * (nz + 2)
    tot = 0
    for q in range(steps):
        j = q % restart
        tot += ((j + 1) if j else 0) + (j + 3)
    return 8.0 * N * tot / steps


def coarse_cycle_model(levels, cpre=1, cpost=1, wave_cols=4608, gal_from=2):
    """Algorithmic bytes and launches of the coarse part of one V-cycle (the
    default MIXED_MG cycle: V(1,1) on the coarse levels, automatic depth,
    Galerkin 27-point operators in FP32 from multigrid level gal_from on (the
    coarse levels are numbered from 1), fused transfers between the
    wavefront-per-column levels, one-pass steps with bf16 rows and factors on
    the larger rediscretised levels, FP32 on the wavefront levels). levels:
    (nx, ny, mz) of the coarse levels, coarsest last. Each step reads what it
    must once: from zero, the rhs and the factors; with an iterate, the
    iterate, the rhs, the coefficient rows (six and the four BC planes, or the
    27 Galerkin rows) and the three factor rows; every step writes its iterate;
    a descent reads the finer iterate, rhs and rows and writes the coarser rhs
    (and, fused, the coarser first step); a return reads the coarser
    correction."""
    geo = []
    for q, (nx, ny, mz) in enumerate(levels, start=1):
        P = (nx + 2) * (ny + 2)
        gal = 0 < gal_from <= q and mz <= 64
        wave = (P <= wave_cols or gal) and mz <= 64
        fb = 4 if wave else 2
        # bytes per row of the operator a step or a residual reads, BC planes per column
        rows, bc = (27 * 4, 0) if gal else (6 * fb, 16)
        geo.append(dict(P=P, N=P * mz, fb=fb, wave=wave, small=min(nx, ny), rows=rows, bc=bc))
    used = 0
    while used < len(geo) and geo[used]["small"] >= 4:
        used += 1
    used = max(used, 1)
    tot = dict(bytes=0.0, launches=0)

    def add(b, n=1):
        tot["bytes"] += b
        tot["launches"] += n

    def zero_step(g):
        add((8 + 3 * g["fb"]) * g["N"])

    def step(g):
        add((12 + g["rows"] + 3 * g["fb"]) * g["N"] + g["bc"] * g["P"])

    def cycle(l, entered):
        g = geo[l]
        if l + 1 == len(geo):
            if not entered:
                zero_step(g)
            return
        if not entered:
            zero_step(g)
        if l + 1 == used:
            for _ in range(cpre + cpost - 1):
                step(g)
            return
        for _ in range(cpre - 1):
            step(g)
        c = geo[l + 1]
        fused = g["wave"] and c["wave"]
        desc = (8 + g["rows"]) * g["N"] + g["bc"] * g["P"] + 4 * c["N"]
        add(desc + ((4 + 3 * c["fb"]) * c["N"] if fused else 0))
        cycle(l + 1, fused)
        if fused:
            step(g)
            add(4 * c["N"], 0)
            post = cpost - 1
        else:
            add(4 * c["N"] + 8 * g["N"])
            post = cpost
        for _ in range(post):
            step(g)

    cycle(0, False)
    return tot


def build_case(nx, ny, nz):
    from hicar_amd import synthetic as S
    return S.fractal_case(nx=nx, ny=ny, nz=nz, dx=50.0, relief=2500.0, speed=8.0, seed=1)


def cpu_info() -> dict:
    """The host's CPU model and sizes (lscpu) and the cores this process may use."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        sockets, per = kv.get("Socket(s)"), kv.get("Core(s) per socket")
        if sockets and per:
            info["host_physical_cores"] = int(sockets) * int(per)
        info["host_logical_cpus"] = int(kv["CPU(s)"]) if "CPU(s)" in kv else None
        info["threads_per_core"] = int(kv["Thread(s) per core"]) if "Thread(s) per core" in kv else None
    except (OSError, ValueError, KeyError, subprocess.SubprocessError):
        pass
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    return info


def cpu_baseline(case, div_host, iters_full: int, sample_iters: int = 50, samples: int = 5,
                 variants=("fastmath",), thread_counts=(None,), full_solve=False, log=None):
    """The reference algorithm (the oracle's C restatement: the same assembly,
    preconditioner, BiCGStab and stop rule) timed on this host. It is built
    here with -O3 -march=native -fopenmp (variant "native") or that plus
    -ffast-math ("fastmath", the reference's release flags), every vector and
    preconditioner loop threaded and dot products as plain parallel sums.
    Per variant and thread count: the median over `samples` runs of the
    first-call set-up (alpha, assembly, diagonal, rhs), of `sample_iters`
    BiCGStab iterations and of the correction; the solve is set-up +
    iters_full iterations (bicgstab_solve's share of calc_iter_winds), and
    calc_iter_winds adds the correction. full_solve: also one solve to the
    1e-8 stop with all threads, to check the extrapolation."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    out_dir = tempfile.mkdtemp(prefix="oracle_baseline_")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "baseline", f"BASE_OUT={out_dir}"], check=True)
    libs = {"native": os.path.join(out_dir, "liboracle_native.so"),
            "fastmath": os.path.join(out_dir, "liboracle_native_fastmath.so")}
    zeros = np.zeros((case.ny + 2, case.nz + 2, case.nx + 2))
    res = {}
    for var in variants:
        o = orc.Oracle(case.nx, case.ny, case.nz, case.dx, lib_path=libs[var])
        o.set_metrics(case.metrics)
        for nt in thread_counts:
            o.lib.orc_set_threads(int(nt) if nt else len(os.sched_getaffinity(0)))
            threads = o.lib.orc_num_threads()
            su, it, co = [], [], []
            for _ in range(samples):
                o.set_lambda(zeros)
                t0 = time.perf_counter()
                o.set_alpha(case.alpha)
                o.assemble()
                o.extract_diagonal()
                o.compute_rhs(div_host)
                t1 = time.perf_counter()
                o.bicgstab(max_iters=sample_iters, tol_rel=1e-30)
                t2 = time.perf_counter()
                o.correct(case.u, case.v, case.w, rho=case.rho, adv_den=True)
                t3 = time.perf_counter()
                su.append(t1 - t0), it.append((t2 - t1) / sample_iters), co.append(t3 - t2)
            setup, per_iter, corr = (float(np.median(a)) for a in (su, it, co))
            r = dict(threads=threads, setup_ms=round(1e3 * setup, 2), ms_per_iter=round(1e3 * per_iter, 3),
                     correction_ms=round(1e3 * corr, 2), solve_ms=round(1e3 * (setup + per_iter * iters_full), 1),
                     calc_iter_winds_ms=round(1e3 * (setup + per_iter * iters_full + corr), 1),
                     iters_spread=[round(1e3 * min(it), 3), round(1e3 * max(it), 3)])
            res[f"{var}_t{threads}"] = r
            if log:
                log(f"cpu baseline {var}, {threads} threads: {r}")
        if full_solve:
            o.lib.orc_set_threads(len(os.sched_getaffinity(0)))
            o.set_lambda(zeros)
            t0 = time.perf_counter()
            o.set_alpha(case.alpha)
            o.assemble()
            o.extract_diagonal()
            o.compute_rhs(div_host)
            rr = o.bicgstab(max_iters=20000, tol_rel=1e-8)
            t1 = time.perf_counter()
            res[f"{var}_full_solve"] = dict(threads=o.lib.orc_num_threads(), ms=round(1e3 * (t1 - t0), 1),
                                            iters=rr["iters"], status=rr["status"])
            if log:
                log(f"cpu baseline {var}, full solve: {res[f'{var}_full_solve']}")
        o.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--inner-tol", type=float, default=0.5)
    ap.add_argument("--arms", default="plain,mixed,line,mg,mg_learned")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=50, help="BiCGStab iterations per CPU-baseline sample")
    ap.add_argument("--cpu-samples", type=int, default=5, help="CPU-baseline samples (median)")
    ap.add_argument("--cpu-study", action="store_true",
                    help="CPU baseline with and without -ffast-math, at all threads and at 1, and one full solve")
    ap.add_argument("--mixed-tol", type=float, default=1e-8,
                    help="rel. residual the mixed-precision arms stop at (default the metric's 1e-8)")
    ap.add_argument("--tight", type=float, default=1e-11,
                    help="rel. residual of the FP64 reference solve the errors are stated against")
    ap.add_argument("--no-tight", action="store_true", help="skip the tight FP64 reference solve")
    ap.add_argument("--no-varying-alpha", action="store_true",
                    help="skip the mixed arms on the spatially varying alpha of the mixed-precision configuration")
    ap.add_argument("--sequence", type=int, default=0,
                    help="run S consecutive solves of the 1024x1024x60 sequence instead of the benchmark")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="no-op kept for the profiling commands: every run first copies a 1 GiB device buffer "
                         "three times (the copy ceiling in `roofline`, and known traffic for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is described in DESIGN.md, not implemented).", file=sys.stderr)
        sys.exit(2)

    import torch
    from hicar_amd import _native as nat
    from hicar_amd import learned
    from hicar_amd.wind_iterative import WindSolver

    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    dev = torch.device("cuda", 0)
    if args.sequence > 0:
        run_sequence(args, log, dev)
        return
    t_gen = time.perf_counter()
    case = build_case(args.nx, args.ny, args.nz)
    log(f"case {case.nx}x{case.ny}x{case.nz} generated in {time.perf_counter() - t_gen:.1f}s")

    def D(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

    s = WindSolver(case.nx, case.ny, case.nz, case.dx, device_fields=True)
    s.set_metrics({k: D(v) for k, v in case.metrics.items()
                   if k in ("dzdx", "dzdy", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w", "dzdx_u",
                            "dzdy_v", "advection_dz", "terrain")})
    u, v, w, rho, alpha = D(case.u), D(case.v), D(case.w), D(case.rho), D(case.alpha)
    div = torch.empty(s.mass_shape, dtype=torch.float64, device=dev)
    s.divergence(u, v, w, density=rho, adv_den=True, use_dqdt=True, out=div)
    torch.cuda.synchronize()
    # the HBM rate a plain device copy reaches on this GPU (1 GiB read + 1 GiB
    # written, best of 3, events on torch's stream): the practical ceiling of
    # the streaming kernels, reported beside the published peak
    src = torch.ones(2 ** 27, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    copy_ms = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        copy_ms.append(e0.elapsed_time(e1))
    del src, dst
    copy_gbs = 2 * 2 ** 30 / (min(copy_ms) * 1e-3) / 1e9
    log(f"device copy of 1 GiB: {min(copy_ms):.3f} ms, {copy_gbs:.0f} GB/s"
        + (" (also the calibration of the PMC traffic passes)" if args.copy_calibration else ""))

    net = learned.Net.load(WEIGHTS) if os.path.exists(WEIGHTS) else None
    arms = {
        "plain": dict(precision=nat.PLAIN_F64, start=nat.START_ZERO),
        "mixed": dict(precision=nat.MIXED, start=nat.START_ZERO),
        "line": dict(precision=nat.MIXED_LINE, start=nat.START_ZERO),
        "line_learned": dict(precision=nat.MIXED_LINE, start=nat.START_LEARNED),
        "mixed_learned": dict(precision=nat.MIXED, start=nat.START_LEARNED),
        "mg": dict(precision=nat.MIXED_MG, start=nat.START_ZERO, inner_tol=0.3),
        "mg_learned": dict(precision=nat.MIXED_MG, start=nat.START_LEARNED, inner_tol=0.3),
    }
    want = [a for a in args.arms.split(",") if a in arms and (net is not None or "learned" not in a)]
    zeros = np.zeros(s.rows_shape)

    def inner_tol_of(arm):
        return arms[arm].get("inner_tol", args.inner_tol) if arms[arm]["precision"] != nat.PLAIN_F64 else None

    def solve(arm, profile=0, tol=None, a=None, correct=False):
        """One solve from a zero lambda (ms); with correct, also the solve
        followed by the u, v, w correction of calc_iter_winds (ms)."""
        o = dict(arms[arm])
        inner = o.pop("inner_tol", args.inner_tol)
        if tol is None:
            tol = 1e-8 if o["precision"] == nat.PLAIN_F64 else args.mixed_tol
        learned.attach(s, net if o["start"] == nat.START_LEARNED else None)
        s.set_lambda(zeros)
        if correct:
            uu, vv, ww = u.clone(), v.clone(), w.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = s.solve(alpha if a is None else a, div, tol_rel=tol, max_iters=50000, inner_tol=inner,
                       profile=profile, **o)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if not correct:
            return 1e3 * (t1 - t0), info
        s.apply_correction(uu, vv, ww, density=rho, adv_den=True)
        torch.cuda.synchronize()
        return 1e3 * (t1 - t0), info, 1e3 * (time.perf_counter() - t0)

    def winds():
        uu, vv, ww = u.clone(), v.clone(), w.clone()
        s.apply_correction(uu, vv, ww, density=rho, adv_den=True)
        dd = torch.empty_like(div)
        s.divergence(uu, vv, ww, density=rho, adv_den=True, use_dqdt=True, out=dd)
        return [t.cpu().numpy() for t in (uu, vv, ww, dd)]

    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))  # noqa: E731

    def errs(lam, wnd, ref):
        return dict(lam=rel(lam, ref[0]), u=rel(wnd[0], ref[1][0]), v=rel(wnd[1], ref[1][1]),
                    w=rel(wnd[2], ref[1][2]), div=rel(wnd[3], ref[1][3]))

    # Tight FP64 reference (untimed): the plain FP64 solve to rel. residual
    # TIGHT, against which every arm's solution error is stated.
    tight = None
    if not args.no_tight:
        ms_t, info_t = solve("plain", tol=args.tight)
        tight = (s.get_lambda(), winds())
        tight_info = dict(tol_rel=args.tight, iters=info_t["iters"], status=info_t["status"],
                          true_rel_res=info_t["true_rel_res"], ms=round(ms_t, 1))
        log(f"tight FP64 reference: {tight_info}")

    res = {}
    ref = None
    for arm in want:
        for _ in range(args.warmup):
            solve(arm)
        times, times_ciw, info = [], [], None
        for _ in range(args.steps):
            ms, info, ms_ciw = solve(arm, correct=True)
            times.append(ms)
            times_ciw.append(ms_ciw)
        lam = s.get_lambda()
        wnd = winds()
        s.profile(reset=True)
        solve(arm, profile=1)
        prof = s.profile(reset=True)
        r = dict(ms=float(np.mean(times)), ms_calc_iter_winds=round(float(np.mean(times_ciw)), 3),
                 iters=info["iters"], outer=info["outer_iters"], status=info["status"],
                 tol_rel=1e-8 if arms[arm]["precision"] == nat.PLAIN_F64 else args.mixed_tol,
                 inner_tol=inner_tol_of(arm),
                 true_rel_res=info["true_rel_res"], start_used=info["start_used"], prof=prof)
        if arm == "plain":
            ref = (lam, wnd)
        elif ref is not None:
            r["err_vs_plain"] = errs(lam, wnd, ref)
        if tight is not None:
            r["err_vs_tight"] = errs(lam, wnd, tight)
        if "learned" in arm:
            r.update(guess_res=info["guess_res"], plain_res=info["plain_res"])
        res[arm] = r
        log(f"{arm}: {r['ms']:.1f} ms/solve iters={r['iters']} outer={r['outer']} status={r['status']} "
            f"true_rel_res={r['true_rel_res']:.3e} start_used={r['start_used']}")

    # The mixed-precision configuration with a spatially varying stability
    # weight: the same terrain and winds, alpha in [0.3, 1] from a seeded
    # lapse-rate field smoothed over 2 km; errors of the mixed arms against a
    # tight FP64 solve and the plain FP64 solve of the same alpha.
    vary = None
    if not args.no_varying_alpha:
        from hicar_amd import synthetic as S
        a2 = D(S.fractal_case(nx=args.nx, ny=args.ny, nz=args.nz, dx=50.0, relief=2500.0, speed=8.0, seed=1,
                              varying_alpha=True).alpha)
        vary = dict(alpha_min=round(float(a2.min()), 4), alpha_max=round(float(a2.max()), 4), arms={})
        ms_t, info_t = solve("plain", tol=args.tight, a=a2)
        tight2 = (s.get_lambda(), winds())
        vary["tight_reference"] = dict(tol_rel=args.tight, iters=info_t["iters"], status=info_t["status"],
                                       true_rel_res=info_t["true_rel_res"])
        ref2 = None
        for arm in [x for x in ("plain", "mixed", "mg", "mg_learned") if x in want]:
            solve(arm, a=a2)
            ms2, info2 = solve(arm, a=a2)
            lam2, wnd2 = s.get_lambda(), winds()
            r2 = dict(ms=round(ms2, 2), iters=info2["iters"], outer=info2["outer_iters"], status=info2["status"],
                      true_rel_res=info2["true_rel_res"], err_vs_tight=errs(lam2, wnd2, tight2))
            if arm == "plain":
                ref2 = (lam2, wnd2)
            elif ref2 is not None:
                r2["err_vs_plain"] = errs(lam2, wnd2, ref2)
            vary["arms"][arm] = r2
            log(f"varying alpha, {arm}: {r2['ms']:.1f} ms/solve iters={r2['iters']} "
                f"lambda err vs tight {r2['err_vs_tight']['lam']:.2e}")

    head = next((a for a in ("mg_learned", "mg", "line_learned", "line") if a in res), next(iter(res)))
    ms_h = res[head]["ms"]

    def kernel_table(prof, arm):
        out = {}
        for kind, (calls, tot) in prof.items():
            if arm.startswith(("line", "mg")) and kind in ("update_s_f32", "update_xrp_f32"):
                kind = kind[:-4] + "_line_f32"  # without the Jacobi sweep's D^-1 terms
            avg = tot / calls
            # the multigrid inner solve reads bf16 coefficient rows
            cb = 2 if arm.startswith("mg") and kind in ("resid_f32", "mv_sigma_f32", "mv_omega_f32",
                                                         "line_solve_f32", "smooth_f32",
                                                         "resid_restrict_f32") else 0
            nb = kernel_bytes(kind, case.nx, case.ny, case.nz, cb)
            if kind == "fg_setup" and arm.startswith(("line", "mg")):
                # the line and multigrid set-up writes r_hat alone (the first inner
                # iteration reads r and p from it): v (FP64 in/out) -> rh32
                N = (case.nx + 2) * (case.ny + 2) * (case.nz + 2)
                nb = N * (16 + 4)
            if kind == "fg_gram_schmidt":
                nb = gram_schmidt_bytes(case.nx, case.ny, case.nz, int(calls))
            if kind == "mg_coarse_f32" and coarse is not None:
                nb = coarse["bytes"]
            gbs = nb / (avg * 1e-3) / 1e9 if nb else None
            out[kind] = dict(calls=int(calls), avg_ms=round(avg, 5), total_ms=round(tot, 2), bytes=nb,
                             GBs=round(gbs, 1) if gbs else None, frac=round(gbs / HBM_PEAK_GBS, 3) if gbs else None)
            if kind == "mg_coarse_f32" and coarse is not None:
                out[kind].update(
                    launches_per_call=coarse["launches"], us_per_launch=round(1e3 * avg / coarse["launches"], 2),
                    note=("one call = the coarse part of one V-cycle (every launch from the fine restriction to "
                          "the fine prolongation, timed as one region); bytes = the cycle's algorithmic bytes "
                          "(coarse_cycle_model). Latency-bound: the levels below 130 x 130 columns hold about a "
                          "microsecond of work per launch"))
            if kind in ("update_s_line_f32", "update_xrp_line_f32"):
                out[kind]["note"] = ("bytes of the full update: the closing iteration of each inner solve "
                                     "updates only x (and r), and the vectors written just before are partly "
                                     "served from the 256 MB Infinity Cache, so frac can exceed 1")
        return out

    coarse = None
    if any(a.startswith("mg") for a in res):
        import ctypes
        lv, q = [], 1
        dims = (ctypes.c_int * 3)()
        while s.lib.windhip_mg_level(s.h, q, dims, None) == 0:  # the hierarchy of the last MIXED_MG solve
            lv.append((dims[0], dims[1], dims[2]))
            q += 1
        coarse = coarse_cycle_model(lv) if lv else None
    tables = {arm: kernel_table(r.pop("prof"), arm) for arm, r in res.items()}
    kh = tables[head]
    # the dominant single kernel (the coarse-level kind mixes every kernel of
    # every coarse level: a region of many launches, reported on its own below)
    dom = max((k for k in kh if kh[k]["GBs"] and k != "mg_coarse_f32"), key=lambda k: kh[k]["total_ms"])
    d = kh[dom]
    traffic = None
    if os.path.exists(TRAFFIC):
        with open(TRAFFIC) as f:
            tr = json.load(f).get("kernels", {}).get(dom)
        if tr:
            traffic = tr["bytes_per_launch"]
    roof = dict(bound="hbm", achieved=d["GBs"], peak=HBM_PEAK_GBS, unit="GB/s", frac=d["frac"], traffic=traffic,
                kernel=dom, bytes_per_launch=d["bytes"], avg_ms=d["avg_ms"],
                copy_ceiling_GBs=round(copy_gbs, 1), frac_of_copy=round(d["GBs"] / copy_gbs, 3))
    # the coarse part of the headline arm's V-cycles: one timed region per cycle
    # (many launches), against its algorithmic bytes
    coarse_roof = None
    if "mg_coarse_f32" in kh and kh["mg_coarse_f32"]["GBs"]:
        cr = kh["mg_coarse_f32"]
        coarse_roof = dict(region="mg_coarse_f32", achieved=cr["GBs"], peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=cr["frac"], avg_ms=cr["avg_ms"], total_ms=cr["total_ms"],
                           launches_per_call=cr.get("launches_per_call"), us_per_launch=cr.get("us_per_launch"),
                           bytes_per_call=cr["bytes"])
    # the plain FP64 solve's matvec (fused with the last Jacobi sweep: v = A p_hat and sigma)
    plain_mv = None
    if "plain" in tables and "sweep_sigma_f64" in tables["plain"]:
        pm = tables["plain"]["sweep_sigma_f64"]
        plain_mv = dict(kernel="sweep_sigma_f64", achieved=pm["GBs"], peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=pm["frac"], frac_of_copy=round(pm["GBs"] / copy_gbs, 3), avg_ms=pm["avg_ms"],
                        bytes_per_launch=kernel_bytes("sweep_sigma_f64", case.nx, case.ny, case.nz))

    netinfo = None
    if net is not None:
        fms = net.time_forward(case.nz, case.ny, case.nx, reps=3)
        fl = learned.conv_flops(case.nz, case.ny, case.nx)
        netinfo = dict(forward_ms=round(fms, 3), forward_tflop=round(fl / 1e12, 3),
                       tflops=round(fl / fms / 1e9, 1), mfma_util=round(fl / fms / 1e9 / BF16_PEAK_TFLOPS, 4))

    cpu = None
    if not args.no_cpu and "plain" in res:
        iters_full = res["plain"]["iters"]
        study = args.cpu_study
        cres = cpu_baseline(case, div.cpu().numpy(), iters_full, sample_iters=args.cpu_iters,
                            samples=args.cpu_samples, variants=("fastmath", "native") if study else ("fastmath",),
                            thread_counts=(None, 1) if study else (None,), full_solve=study, log=log)
        main_key = next(k for k in cres if k.startswith("fastmath_t") and cres[k]["threads"] > 1) \
            if any(k.startswith("fastmath_t") and cres[k]["threads"] > 1 for k in cres) else next(iter(cres))
        m = cres[main_key]
        cpu = dict(value=m["solve_ms"], unit="ms/solve", cores=m["threads"], kind="port",
                   sample=f"the reference algorithm restated in C (oracle/wind_oracle.c, -O3 -march=native "
                          f"-fopenmp -ffast-math, threaded loops, parallel dot products) at {case.nx}x{case.ny}x"
                          f"{case.nz}: median of {args.cpu_samples} runs of the first-call set-up and "
                          f"{args.cpu_iters} BiCGStab iterations, scaled to the {iters_full} iterations of the "
                          f"plain FP64 solve", calc_iter_winds_ms=m["calc_iter_winds_ms"],
                   ms_per_iter=m["ms_per_iter"], setup_ms=m["setup_ms"], correction_ms=m["correction_ms"],
                   cpu=cpu_info(), runs=cres)
        log(f"cpu baseline: {cpu['value']:.0f} ms/solve ({cpu['cores']} threads)")

    faithful = None
    if os.path.exists(FAITHFUL):
        with open(FAITHFUL) as f:
            fd = json.load(f)
        faithful = {"what": fd["what"],
                    **{k: {q: v[q] for q in ("label", "tol_rel", "lam", "u", "v", "w", "div")}
                       for k, v in fd["cases"].items()}}
    plain_ms = res["plain"]["ms"] if "plain" in res else None
    # what the learned start costs or saves: the same multigrid solve from the
    # zero start beside it, and the network's forward pass alone
    learned_cost = None
    if "mg" in res and "mg_learned" in res:
        learned_cost = dict(mg_ms=round(res["mg"]["ms"], 3), mg_learned_ms=round(res["mg_learned"]["ms"], 3),
                            net_cost_ms=round(res["mg_learned"]["ms"] - res["mg"]["ms"], 3),
                            iters_mg=res["mg"]["iters"], iters_mg_learned=res["mg_learned"]["iters"],
                            outer_mg=res["mg"]["outer"], outer_mg_learned=res["mg_learned"]["outer"],
                            start_used=res["mg_learned"]["start_used"],
                            network_forward_ms=netinfo["forward_ms"] if netinfo else None)
    line = {
        "metric": METRIC,
        "value": round(ms_h, 3),
        "ms_calc_iter_winds": res[head]["ms_calc_iter_winds"],
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_h, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32+bf16",
        "dtype_detail": "FP64 outer FGMRES, residual and lambda; FP32 inner BiCGStab and multigrid arithmetic, "
                        "the fine-level coefficient rows and line factors stored as bf16 (widened to FP32 on "
                        "load); network bf16 MFMA with FP32 accumulation",
        "data": "synthetic",
        "config": {"workload": f"wind solve {case.nx}x{case.ny}x{case.nz} fractal terrain, rel. residual 1e-8, "
                               f"arm '{head}'", "grid": [case.nx, case.ny, case.nz], "dx_m": case.dx,
                   "tol_rel": 1e-8, "tol_rel_mixed_arms": args.mixed_tol,
                   "inner_tol": {a: inner_tol_of(a) for a in res},
                   "cells": case.cells,
                   "learned_weights": os.path.relpath(WEIGHTS, ROOT) if net is not None else None},
        "roofline": roof,
        "coarse_cycle_roofline": coarse_roof,
        "plain_matvec_roofline": plain_mv,
        "cpu_baseline": cpu,
        "speedup_vs_plain": round(plain_ms / ms_h, 3) if plain_ms else None,
        "learned_start_cost": learned_cost,
        "faithful_distance": faithful,
        "cells_per_s": case.cells / (ms_h * 1e-3),
        "tight_reference": tight_info if tight is not None else None,
        "arms": res,
        "varying_alpha": vary,
        "network": netinfo,
        "kernels": tables,
    }
    print(json.dumps(line))


def run_sequence(args, log, dev):
    """The solve sequence of the 1024x1024x60 case (see the module docstring)."""
    import torch
    from hicar_amd import _native as nat
    from hicar_amd import learned
    from hicar_amd import synthetic as S
    from hicar_amd.wind_iterative import WindSolver

    t_gen = time.perf_counter()
    case = S.fractal_case(nx=1024, ny=1024, nz=60, dx=50.0, relief=3000.0, speed=8.0, seed=4)
    log(f"case 1024x1024x60 generated in {time.perf_counter() - t_gen:.1f}s")
    m = case.metrics
    nz = case.nz

    def D(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

    s = WindSolver(case.nx, case.ny, case.nz, case.dx, device_fields=True)
    s.set_metrics({k: D(v) for k, v in m.items()
                   if k in ("dzdx", "dzdy", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w", "dzdx_u",
                            "dzdy_v", "advection_dz", "terrain")})
    # the mean power-law flow (25 deg shear over the column) and the fixed noise
    theta0 = torch.tensor(np.deg2rad(25.0 * np.arange(nz) / (nz - 1)), device=dev)[None, :, None]
    su = D(8.0 * (np.maximum(m["z_agl_u"], 0.1) / 10.0) ** (1.0 / 7.0))
    sv = D(8.0 * (np.maximum(m["z_agl_v"], 0.1) / 10.0) ** (1.0 / 7.0))
    nu = D(case.u) - su * torch.cos(theta0)
    nv = D(case.v) - sv * torch.sin(theta0)
    w, rho, alpha = D(case.w), D(case.rho), D(case.alpha)
    div = torch.empty(s.mass_shape, dtype=torch.float64, device=dev)
    scales = np.random.default_rng(7).uniform(0.95, 1.05, size=args.sequence)
    net = learned.Net.load(WEIGHTS)
    starts = {"learned": nat.START_LEARNED, "warm": nat.START_WARM, "zero": nat.START_ZERO}
    peak = handle_peak = 0.0
    out = {}
    for name, start in starts.items():
        learned.attach(s, net if start == nat.START_LEARNED else None)
        s.set_lambda(np.zeros(s.rows_shape))
        ms, its, worst = [], [], 0.0
        scale = 1.0
        for step in range(args.sequence):
            scale *= scales[step]
            th = theta0 + np.deg2rad(float(step))
            u = scale * su * torch.cos(th) + nu
            v = scale * sv * torch.sin(th) + nv
            s.divergence(u, v, w, density=rho, adv_den=True, use_dqdt=True, out=div)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = s.solve(alpha, div, precision=nat.MIXED_MG, start=start, tol_rel=args.mixed_tol,
                           max_iters=3000, inner_tol=0.3)
            torch.cuda.synchronize()
            ms.append(1e3 * (time.perf_counter() - t0))
            its.append(info["iters"])
            log(f"{name} step {step}: {ms[-1]:.1f} ms, {info['iters']} iterations, {info['outer_iters']} FGMRES steps, "
                f"true rel. residual {info['true_rel_res']:.2e}")
            worst = max(worst, info["true_rel_res"])
            if info["status"] != 0:
                raise RuntimeError(f"sequence step {step} ({name} start) did not converge: {info}")
            peak = max(peak, s.memory()[1])
            handle_peak = max(handle_peak, s.memory()[0])
        out[name] = dict(ms_total=round(sum(ms), 1), ms_mean=round(float(np.mean(ms)), 2),
                         iters_total=int(sum(its)), iters_first=its[0], iters_mean=round(float(np.mean(its)), 2),
                         worst_true_rel_res=worst)
        log(f"{name}: {out[name]}")
    line = {"metric": "ms per wind solve in a 1024x1024x60 sequence", "steps": args.sequence,
            "config": {"workload": f"{args.sequence} wind solves 1024x1024x60 fractal terrain (3000 m), forcing "
                                   "turned 1 deg and rescaled U[0.95, 1.05] per step", "tol_rel": args.mixed_tol,
                       "method": "MIXED_MG (FP64 FGMRES, FP32 multigrid-preconditioned BiCGStab)"},
            "starts": out, "peak_device_bytes_in_use": peak, "solver_handle_bytes": handle_peak,
            "torch_peak_bytes": int(torch.cuda.max_memory_allocated())}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
