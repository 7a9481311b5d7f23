"""Benchmark: implicit filament timesteps/s at 256^3 with 30,000 segments.

Workload (one GPU): 1,000 swimming filaments x 30 segments in a periodic
256^3 box (dx = a/1.86, ngd = 12), travelling-wave preferred curvature
(amplitude 0.5/L, wavenumber 2 pi/L, period 1), bending and twist moduli
1000 (the reference swimmer modulus, config.hpp:100), dt = 0.005, barrier on,
BDF2 + bad Broyden with TOL 1e-6. A step is one whole implicit timestep
(initial guess, forces, Jacobian, cell list, barrier, Broyden iterations of
mobility solves, state step), all resident in HBM.

Reported `value`: timesteps/s of the mixed-precision path (FCM stages in
FP32, Broyden/tensions/BDF2 state in FP64). Also reported: the plain FP64
path on the same system, mobility solves/s of the 256^3 / 30,000-segment
FCM microbenchmark in FP64 and FP32, and the roofline of the dominant
kernel measured with HIP events on the solver's own stream.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_system(precision, nfil=1000, nw=30, grid=256, seed=20261202, kap=1000.0,
                 tol=1e-6, dt=0.005):
    import numpy as np
    from filsim_amd import api, params
    fp = params.fcm_params(grid)
    sp = params.sim_params(fp, dt=dt, tol=tol, maxiter=100,
                           precision=precision, weight=(0.0, 0.0, 0.0),
                           barrier_on=1)
    rng = np.random.default_rng(seed)
    props = params.swimming_props(nfil, nw, kap=kap,
                                  phases=list(rng.uniform(0, 2 * math.pi,
                                                          nfil)))
    X, q = api.make_filaments(params.box(fp), [nw] * nfil, dL=2.2,
                              min_sep=2.2, seed=seed)
    s = api.System(sp, [nw] * nfil, props)
    s.set_state(X=X, Xs=X, Xt=X, q=q, qs=q, qt=q)
    return s, fp


def time_steps(s, warmup, steps, torch):
    its = []
    nt = 0
    for _ in range(warmup):
        nt += 1
        s.timestep(nt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        nt += 1
        its.append(s.timestep(nt))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, its, nt


def mobility_bench(precision, reps, torch, np_=30000, grid=256, seed=7):
    import numpy as np
    from filsim_amd import api, params
    fp = params.fcm_params(grid)
    rng = np.random.default_rng(seed)
    L = np.array(params.box(fp))
    Y = torch.tensor(rng.uniform(0, 1, (np_, 3)) * L, dtype=torch.float64,
                     device="cuda")
    F = torch.tensor(rng.standard_normal((np_, 3)), dtype=torch.float64,
                     device="cuda")
    T = torch.tensor(rng.standard_normal((np_, 3)), dtype=torch.float64,
                     device="cuda")
    V = torch.zeros_like(Y)
    W = torch.zeros_like(Y)
    fcm = api.Fcm(fp, np_, precision=precision)
    for _ in range(3):
        fcm.mobility(Y, F, T, V, W)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fcm.mobility(Y, F, T, V, W)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    # per-stage device times (HIP events on the solver's stream)
    fcm.set_timing(True)
    stages = np.zeros(7)
    nrep = 10
    for _ in range(nrep):
        fcm.mobility(Y, F, T, V, W)
        stages += fcm.stage_times()
    fcm.set_timing(False)
    stages /= nrep
    fcm.close()
    return t, stages, fp


def cpu_baseline(fp_iters):
    """FP64 oracle (CPU restatement, one thread) on one 256^3 / 30,000-
    segment mobility solve, scaled to timesteps/s with the GPU FP64 path's
    mean Broyden iterations per step (each step = iterations + 1 solves)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from filsim_amd import params
    fp = params.fcm_params(256)
    rng = np.random.default_rng(7)
    L = np.array(params.box(fp))
    Y = rng.uniform(0, 1, (30000, 3)) * L
    F = rng.standard_normal((30000, 3))
    T = rng.standard_normal((30000, 3))
    t0 = time.perf_counter()
    O.mobility(fp, Y, F, T)
    t_solve = time.perf_counter() - t0
    solves_per_step = fp_iters + 1
    return t_solve, 1.0 / (t_solve * solves_per_step), solves_per_step


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mob-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "mixed", "fp64", "mobility"],
                    default="all")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the FCM grid and every "
              "filament of this workload live on one MI355X; multi-GPU "
              "sharding is a design note, not a measured path.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(2)
    from filsim_amd import abi
    out = {}

    # FCM mobility microbenchmark (256^3, 30,000 segments)
    t64, st64, fp = mobility_bench(abi.FS_FP64, args.mob_reps, torch)
    t32, st32, _ = mobility_bench(abi.FS_FP32, args.mob_reps, torch)
    log("mobility fp64 %.3f ms (%.1f solves/s), fp32 %.3f ms (%.1f/s)" %
        (1e3 * t64, 1 / t64, 1e3 * t32, 1 / t32))
    log("fp64 stages ms", [round(float(x), 4) for x in st64])
    log("fp32 stages ms", [round(float(x), 4) for x in st32])
    N = 256 ** 3
    out["mobility"] = {
        "fp64_solves_per_s": 1 / t64, "fp32_solves_per_s": 1 / t32,
        "fp64_ms": 1e3 * t64, "fp32_ms": 1e3 * t32,
        "fp64_algorithmic_GBps": 24 * N * 8 / t64 / 1e9,
        "fp32_algorithmic_GBps": 24 * N * 4 / t32 / 1e9,
        "fp64_stage_ms": [float(x) for x in st64],
        "fp32_stage_ms": [float(x) for x in st32]}
    # dominant kernel: the fused x pass (forward x FFT + Stokes projection +
    # inverse x FFT over the 3 half-spectrum components), stage bucket 4.
    # Algorithmic bytes: read + write 3 components of 256*256*129 complex.
    nc = 256 * 256 * 129
    xk_bytes = 2 * 3 * nc * 2 * 4  # fp32 path (the headline precision)
    xk_ms = float(st32[4])

    res = {}
    if args.only in ("all", "fp64", "mixed"):
        for name, prec in (("mixed", abi.FS_FP32), ("fp64", abi.FS_FP64)):
            if args.only not in ("all", name):
                continue
            s, _ = build_system(prec)
            el, its, nt = time_steps(s, args.warmup, args.steps, torch)
            s.close()
            res[name] = (el, its)
            log("%s: %d steps in %.3f s -> %.3f steps/s, iters %s" %
                (name, args.steps, el, args.steps / el, its))

    value = args.steps / res["mixed"][0] if "mixed" in res else None
    if "fp64" in res:
        out["fp64_timesteps_per_s"] = args.steps / res["fp64"][0]
        out["fp64_iters"] = res["fp64"][1]
    if "mixed" in res:
        out["mixed_iters"] = res["mixed"][1]
    if "fp64" in res and value:
        out["speedup_vs_fp64"] = value / out["fp64_timesteps_per_s"]

    cpu = None
    if not args.no_cpu:
        iters = res["fp64"][1] if "fp64" in res else [10]
        mean_it = sum(iters) / len(iters)
        t_solve, tsps, sps = cpu_baseline(mean_it)
        cpu = {"value": tsps, "unit": "timesteps/s", "cores": 1,
               "kind": "port",
               "sample": ("one FP64 256^3 / 30,000-segment mobility solve "
                          "by the oracle (%.1f s), scaled by %.1f solves per "
                          "step (GPU FP64 mean Broyden iterations + 1)" %
                          (t_solve, sps))}
        log("cpu baseline", cpu)

    ms = 1e3 * res["mixed"][0] / args.steps if "mixed" in res else None
    line = {
        "metric": "timesteps/s at 256^3, 30k segs: mixed+learned vs plain "
                  "FP64 on GPU, at fixed error",
        "value": value, "unit": "timesteps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 FCM + f64 Broyden", "data": "synthetic",
        "config": {"workload": "1000 swimming filaments x 30 segments, "
                               "256^3 periodic FCM grid, dx=a/1.86, ngd=12, "
                               "KB=1000, dt=0.005, TOL=1e-6, barrier on",
                   "global_batch": 30000, "seq_len": 256,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_fft_x_project<float>",
                     "achieved": xk_bytes / (xk_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": xk_bytes / (xk_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None},
        "cpu_baseline": cpu,
        "extra": out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
