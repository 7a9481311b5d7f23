"""Benchmark: implicit filament timesteps/s at 256^3 with 30,000 segments.

Workload (one GPU): 1,000 swimming filaments x 30 segments in a periodic
256^3 box (dx = a/1.86, ngd = 12), travelling-wave preferred curvature
(amplitude 0.5/L, wavenumber 2 pi/L, period 1), bending and twist moduli
1000 (the reference swimmer modulus, config.hpp:100), dt = 0.005, barrier on,
BDF2 + bad Broyden with TOL 1e-6. A step is one whole implicit timestep
(initial guess, forces, Jacobian, cell list, barrier, Broyden iterations of
mobility solves, state step), all resident in HBM.

Reported `value`: timesteps/s of the mixed-precision + learned-warm-start
path (FCM stages in FP32, Broyden/tensions/BDF2 state in FP64, first
Broyden iterate from an MLP trained on the GPU on --record steps of the
plain solver's own trajectory, kept only when its residual is lower).
Also reported, from the same state and TOL: plain mixed precision and
plain FP64 timesteps/s, Broyden iterations, the position deviation from the
FP64 path, mobility solves/s of the 256^3 / 30,000-segment FCM
microbenchmark in FP64 and FP32, and the roofline of the dominant kernel
measured with HIP events on the solver's own stream.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch of the FP32 y pass (k_fft_y<float, 8, 8>, the
# timestep's largest kernel by total time) at 256^3, measured with rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE in separate runs, each scaled by the same
# counter's reading for a 1 GiB copy
# (profiles/fcm_mobility_256cubed_30k_hbm_bytes.txt): 203.0 MB read + 202.9
# MB written, the half spectrum once each way
YPASS_FP32_TRAFFIC_BYTES = 405.9e6


def api_kernels():
    from filsim_amd import api
    return api.Fcm.KERNELS


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(precision, nfil=1000, nw=30, grid=256, seed=20261202,
             kap=1000.0, tol=1e-6, dt=0.005):
    """The swimming workload's parameters, filament properties and initial
    positions/orientations."""
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
    return fp, sp, [nw] * nfil, props, X, q


def build_system(precision, **kw):
    from filsim_amd import api
    fp, sp, nws, props, X, q = workload(precision, **kw)
    s = api.System(sp, nws, props)
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
    # per-stage and per-kernel device times (HIP events on the solver's
    # stream)
    fcm.set_timing(True)
    stages = np.zeros(7)
    kernels = np.zeros(10)
    nrep = 10
    for _ in range(nrep):
        fcm.mobility(Y, F, T, V, W)
        stages += fcm.stage_times()
        kernels += fcm.kernel_times()
    fcm.set_timing(False)
    stages /= nrep
    kernels /= nrep
    fcm.close()
    return t, stages, fp, kernels


CPU_BUCKETS = ("assign_F", "build_Jinv", "collision_barrier", "FCM",
               "error_check", "solve_Jinv", "timestep_total", "FCM_spread",
               "FCM_fft_forward", "FCM_flow_solve", "FCM_fft_backward",
               "FCM_particle_velocities_rotations")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(state, nt0, max_steps=3, budget_s=25.0, wl=None):
    """The oracle (the FP64 CPU restatement of main.cpp:284-754, its FFT,
    spread, interpolation and per-filament loops on OpenMP threads) steps
    the workload from `state` at timestep nt0 -- the state the GPU paths
    start from -- for max_steps real timesteps, or fewer once budget_s of
    wall time has passed. Returns timesteps/s, the iterations per step,
    the profiler buckets (profilers.hpp:24-55) in seconds per step, the
    thread count and the CPU model."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from filsim_amd import abi
    wl = wl or {}
    fp, sp, nws, props, _, _ = workload(abi.FS_FP64, **wl)
    o = O.Sim(sp, nws, props)
    o.set_state(**state)
    its, tb = [], np.zeros(len(CPU_BUCKETS))
    nt = nt0
    t0 = time.perf_counter()
    while len(its) < max_steps:
        its.append(o.timestep(nt))
        tb += o.bucket_times()
        nt += 1
        log("cpu oracle step %d: %d iterations, %.1f s so far" %
            (nt - 1, its[-1], time.perf_counter() - t0))
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    n = len(its)
    return {"timesteps_per_s": n / el, "steps": n, "first_step": nt0,
            "iterations": its, "seconds": el,
            "buckets_s_per_step": {k: round(float(v) / n, 4)
                                   for k, v in zip(CPU_BUCKETS, tb)},
            "fcm_solve_s_mean": float(tb[3]) / (sum(its) + n),
            "threads": O.num_threads(), "cpu_model": cpu_model()}


def c1_cpu_gpu(args, torch):
    """The CPU-runnable case in full (16 filaments x 15 segments, 64^3,
    gravity, barrier on, TOL 1e-4, dt 0.01, 100 steps, FP64) on the oracle
    and on the GPU from the same state: timesteps/s of each, iterations
    per step and the oracle's profiler buckets."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from filsim_amd import abi, api, params
    fp = params.fcm_params(64)
    sp = params.sim_params(fp, dt=0.01, tol=1e-4, maxiter=100,
                           precision=abi.FS_FP64, weight=(0.0, 0.0, -1.0),
                           barrier_on=1)
    nws = [15] * 16
    props = params.fil_props(16, kap=1.0)
    X, q = api.make_filaments(params.box(fp), nws, dL=2.2, min_sep=2.2,
                              seed=20261202)
    st = dict(X=X, Xs=X, Xt=X, q=q, qs=q, qt=q)
    out = {"cpu_model": cpu_model(), "threads": O.num_threads()}
    o = O.Sim(sp, nws, props)
    o.set_state(**st)
    its, tb = [], np.zeros(len(CPU_BUCKETS))
    t0 = time.perf_counter()
    for nt in range(1, 101):
        its.append(o.timestep(nt))
        tb += o.bucket_times()
        if nt % 20 == 0:
            log("c1 oracle step %d (%.1f s)" % (nt, time.perf_counter() - t0))
    el = time.perf_counter() - t0
    out["c1_cpu"] = {"timesteps_per_s": 100 / el, "iterations": its,
                     "buckets_s_per_step": {
                         k: round(float(v) / 100, 5)
                         for k, v in zip(CPU_BUCKETS, tb)}}
    g = api.System(sp, nws, props)
    g.set_state(**st)
    gits = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for nt in range(1, 101):
        gits.append(g.timestep(nt))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    g.close()
    out["c1_gpu_fp64"] = {"timesteps_per_s": 100 / el, "iterations": gits}
    return out


def batch_bench(args, torch, n=64, seed=4000):
    """Batched sweep: n suspensions (grid from {32..128}^3, 8-200
    filaments of 15-40 segments capped by box size, KB log-uniform in
    [0.1, 10], gravity N(0, 0.2^2)), FP32 FCM + FP64 Broyden, TOL 1e-6,
    dt 0.01, each suspension's FCM autotuned. Every mode (the batch in each
    launch mode, and the same suspensions stepped one after another) runs a
    fresh copy of the suspensions from the same initial state over the same
    steps, so the modes do the same work."""
    from filsim_amd import abi, api, params
    pts = [params.sweep_point(seed + k, precision=abi.FS_FP32, tol=1e-6)
           for k in range(n)]
    t0 = time.perf_counter()

    def fresh(tunes=None):
        g = []
        for k, (fp, sp, nws, props, st, info) in enumerate(pts):
            s = api.System(sp, nws, props)
            s.set_state(**st)
            if tunes is not None:
                s.set_tuning(tunes[k])
            g.append(s)
        return g

    tuner = fresh()
    tunes = [s.autotune(reps=2) for s in tuner]
    for s in tuner:
        s.close()
    t_setup = time.perf_counter() - t0
    res, finals, t_capture = {}, {}, 0.0
    steps = max(2, args.steps // 2)
    for name in (["fused_lanes%d" % k for k in (1, 4)] +
                 ["batched_lanes%d" % k for k in (0, 4, 8)] + ["sequential"]):
        grp = fresh(tunes)
        batch = None
        if name != "sequential":
            t0 = time.perf_counter()
            batch = api.Batch(grp)
            t_capture = time.perf_counter() - t0
            batch.set_fused(name.startswith("fused"))
            batch.set_lanes(int(name.split("_lanes")[1]))

        def step(nt):
            if batch is not None:
                return batch.timestep(nt)
            return [s.timestep(nt) for s in grp]

        nt = 1
        for _ in range(args.warmup):
            step(nt)
            nt += 1
        torch.cuda.synchronize()
        its_all = []
        t0 = time.perf_counter()
        for _ in range(steps):
            its_all.append(step(nt))
            nt += 1
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[name] = {"suspension_steps_per_s": n * steps / el,
                     "s_per_batch_step": el / steps,
                     "mean_iters": float(np.mean(its_all))}
        log(name, res[name])
        finals[name] = [s.get_state()["X"] for s in grp]
        if batch is not None:
            batch.close()
        for s in grp:
            s.close()
    best = max((k for k in res if k != "sequential"),
               key=lambda k: res[k]["suspension_steps_per_s"])
    same = all(np.array_equal(a, b)
               for k in res if k != "sequential"
               for a, b in zip(finals[k], finals["sequential"]))
    res["batched"] = dict(res[best], mode=best)
    info = [p[5] for p in pts]
    return {"metric": "batched sweep: suspension-timesteps/s, %d suspensions"
                      % n,
            "value": res["batched"]["suspension_steps_per_s"],
            "unit": "suspension-timesteps/s", "n_gpus": 1, "steps": steps,
            "warmup": args.warmup, "higher_is_better": True,
            "dtype": "f32 FCM + f64 Broyden", "data": "synthetic",
            "config": {"workload": "%d suspensions, grids %s, %d segments "
                                   "total, TOL 1e-6, dt 0.01" %
                                   (n, sorted(set(i["grid"] for i in info)),
                                    sum(i["segments"] for i in info))},
            "batched": res["batched"], "sequential": res["sequential"],
            "modes": {k: v for k, v in res.items()
                      if k.startswith("batched_") or k.startswith("fused_")},
            "speedup": res["batched"]["suspension_steps_per_s"] /
                       res["sequential"]["suspension_steps_per_s"],
            "bitwise_equal_final_positions": bool(same),
            "setup_s": t_setup, "capture_s": t_capture}


def c5_bench(args, torch):
    """Learned warm start at scale: 8,000 swimming filaments x 30 segments
    (240,000) at the 1,000-filament 256^3 workload's density, i.e. on a
    512^3 grid (dx = a/1.86, ngd = 12), random non-overlapping placement,
    random phases, KB = 1000, dt = 0.005, TOL 1e-6, barrier on, FP32 FCM +
    FP64 Broyden, placed by the GPU generator. (On a 384^3 grid the same
    filaments fill 11% of the box and the per-filament approximate Jacobian
    no longer makes Broyden converge: the iterates diverge in the first
    steps, on the CPU oracle as on the GPU.) The network is trained on
    56,000 filament-steps (7 recorded steps after 40 start-up steps, in
    which the swimming gait develops from straight filaments and the
    iteration count settles) of the plain mixed path, then mixed + learned
    and plain mixed are timed from the same state."""
    from filsim_amd import abi, api, params, warmstart
    fp = params.fcm_params(512)
    nfil, nw = 8000, 30
    rng = np.random.default_rng(9)
    phases = list(rng.uniform(0, 2 * math.pi, nfil))
    X0, q0 = api.make_filaments(params.box(fp), [nw] * nfil, dL=2.2,
                                min_sep=2.2, seed=20261202, gpu=True)

    def build():
        sp = params.sim_params(fp, dt=0.005, tol=1e-6, maxiter=100,
                               precision=abi.FS_FP32, weight=(0.0, 0.0, 0.0),
                               barrier_on=1)
        props = params.swimming_props(nfil, nw, kap=1000.0, phases=phases)
        s = api.System(sp, [nw] * nfil, props)
        return s

    log("c5: building 8000 x 30 segments at 512^3")
    s_l = build()
    s_l.set_state(X=X0, Xs=X0, Xt=X0, q=q0, qs=q0, qt=q0)
    log("c5: built")
    t0 = time.perf_counter()
    skip, rec = 40, 7
    X, T, its_rec = warmstart.record(s_l, rec, nt0=1, torch=torch, skip=skip,
                                     log=log)
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    mlp, hist = warmstart.train(X, T, torch, epochs=args.epochs, batch=1024,
                                lr=1e-3, log=log)
    t_train = time.perf_counter() - t0
    log("c5: recorded %d samples in %.1fs, trained in %.1fs, holdout %.3f" %
        (X.shape[0], t_rec, t_train, hist[-1][2]))
    del X, T
    state = s_l.get_state()
    keep = {k: state[k] for k in ("X", "Xt", "U", "Ut", "q", "qt", "lam",
                                  "lam1", "lam2")}
    keep.update(Xs=state["X"], Us=state["U"], qs=state["q"])
    api.system_set_warmstart(s_l, mlp)
    api.system_set_warmstart_check(s_l, args.ws_check)
    api.system_set_history_carry(s_l, args.c5_hist_carry)
    nt0 = 1 + skip + rec
    res = {}
    for name in ("learned", "mixed"):
        if name == "learned":
            s = s_l
        else:
            s = build()
            s.set_state(**keep)
        nt = nt0
        for _ in range(args.warmup):
            s.timestep(nt)
            nt += 1
        torch.cuda.synchronize()
        its = []
        steps = max(3, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(steps):
            its.append(s.timestep(nt))
            log("c5 %s step %d: %d iterations" % (name, nt, its[-1]))
            nt += 1
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[name] = {"timesteps_per_s": steps / el, "iters": its,
                     "mean_iters": float(np.mean(its))}
        log("c5", name, res[name])
        s.close()
    mlp.close()
    return {"metric": "timesteps/s at 512^3, 240k segs: mixed+learned vs "
                      "mixed", "value": res["learned"]["timesteps_per_s"],
            "unit": "timesteps/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32 FCM + f64 Broyden + bf16 MLP", "data": "synthetic",
            "config": {"workload": "8000 swimming filaments x 30 segments, "
                                   "512^3, KB=1000, dt=0.005, TOL=1e-6",
                       "warmstart_check_every": args.ws_check,
                       "broyden_history_carry_cols": args.c5_hist_carry},
            "learned": res["learned"], "mixed": res["mixed"],
            "speedup_vs_mixed": res["learned"]["timesteps_per_s"] /
                                res["mixed"]["timesteps_per_s"],
            "train": {"samples": 56000, "record_s": t_rec,
                      "train_s": t_train, "holdout": hist[-1][2]}}


def c3_bench(args, torch):
    """Mixed precision against FP64 over the whole swimming run: 1,000
    swimming filaments x 30 segments, 256^3, KB = 1000, dt = 0.005,
    TOL 1e-6, barrier on, args.c3_steps steps from one initial state, the
    FP32-FCM path and the FP64 path (whose mobility solve matches the FP64
    oracle to 1e-12) side by side. Every 250 steps: the largest position and
    orientation differences (positions relative to the box) and the
    relative differences of the segment velocities U and angular velocities
    Omega of the last mobility solve."""
    import ctypes as C
    from filsim_amd import abi, params
    hip = C.CDLL("libamdhip64.so")

    def d2h(ptr, n):
        out = np.zeros(n)
        rc = hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptr,
                           C.c_size_t(8 * n), 2)
        assert rc == 0, "hipMemcpy failed"
        return out

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    s32, fp = build_system(abi.FS_FP32)
    s64, _ = build_system(abi.FS_FP64)
    L = max(params.box(fp))
    n3 = 3 * 30000
    rows, it32, it64 = [], 0, 0
    t0 = time.perf_counter()
    for nt in range(1, args.c3_steps + 1):
        it32 += s32.timestep(nt)
        it64 += s64.timestep(nt)
        if nt % 250 == 0 or nt == args.c3_steps:
            a, b = s32.get_state(), s64.get_state()
            V32, W32 = (d2h(p, n3) for p in s32.fcm_ptrs()[3:5])
            V64, W64 = (d2h(p, n3) for p in s64.fcm_ptrs()[3:5])
            row = {"step": nt,
                   "pos_max_over_box": float(np.abs(a["X"] - b["X"]).max() / L),
                   "q_max": float(np.abs(a["q"] - b["q"]).max()),
                   "U_rel": rel(V32, V64), "Omega_rel": rel(W32, W64),
                   "mean_iters_fp32": it32 / nt, "mean_iters_fp64": it64 / nt}
            rows.append(row)
            log("c3", row, "%.0fs" % (time.perf_counter() - t0))
    s32.close()
    s64.close()
    return {"metric": "mixed precision vs FP64 over the swimming run at "
                      "256^3, 30k segments (errors; not a rate)",
            "n_gpus": 1, "steps": args.c3_steps, "data": "synthetic",
            "config": {"workload": "1000 swimming filaments x 30 segments, "
                                   "256^3, KB=1000, dt=0.005, TOL=1e-6, "
                                   "barrier on"},
            "reference_path": "FP64 on the GPU (mobility = FP64 oracle to "
                              "1e-12 per solve)",
            "rows": rows, "wall_s": time.perf_counter() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mob-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--record", type=int, default=100,
                    help="plain steps recorded to train the warm start")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--refresh", type=int, default=100,
                    help="learned path: fine-tune on its own last REFRESH "
                         "steps every REFRESH steps (0: never)")
    ap.add_argument("--ws-check", type=int, default=10,
                    help="learned path: solve the extrapolated guess beside "
                         "the network's every WS_CHECK steps (and when the "
                         "network's guess loses); 1: every step")
    ap.add_argument("--hist-carry", type=int, default=96,
                    help="learned path: carry the Broyden history across "
                         "steps while it holds at most this many columns "
                         "(0: restart every step)")
    ap.add_argument("--c5-hist-carry", type=int, default=0,
                    help="--only c5: Broyden history carried across steps "
                         "(measured slower there: 6.0 vs 7.7 timesteps/s "
                         "with 96 columns)")
    ap.add_argument("--c3-steps", type=int, default=2000,
                    help="steps of the mixed-vs-FP64 swimming run (--only c3)")
    ap.add_argument("--skip", type=int, default=10,
                    help="start-up steps run before recording")
    ap.add_argument("--only", choices=["all", "mixed", "fp64", "mobility",
                                       "copy", "batch", "c5", "c3", "cpu"],
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
    if args.only == "batch":
        print(json.dumps(batch_bench(args, torch)), flush=True)
        return
    if args.only == "c5":
        print(json.dumps(c5_bench(args, torch)), flush=True)
        return
    if args.only == "c3":
        print(json.dumps(c3_bench(args, torch)), flush=True)
        return
    if args.only == "cpu":
        print(json.dumps(c1_cpu_gpu(args, torch)), flush=True)
        return
    if args.only == "copy":
        # calibration for HBM byte counters: a plain 1 GiB FP32 scale-by-one
        # copy (reads and writes exactly 2^30 bytes per dispatch)
        src = torch.rand(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 1 << 30,
                          "copy_bytes_written": 1 << 30}), flush=True)
        return

    # FCM mobility microbenchmark (256^3, 30,000 segments)
    t64, st64, fp, kt64 = mobility_bench(abi.FS_FP64, args.mob_reps, torch)
    t32, st32, _, kt32 = mobility_bench(abi.FS_FP32, args.mob_reps, torch)
    log("mobility fp64 %.3f ms (%.1f solves/s), fp32 %.3f ms (%.1f/s)" %
        (1e3 * t64, 1 / t64, 1e3 * t32, 1 / t32))
    log("fp64 stages ms", [round(float(x), 4) for x in st64])
    log("fp32 stages ms", [round(float(x), 4) for x in st32])
    log("fp64 kernels ms", [round(float(x), 4) for x in kt64])
    log("fp32 kernels ms", [round(float(x), 4) for x in kt32])
    N = 256 ** 3
    out["mobility"] = {
        "fp64_solves_per_s": 1 / t64, "fp32_solves_per_s": 1 / t32,
        "fp64_ms": 1e3 * t64, "fp32_ms": 1e3 * t32,
        "fp64_algorithmic_GBps": 24 * N * 8 / t64 / 1e9,
        "fp32_algorithmic_GBps": 24 * N * 4 / t32 / 1e9,
        "fp64_stage_ms": [float(x) for x in st64],
        "fp32_stage_ms": [float(x) for x in st32],
        "kernels": list(api_kernels()),
        "fp64_kernel_ms": [float(x) for x in kt64],
        "fp32_kernel_ms": [float(x) for x in kt32]}
    # dominant kernel of the timestep by total time: the FP32 y pass, two
    # launches per solve (forward and inverse; HIP events on the solver's
    # stream around each launch, averaged). Algorithmic bytes: one read and
    # one write of the 3 half-spectrum components (3 * 256 * 256 * 129
    # complex floats each way).
    xk_bytes = 2 * 3 * 256 * 256 * 129 * 8
    xk_ms = 0.5 * float(kt32[4] + kt32[6])

    res = {}
    info = {}
    cpu_state, nt0 = None, 1
    if args.only in ("all", "fp64", "mixed"):
        from filsim_amd import api, warmstart
        # 1) the mixed-precision solver generates the training trajectory
        s_l, _ = build_system(abi.FS_FP32)
        t0 = time.perf_counter()
        X, T, its_rec = warmstart.record(s_l, args.record, nt0=1, torch=torch,
                                         skip=args.skip)
        t_rec = time.perf_counter() - t0
        log("recorded %d samples in %.1f s (iters %s)" %
            (X.shape[0], t_rec, its_rec[-5:]))
        t0 = time.perf_counter()
        mlp, hist = warmstart.train(X, T, torch, epochs=args.epochs,
                                    batch=1024, lr=1e-3, log=log, linear=True)
        t_train = time.perf_counter() - t0
        del X, T
        nt0 = args.skip + args.record + 1
        state = s_l.get_state()
        keep = {k: state[k] for k in ("X", "Xt", "U", "Ut", "q", "qt", "lam",
                                      "lam1", "lam2")}
        keep.update(Xs=state["X"], Us=state["U"], qs=state["q"])
        cpu_state = keep
        api.system_set_warmstart(s_l, mlp)
        api.system_set_warmstart_check(s_l, args.ws_check)
        api.system_set_history_carry(s_l, args.hist_carry)
        info["train"] = {"samples": args.record * 1000, "skip": args.skip,
                         "record_s": t_rec,
                         "train_s": t_train, "epochs": args.epochs,
                         "holdout_loss_every_10_epochs":
                             [h[2] for h in hist[::10]] + [hist[-1][2]],
                         "holdout_loss_network": hist[-1][2],
                         "holdout_loss_ridge_alone": mlp.ridge_holdout,
                         "holdout_loss_ridge_through_network":
                             mlp.ridge_holdout_net,
                         "holdout_loss_extrapolation_alone":
                             mlp.zero_holdout,
                         "online_refresh_every_steps": args.refresh}
        # 2) the same state, three paths, same TOL
        finals = {}
        for name, prec in (("learned", None), ("mixed", abi.FS_FP32),
                           ("fp64", abi.FS_FP64)):
            if name == "learned":
                s = s_l
            else:
                s, _ = build_system(prec)
                s.set_state(**keep)
            its, used, solved, wits = [], 0, 0, []
            nt = nt0
            for _ in range(args.warmup):
                wits.append(s.timestep(nt))
                nt += 1
            info["warmup_iters_" + name] = wits
            ref = None
            if name == "learned" and args.refresh > 0:
                ref = warmstart.OnlineRefresh(s, mlp, torch,
                                              window=args.refresh)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                its.append(s.timestep(nt))
                if name == "learned":
                    u, rp, _ = api.system_last_step_info(s)
                    used += u
                    solved += int(np.isfinite(rp))
                    if ref is not None:
                        ref.after_step()
                nt += 1
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            if ref is not None:
                info["online_refreshes"] = ref.refreshes
                ref.close()
            w = max(1, min(100, args.steps))
            info["mean_iters_per_%d_steps_%s" % (w, name)] = [
                round(float(np.mean(its[i:i + w])), 2)
                for i in range(0, len(its), w)]
            finals[name] = s.get_state()["X"]
            res[name] = (el, its)
            # one more step, profiled (outside the timed region), in the
            # reference profiler's buckets
            s.set_profiling(True)
            s.timestep(nt)
            info["profile_ms_" + name] = {
                k: round(v, 4) for k, v in s.profile().items()}
            s.set_profiling(False)
            log("%s: %d steps in %.3f s -> %.3f steps/s, iters %s%s" %
                (name, args.steps, el, args.steps / el, its,
                 " (network guess used %d/%d, extrapolated guess solved "
                 "on %d)" % (used, args.steps, solved)
                 if name == "learned" else ""))
            if name == "learned":
                info["learned_used"] = used
                info["learned_extrapolated_solves"] = solved
                info["warmstart_check_every"] = args.ws_check
                info["broyden_history_carry_cols"] = args.hist_carry
            s.close()
        ref = finals["fp64"]
        for name in ("learned", "mixed"):
            info["pos_rel_err_vs_fp64_" + name] = float(
                np.abs(finals[name] - ref).max() / np.abs(ref).max())
        mlp.close()

    value = args.steps / res["learned"][0] if "learned" in res else None
    for name in ("mixed", "fp64", "learned"):
        if name in res:
            out[name + "_timesteps_per_s"] = args.steps / res[name][0]
            out[name + "_iters"] = res[name][1]
    if "fp64" in res and value:
        out["speedup_vs_fp64"] = value / out["fp64_timesteps_per_s"]
        out["mixed_speedup_vs_fp64"] = (out["mixed_timesteps_per_s"] /
                                        out["fp64_timesteps_per_s"])
    out.update(info)

    cpu = None
    if not args.no_cpu and cpu_state is not None:
        cb = cpu_baseline(cpu_state, nt0, max_steps=2, budget_s=30.0)
        cpu = {"value": cb["timesteps_per_s"], "unit": "timesteps/s",
               "cores": cb["threads"], "kind": "port",
               "cpu_model": cb["cpu_model"],
               "sample": ("the oracle's FP64 timesteps %d-%d of this "
                          "workload (%s Broyden iterations; the GPU FP64 "
                          "path's first steps from the same state took %s) "
                          "on %d OpenMP threads, %.1f s" %
                          (cb["first_step"], cb["first_step"] + cb["steps"]
                           - 1, cb["iterations"],
                           (info.get("warmup_iters_fp64", []) +
                            res["fp64"][1])[:cb["steps"]]
                           if "fp64" in res else "-",
                           cb["threads"], cb["seconds"])),
               "buckets_s_per_step": cb["buckets_s_per_step"],
               "fcm_solve_s_mean": cb["fcm_solve_s_mean"]}
        log("cpu baseline", cpu)

    ms = 1e3 * res["learned"][0] / args.steps if "learned" in res else None
    line = {
        "metric": "timesteps/s at 256^3, 30k segs: mixed+learned vs plain "
                  "FP64 on GPU, at fixed error",
        "value": value, "unit": "timesteps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 FCM + f64 Broyden + bf16 MLP", "data": "synthetic",
        "config": {"workload": "1000 swimming filaments x 30 segments, "
                               "256^3 periodic FCM grid, dx=a/1.86, ngd=12, "
                               "KB=1000, dt=0.005, TOL=1e-6, barrier on",
                   "global_batch": 30000, "seq_len": 256,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_fft_y<float, 8, 8>",
                     "achieved": xk_bytes / (xk_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": xk_bytes / (xk_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": YPASS_FP32_TRAFFIC_BYTES,
                     "algorithmic_bytes": xk_bytes},
        "cpu_baseline": cpu,
        "extra": out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
