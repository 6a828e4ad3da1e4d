"""Benchmark: RK-Merson steps/s of the plasma-transport hot path on one MI355X.

Workload: 512x512 argon discharge, FP64, 2 Torr, 400 V trapezoid pulse
(10 ns rise), geometric x refinement toward both electrodes, Gaussian + 1%
seeded noise initial state.  A "step" is one accepted pass of the main loop
(main.F90:74-121): ptcl_step (5 stages, circuit, controller, restarts),
lapl_solve (assembly + PCG to a relative residual of 1e-10) and accept.

Prints ONE JSON line (plus the roofline of the dominant kernel and the CPU
oracle timed on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def gpu_run(args):
    from fluidsim_amd import Discharge, configs

    cfg = configs.config1(args.n, rtol=args.rtol, chunk=args.chunk)
    d = Discharge(cfg)
    d.init_gaussian(seed=0)
    # warm-up: the two Poisson-only passes (main.F90:75) and the first steps,
    # whose controller restarts shrink dt from dt0
    d.run(args.warmup)
    s0 = d.scalars()
    log0 = len(d.log())
    d.set_timing(True)
    t0 = time.perf_counter()
    d.run(args.steps)
    t1 = time.perf_counter()
    s1 = d.scalars()
    log = d.log()[log0:]
    timers = d.timers()  # ms: stage kernels, circuit+controller, assembly, pcg
    accepted = int(round(s1["ts"] - s0["ts"]))
    attempts = len(log)
    iters = [float(r[7]) for r in log if r[6] >= 0]
    elapsed = t1 - t0
    # per-launch kernel durations on the live state (HIP events on the stream)
    N = args.n * args.n
    k = {
        "stage5": d.time_kernel(5, 50),
        "stage3": d.time_kernel(3, 50),
        "assembly": d.time_kernel(10, 50),
        "pcg_it": d.time_kernel(11, 400),
    }
    from fluidsim_amd.discharge import copy_bandwidth

    copy_gbs = copy_bandwidth(1 << 30, 20)
    return dict(cfg=cfg, elapsed=elapsed, accepted=accepted, attempts=attempts,
                iters=iters, timers=timers, kernels=k, N=N, s1=s1, copy_gbs=copy_gbs)


def cpu_baseline(args):
    """The CPU oracle (C restatement of the reference, OpenMP over rows) on a
    bounded sample: the first three passes untimed, then `cpu_steps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from fluidsim_amd import configs
    from oracle import oracle as O

    nth = args.cpu_threads or min(os.cpu_count() or 1, 16)
    cfg = configs.config1(args.n, rtol=args.rtol)
    ocfg = dict(cfg, rtol=args.rtol, maxit=cfg["lapl_maxit"], solver=2, nthreads=nth)
    o = O.Oracle(ocfg)
    o.init_gaussian(seed=0)
    o.run(3)
    t0 = time.perf_counter()
    o.run(args.cpu_steps)
    dt = time.perf_counter() - t0
    return dict(value=args.cpu_steps / dt, unit="steps/s", cores=nth, kind="port",
                sample=f"{args.cpu_steps} steps of the {args.n}x{args.n} workload after 3 "
                       f"untimed passes; Jacobi-PCG Poisson to 1e-10, {nth} OpenMP threads")


def profiled_traffic(kernel: str, n: int):
    """HBM bytes per dispatch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{n}x{n}_pulse_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == kernel and len(cells) >= 7:
            return float(cells[6]) * 1e6
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: one discharge per device; "
                                   "multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    r = gpu_run(args)
    steps_per_s = r["accepted"] / r["elapsed"]
    N = r["N"]
    plane = 8.0 * N
    # algorithmic bytes per launch (one FP64 plane = 8*N bytes):
    #  PCG iteration kernel reads r, w, s, p, x, d, cx, cy and writes p, x, s,
    #  r, w: 13 planes;
    #  fused stage kernel at stage 5 reads 4 densities, ph, n3 x4, k1, k3, k4 x4
    #  and writes 4 densities: 25 planes
    bytes_pcg = 13 * plane
    bytes_stage5 = 25 * plane
    ka = r["kernels"]["pcg_it"] * 1e-3
    ks = r["kernels"]["stage5"] * 1e-3
    ach_a = bytes_pcg / ka / 1e9
    ach_s = bytes_stage5 / ks / 1e9
    timers = r["timers"]
    tot_ms = r["elapsed"] * 1e3
    iters = r["iters"]
    line = {
        "metric": "RK-Merson steps/s at 512^2 FP64; Poisson time-to-1e-10, learned vs plain PCG",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tot_ms / max(r["accepted"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Gaussian + seeded 1% noise initial state)",
        "config": {"workload": f"argon DC-glow pulse, {args.n}x{args.n} FP64, 2 Torr, 400 V "
                               f"trapezoid, geometric x grid",
                   "nx": args.n, "ny": args.n, "poisson_rtol": args.rtol,
                   "poisson_solver": "matrix-free Jacobi-PCG (Chronopoulos-Gear)"},
        "accepted_steps": r["accepted"],
        "ptcl_attempts": r["attempts"],
        "pcg_iters_mean": sum(iters) / max(len(iters), 1),
        "time_split_ms": {"stage_kernels": timers[0], "circuit_controller": timers[1],
                          "assembly": timers[2], "pcg": timers[3], "wall": tot_ms},
        "kernel_ms": r["kernels"],
        "roofline": {"kernel": "k_pcg_it", "bound": "hbm", "achieved": ach_a,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach_a / PEAK_HBM_GBS,
                     "traffic": profiled_traffic("k_pcg_it", args.n),
                     "algorithmic_bytes": bytes_pcg},
        "roofline_stage": {"kernel": "k_stage (stage 5)", "bound": "hbm", "achieved": ach_s,
                           "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach_s / PEAK_HBM_GBS},
        "hbm_copy_GBs_measured": r["copy_gbs"],
        "poisson": {"plain_pcg_ms_per_solve": timers[3] / max(r["accepted"], 1),
                    "plain_pcg_iters": sum(iters) / max(len(iters), 1), "learned": None},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
