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


def warm_mode(args):
    """fs_params.lapl_warm of the run: 0 with --cold; else --warm, or by default
    the projection onto the last corrections (2) for the 2048^2 workload, whose
    extrapolated starts leave 18 iterations per solve (12 projected), and the
    256-member ensemble (9577 against 7379 aggregate steps/s), and the
    extrapolation (1) at 512^2, where it is already within 0.2 iterations of
    the projection and the projection's two kernels cost more than that."""
    if args.cold:
        return 0
    if args.warm:
        return args.warm
    return 2 if args.config in (2, 3) else 1


WARM_TEXT = {0: "zero", 1: "extrapolated from the last two passes' corrections",
             2: "energy-norm projection onto the last three passes' corrections"}


def gpu_run(args):
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.lib import FsError

    if args.config == 2:
        cfg = configs.config2(args.n, rtol=args.rtol, chunk=args.chunk, warm=warm_mode(args),
                              graphs=int(not args.no_graphs))
    else:
        cfg = configs.config1(args.n, rtol=args.rtol, chunk=args.chunk, warm=warm_mode(args),
                              graphs=int(not args.no_graphs))
    d = Discharge(cfg)
    d.init_gaussian(seed=0)
    # warm-up: the two Poisson-only passes (main.F90:75) and the first steps,
    # whose controller restarts shrink dt from dt0
    d.run(args.warmup)
    s0 = d.scalars()
    log0 = len(d.log())
    t0 = time.perf_counter()
    d.run(args.steps)
    t1 = time.perf_counter()
    s1 = d.scalars()
    log = d.log()[log0:]
    # the split of a pass over its phases, from HIP events between the phases
    # in a further run (the events themselves cost GPU idle time, so they stay
    # out of the timed run)
    split_steps = max(1, args.steps // 4)
    d.set_timing(True)
    ts0 = time.perf_counter()
    d.run(split_steps)
    ts1 = time.perf_counter()
    timers = d.timers()  # ms: stage kernels, circuit+controller, assembly, pcg
    d.set_timing(False)
    timers = [t / split_steps for t in timers] + [1e3 * (ts1 - ts0) / split_steps, split_steps]
    accepted = int(round(s1["ts"] - s0["ts"]))
    attempts = len(log)
    iters = [float(r[7]) for r in log if r[6] >= 0]
    elapsed = t1 - t0
    N = args.n * args.n
    # the state after the timed steps: the CPU baseline and the same-method
    # leg below start from it
    snap = snapshot(d)
    same = None if args.no_cpu else same_method_gpu(d, snap, args)
    net, train = (None, None) if args.no_learned else learned_setup(args)
    poisson = poisson_compare(d, args.solve_reps, net)
    if net is not None:
        poisson["learned"]["train"] = train
        if args.learned_n2:
            poisson["learned_at"] = {f"{args.learned_n2}x{args.learned_n2}":
                                     solve_at(args.learned_n2, net, args.solve_reps)}
    # per-launch kernel durations on the live state (HIP events on the stream)
    k = {f"stage{s}": d.time_kernel(s, 50) for s in (1, 2, 3, 4, 5)}
    k["assembly"] = d.time_kernel(10, 50)
    # multigrid-PCG: the down / up kernel of each tile level, the one-workgroup
    # coarse kernel and one whole iteration
    for lvl in range(16):
        try:
            k[f"mg_down{lvl}"] = d.time_kernel(20 + lvl, 200)
            k[f"mg_up{lvl}"] = d.time_kernel(40 + lvl, 200)
        except FsError:
            break
    k["mg_coarse"] = d.time_kernel(18, 200)
    k["mg_iteration"] = d.time_kernel(17, 100)
    k["jacobi_pcg_it"] = d.time_kernel(11, 400)
    from fluidsim_amd.discharge import copy_bandwidth

    copy_gbs = copy_bandwidth(1 << 30, 20)
    return dict(cfg=cfg, elapsed=elapsed, accepted=accepted, attempts=attempts,
                iters=iters, timers=timers, kernels=k, N=N, s1=s1, copy_gbs=copy_gbs,
                poisson=poisson, snap=snap, same=same, stage_tile=d.stage_tile())


def snapshot(d):
    """Fields (levels 1-3 of the four species, the potential) and scalars."""
    return {"fields": {(n, lev): d.get(n, lev) for n in ("ni", "ne", "nt", "nm")
                       for lev in (1, 2, 3)},
            "ph": d.get("ph"), "scalars": d.scalars()}


def restore(x, snap):
    """Put a snapshot into a Discharge or an oracle instance (same interface)."""
    for (n, lev), a in snap["fields"].items():
        x.set(n, lev, a)
    x.set("ph", 1, snap["ph"])
    x.set_scalars(snap["scalars"])


def same_method_gpu(d, snap, args, passes=5):
    """The device with the CPU baseline's Poisson method: Jacobi-PCG from zero to
    the same tolerance, `passes` passes from the snapshot; the handle is then
    returned to the snapshot and to multigrid-PCG."""
    restore(d, snap)
    d.set_solver(0)
    s0 = d.scalars()["ts"]
    t0 = time.perf_counter()
    d.run(passes)
    el = time.perf_counter() - t0
    acc = d.scalars()["ts"] - s0
    d.set_solver(1)
    restore(d, snap)
    return {"poisson": f"Jacobi-PCG from zero to {args.rtol:g}", "passes": passes,
            "value": acc / el, "unit": "steps/s"}


def learned_setup(args):
    """Harvest residuals from seeded runs of the 512^2 pulse workload (three
    quarters of the set) and of the parameter-sweep ensemble (a quarter), train
    the network on the device (the library's own kernels) and return it."""
    from fluidsim_amd.learn import (LearnedPreconditioner, harvest_ensemble, harvest_runs,
                                    train_schedule)

    net = LearnedPreconditioner(args.learn_samples, 64, 64, seed=1)
    n_ens = args.learn_samples // 4
    th = harvest_runs(net, 512, seeds=(0, 1, 2, 3), samples=args.learn_samples - n_ens)
    n1 = net.samples
    he = harvest_ensemble(net, 32, seeds=(0, 1), samples=n_ens)
    steps = args.learn_steps
    loss, tt = train_schedule(net, ((steps // 2, 16, 1e-3), (steps - steps // 2, 16, 2.5e-4)))
    return net, {"samples": net.samples, "samples_config1": n1,
                 "samples_ensemble": net.samples - n1, "ensemble_members": he["members_used"],
                 "ensemble_members_skipped": he["members_skipped"],
                 "harvest_s": th + he["seconds"], "steps": steps, "batch": 16,
                 "train_s": tt, "loss_first": float(loss[:50].mean()),
                 "loss_last": float(loss[-50:].mean())}


def poisson_compare(d, reps, net=None):
    """Time-to-1e-10 of the same assembled system (the final state of the run)
    with each preconditioner: wall seconds of fs_lapl_solve (assembly + solve
    + the true-residual check), median over `reps`, from the same phi."""
    import numpy as np

    from fluidsim_amd.learn import LearnedPreconditioner

    ph0 = d.get("ph")
    out, sols = {}, {}
    cases = [("mg_pcg", 1, None), ("jacobi_pcg", 0, None)]
    if net is not None:
        cases += [("learned", 2, net), ("learned_level_without_network", 2, False)]
    for name, solver, nn in cases:
        if nn:
            nn.attach(d)
        elif nn is False:
            LearnedPreconditioner.detach(d)
        d.set_solver(solver)
        secs, its, res = [], [], []
        for _ in range(reps):
            d.set("ph", 1, ph0)
            lerr, st = d.lapl_solve()
            secs.append(st.seconds)
            its.append(st.iters)
            res.append(st.resid)
        sols[name] = d.get("ph")
        out[name] = {"ms_median": 1e3 * float(np.median(secs)), "iters": int(np.median(its)),
                     "true_rel_residual": float(max(res)), "converged": lerr == 0}
    d.set("ph", 1, ph0)
    d.set_solver(1)
    b = sols["jacobi_pcg"][1:-1, 1:-1] - ph0[1:-1, 1:-1]
    for name in sols:
        if name != "jacobi_pcg":
            a = sols[name][1:-1, 1:-1] - ph0[1:-1, 1:-1]
            out[name]["correction_rel_l2_vs_jacobi"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["plain_over_mg_time"] = out["jacobi_pcg"]["ms_median"] / out["mg_pcg"]["ms_median"]
    if net is not None:
        out["plain_over_learned_time"] = out["jacobi_pcg"]["ms_median"] / out["learned"]["ms_median"]
    else:
        out["learned"] = None
    return out


def solve_at(n, net, reps):
    """The same comparison on the final system of a short n x n pulse run."""
    from fluidsim_amd import Discharge, configs

    d = Discharge(configs.config1(n))
    d.init_gaussian(seed=0)
    d.run(15)
    return poisson_compare(d, reps, net)


def ensemble_run(args):
    """The parameter-sweep ensemble: every member advances `steps` passes in the
    ensemble's batched launches; aggregate accepted steps per second, next to
    the same members run one after the other on their own (a sample)."""
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.ensemble import Ensemble

    cfgs = configs.config3(args.members)
    for c in cfgs:
        c["lapl_warm"] = warm_mode(args)
        c["use_graphs"] = int(not args.no_graphs)
    t0 = time.perf_counter()
    E = Ensemble(cfgs)
    for m, c in zip(E.members, cfgs):
        m.init_gaussian(seed=c["seed"])
    t_setup = time.perf_counter() - t0
    E.run(args.warmup)
    ts0 = [m.scalars()["ts"] for m in E.members]
    t0 = time.perf_counter()
    E.run(args.steps)
    el = time.perf_counter() - t0
    acc = sum(m.scalars()["ts"] - a for m, a in zip(E.members, ts0))
    attempts = sum(len(m.log()) for m in E.members)
    cells = sum(c["nx"] * c["ny"] for c in cfgs)
    stage_tile = E.stage_tile()
    # solo reference: a sample of members, each run alone for the same passes
    k = min(args.solo_sample, len(cfgs))
    solo_t, solo_acc = 0.0, 0.0
    for c in cfgs[:k]:
        d = Discharge(c)
        d.init_gaussian(seed=c["seed"])
        d.run(args.warmup)
        a = d.scalars()["ts"]
        t1 = time.perf_counter()
        d.run(args.steps)
        solo_t += time.perf_counter() - t1
        solo_acc += d.scalars()["ts"] - a
    E.close()
    return dict(elapsed=el, accepted=acc, attempts=attempts, cells=cells, setup_s=t_setup,
                solo_rate=solo_acc / solo_t if solo_t > 0 else None, solo_sample=k,
                stage_tile=stage_tile)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, snap):
    """The CPU oracle (the C restatement of the reference, OpenMP over rows,
    bitwise equal to its serial run) on a bounded sample of the same workload,
    from the GPU run's state after its timed steps:
      transport  one ptcl_step attempt (five RK-Merson stages with ptclEqn,
                 circuit and controller; a replay record accepts it): median
                 of `cpu_reps` on all threads, and once on one thread;
      Poisson    one lapl_solve to the same tolerance from the state that
                 attempt leaves, with BiCGStab + ILU(0) (the reference's KSP
                 IBCGS with PETSc's default preconditioner) and with
                 Jacobi-PCG, all threads.
    value = 1 / (transport + the faster Poisson), all threads."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from fluidsim_amd import configs
    from oracle import oracle as O

    nth = args.cpu_threads or min(os.cpu_count() or 1, 16)
    cfg = configs.config1(args.n, rtol=args.rtol)
    dt = snap["scalars"]["dt"]
    rec = np.zeros((1, 10))
    rec[0, 4], rec[0, 5] = dt, 1.0

    def oracle(solver, threads):
        o = O.Oracle(dict(cfg, rtol=args.rtol, maxit=cfg["lapl_maxit"], solver=solver,
                          nthreads=threads))
        restore(o, snap)
        return o

    def transport(o):
        restore(o, snap)
        o.set_replay(rec)
        t0 = time.perf_counter()
        st = o.ptcl_step()
        t = time.perf_counter() - t0
        if st != 0:
            raise RuntimeError(f"oracle ptcl_step status {st}")
        return t

    def poisson(o):
        t0 = time.perf_counter()
        st, lerr, stats = o.lapl_solve()
        t = time.perf_counter() - t0
        return t, int(stats[0]), lerr

    ob = oracle(1, nth)
    t_all = sorted(transport(ob) for _ in range(args.cpu_reps))[args.cpu_reps // 2]
    t_bi, it_bi, lerr_bi = poisson(ob)
    oj = oracle(2, nth)
    transport(oj)
    t_ja, it_ja, lerr_ja = poisson(oj)
    t_ser = transport(oracle(1, 1))
    t_pois = min(t_bi, t_ja)
    return dict(value=1.0 / (t_all + t_pois), unit="steps/s", cores=nth, kind="port",
                sample=(f"from the {args.n}x{args.n} state after the timed GPU steps: one "
                        f"ptcl_step attempt (median of {args.cpu_reps}) + one Poisson solve to "
                        f"{args.rtol:g} (the faster of BiCGStab+ILU(0) and Jacobi-PCG), "
                        f"{nth} OpenMP threads"),
                nproc=os.cpu_count(), cpu_model=cpu_model(),
                transport_s=t_all, transport_s_serial=t_ser,
                poisson_s={"bicgstab_ilu0": t_bi, "jacobi_pcg": t_ja},
                poisson_iters={"bicgstab_ilu0": it_bi, "jacobi_pcg": it_ja},
                poisson_converged={"bicgstab_ilu0": lerr_bi == 0, "jacobi_pcg": lerr_ja == 0},
                value_jacobi=1.0 / (t_all + t_ja),
                value_serial_transport_estimate=1.0 / (t_ser + t_pois))


def profiled_traffic(kernel: str, n: int):
    """HBM bytes per dispatch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{n}x{n}_pulse_traffic.md")
    if not os.path.exists(path):
        return None
    col = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if "total MB" in cells:
            col = cells.index("total MB")
        elif col is not None and cells and cells[0] == kernel and len(cells) > col:
            return float(cells[col]) * 1e6
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=0, help="grid side (default 512, 2048 for config 2)")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solve-reps", type=int, default=3)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned preconditioner")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel from the host instead of replaying HIP graphs "
                         "(for kernel tracers that fail on large graph launches)")
    ap.add_argument("--cold", action="store_true",
                    help="start every multigrid-PCG solve of the run from zero (no warm start)")
    ap.add_argument("--warm", type=int, choices=(1, 2), default=0,
                    help="warm start: 1 extrapolation, 2 projection onto the last corrections "
                         "(default: 2 for --config 2 and 3, 1 otherwise)")
    ap.add_argument("--learn-samples", type=int, default=20000)
    ap.add_argument("--learn-steps", type=int, default=6000)
    ap.add_argument("--learned-n2", type=int, default=1024,
                    help="second grid side for the learned/plain comparison (0: none)")
    ap.add_argument("--config", type=int, default=1, choices=(1, 2, 3),
                    help="1: 512^2 pulse (the metric's workload); 2: 2048^2 DC, 5 cm, 1 Torr; "
                         "3: 256-member parameter-sweep ensemble")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--solo-sample", type=int, default=16)
    args = ap.parse_args()
    if args.n == 0:
        args.n = 2048 if args.config == 2 else 512
    if args.config == 2:
        args.no_cpu = True  # a CPU step at 2048^2 takes about a minute; not sampled
        args.no_learned = True
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: one discharge per device; "
                                   "multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.config == 3:
        e = ensemble_run(args)
        rate = e["accepted"] / e["elapsed"]
        print(json.dumps({
            "metric": "aggregate RK-Merson steps/s, 256-discharge parameter-sweep ensemble, FP64",
            "value": rate, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * e["elapsed"] / max(args.steps, 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (Gaussian + seeded 1% noise, member seed = id)",
            "config": {"workload": f"{args.members} argon discharges, sides U{{64..256}}, "
                                   "150-400 V DC, 0.5-10 Torr log-uniform, gamma 0.05-0.15",
                       "members": args.members, "cells_total": e["cells"],
                       "poisson_initial_guess": WARM_TEXT[warm_mode(args)],
                   "hip_graphs": not args.no_graphs},
            "accepted_steps": e["accepted"], "ptcl_attempts": e["attempts"],
            "member_passes_per_s": args.members * args.steps / e["elapsed"],
            "solo_sequential_steps_per_s": e["solo_rate"], "solo_sample": e["solo_sample"],
            "batched_over_sequential": rate / e["solo_rate"] if e["solo_rate"] else None,
            "setup_s": e["setup_s"], "stage_tile": f"64x{e['stage_tile']}",
            "cpu_baseline": None}))
        return
    r = gpu_run(args)
    steps_per_s = r["accepted"] / r["elapsed"]
    N = r["N"]
    plane = 8.0 * N
    kern = r["kernels"]
    # Algorithmic bytes (one FP64 plane = 8*N bytes).  Transport per accepted
    # step: 133 planes (RHS 9 per stage x 5 + RK updates 22 per species x 4),
    # carried by the five launches of the fused stage kernel.
    bytes_step = 133 * plane
    t_stages = sum(kern[f"stage{s}"] for s in (1, 2, 3, 4, 5)) * 1e-3
    ach_s = bytes_step / t_stages / 1e9
    stage_traffic = profiled_traffic("k_stage", args.n)
    # multigrid finest level with the fused CG update, timed in its
    # first-iteration form: down reads r, w, d, cx, cy, u, x, writes p, s, x, r,
    # x1 and the quarter-size coarse residual (12.25 planes; later iterations
    # also read s and p: 14.25); up reads x1, r, d, cx, cy and the coarse
    # correction, writes u and w = M u (7.25 planes)
    ach_d = 12.25 * plane / (kern["mg_down0"] * 1e-3) / 1e9
    ach_u = 7.25 * plane / (kern["mg_up0"] * 1e-3) / 1e9
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
        "config": {"workload": (f"argon DC-glow pulse, {args.n}x{args.n} FP64, 2 Torr, 400 V "
                                f"trapezoid, geometric x grid" if args.config == 1 else
                                f"argon DC glow, {args.n}x{args.n} FP64, 5 cm x 5 cm uniform, "
                                f"1 Torr, 300 V through 5e4 ohm"),
                   "nx": args.n, "ny": args.n, "poisson_rtol": args.rtol,
                   "poisson_solver": "CG preconditioned by an aggregation multigrid V(2,2) "
                                     "cycle with red-black Gauss-Seidel smoothing",
                   "poisson_initial_guess": WARM_TEXT[warm_mode(args)],
                   "hip_graphs": not args.no_graphs},
        "accepted_steps": r["accepted"],
        "ptcl_attempts": r["attempts"],
        "pcg_iters_mean": sum(iters) / max(len(iters), 1),
        "stage_tile": f"64x{r['stage_tile']}",
        "time_split_ms_per_step": {"stage_kernels": timers[0], "circuit_controller": timers[1],
                                   "assembly": timers[2], "poisson_solve": timers[3],
                                   "wall_with_events": timers[4], "steps": timers[5]},
        "kernel_ms": kern,
        "roofline": {"kernel": "k_stage (5 launches per step)", "bound": "hbm",
                     "achieved": ach_s, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": ach_s / PEAK_HBM_GBS,
                     "traffic": None if stage_traffic is None else stage_traffic,
                     "algorithmic_bytes_per_step": bytes_step},
        "roofline_poisson": {"kernels": "k_mg_down / k_mg_up, finest level", "bound": "hbm",
                             "achieved_down": ach_d, "achieved_up": ach_u,
                             "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac_down": ach_d / PEAK_HBM_GBS, "frac_up": ach_u / PEAK_HBM_GBS},
        "hbm_copy_GBs_measured": r["copy_gbs"],
        "poisson": r["poisson"],
    }
    if args.config == 2:
        line["metric"] = "RK-Merson steps/s at 2048^2 FP64 (large single instance)"
    if not args.no_cpu:
        cb = cpu_baseline(args, r["snap"])
        line["cpu_baseline"] = cb
        same = r["same"]
        same["over_cpu_jacobi"] = same["value"] / cb["value_jacobi"]
        line["gpu_same_poisson_method"] = same
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
