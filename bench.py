"""Benchmark: aggregate simulation steps per second of the HIP batch engine, bit-exact portable numerics (FP64).

Workload: the 1,024-instance titanic64 batch (scenario "titanic", dt 0.25 s): per instance a log-uniform [0.5, 2]
scale on every connection kind's area and each watertight door opened with p = 0.2, run to foundering or the 4 h
cap, with the 30 s readout curves recorded (sinksim_amd.workloads.titanic_batch_sweep, seed 20261008).

One bench step = one whole batch run (Batch.run(): every instance to its end, inputs already in HBM). value =
sum over instances of steps / wall time of the timed runs. The roofline is that of the dominant kernel (run_batch,
FP64): algorithmic FP64 operations per simulation step (measured with the oracle's operation counter on a sample of
the same workload) x steps, over the kernel's HIP-event duration. The CPU baseline is the reference's own CPU
engine (oracle/_ref/sinksim_sweep, compiled from the reference's sources) on a bounded sample of the same sweep;
if that binary is absent, the oracle's restatement.

    python bench.py [--steps K] [--warmup W] [--count 1024] [--numerics portable|portable32] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector/matrix peak (FMA counted as two operations)
# Algorithmic FP64 operations per simulation step on this workload, counted by the oracle's operation counter over
# the first 32 instances of the sweep, each to its own end (per step: 4,062 wet and 346 dry column evaluations, 2,865
# partial, 1,229 full and 5,258 dry segment evaluations, 70.4 node passes, 44.3 active connections, 24.1 pow calls)
# and priced by oracle/pyoracle.py:flops_per_step: 2.34e5. tests/test_bench_constants.py pins the counter against an
# independent count of the reference's own kernels on the 1912 run and re-counts a sample of this workload.
FLOPS_PER_STEP = 2.34e5
# HBM bytes per launch on this workload, a launch being one epoch of the resident run (one graph replay: the batch
# kernel at 256, 512 and 1,024 threads per simulation, all but one of them idle, and plan_epoch), from FETCH_SIZE +
# WRITE_SIZE summed over the epoch's kernels of a profiled run and calibrated against a plain copy kernel (FETCH_SIZE
# x 2) (profiles/run_batch_fp64_titanic64_batch1024.md).
TRAFFIC_BYTES_PER_LAUNCH = 13.55e6
TRAFFIC_SOURCE = "profiles/run_batch_fp64_titanic64_batch1024.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="titanic", choices=["titanic", "uneven", "adjoint", "calibration", "calibration-titanic",
                                                           "calibration-starts"],
                    help="titanic: the 1,024-instance titanic64 batch (the headline); uneven: the persistent batch of "
                         "seeded synthetic hulls of uneven size; adjoint: gradients of the twin-experiment loss of 256 "
                         "titanic64 instances; calibration: twin experiments with 24 kind-zone factors on the 16,680-"
                         "column calibration hull, the 1,320-run search per problem against batched L-BFGS with adjoint "
                         "gradients, device time per problem; calibration-titanic: the reference's search against "
                         "gradient L-BFGS on the reference objective (one problem); calibration-starts: one twin problem "
                         "of the calibration hull, the search against L-BFGS from 8,192 starts with a portable32 forward")
    ap.add_argument("--starts", type=int, default=None, help="L-BFGS starts (calibration-titanic: 256; calibration-starts: 8192)")
    ap.add_argument("--problems", type=int, default=256, help="calibration: twin problems the L-BFGS batch calibrates")
    ap.add_argument("--search-problems", type=int, default=8, help="calibration: problems the search also runs on")
    ap.add_argument("--iterations", type=int, default=None, help="calibration: adjoint calls of the L-BFGS at most (default 16; 12 titanic)")
    ap.add_argument("--checkpoint", type=int, default=256, help="adjoint: steps between checkpoints (0: automatic)")
    ap.add_argument("--hull", default="titanic", choices=["titanic", "calibration"],
                    help="adjoint: titanic64 twin experiments (12 parameters), or the calibration hull's twin problems "
                         "(24 kind-zone factors, 20,000 steps)")
    ap.add_argument("--polish-rounds", type=int, default=8, help="calibration: forward-only sampling rounds after the L-BFGS")
    ap.add_argument("--exact", action="store_true", help="adjoint: the discrete program's derivative instead of the relaxed one")
    ap.add_argument("--count", type=int, default=None, help="instances (default 1024 titanic, 4096 uneven)")
    ap.add_argument("--tmax", type=float, default=None, help="s (default 4 h titanic, 2 h uneven)")
    ap.add_argument("--numerics", default="portable", choices=["portable", "portable32"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=8, help="instances per host thread in the CPU-baseline sample")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first fill and copy MIB MiB with the probe library's plain kernels (ssp_fill, ssp_copy): the "
                         "known byte counts that calibrate a profiler's FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sw, sim, per_thread: int, numerics: str) -> dict:
    """The reference's CPU sweep runner (or the oracle's restatement) on the first `per_thread` x cores instances of the
    same sweep, every host core of a one-GPU host's share, in portable and portable32 numerics. The rate is the tool's
    own aggregate line (run_cpu alone: no file IO, no process start); the leg of the bench's numerics is the value."""
    import re
    from sinksim_amd import sweep as S
    from sinksim_amd.data_paths import SHIP, sim_path
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    cores = min(cores, 16)  # a one-GPU host's CPU share
    count = min(len(sw.instances), per_thread * cores)
    sample = S.Sweep(ship=sw.ship, sim=sw.sim, tMax=sw.tMax, every=sw.every, instances=sw.instances[:count])
    ref = os.path.join(ROOT, "oracle", "_ref", "sinksim_sweep")
    legs = {}
    if os.path.exists(ref):
        kind = "reference"
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "sample.sweep.json")
            S.save_sweep(sample, path, sim)
            for num in ("portable", "portable32"):
                out = subprocess.run([ref, SHIP, sim_path(sim.id), path, "--numerics", num, "--threads", str(cores)],
                                     check=True, capture_output=True, text=True).stdout
                m = re.search(r"(\d+) steps in ([0-9.]+) s \(([0-9]+) steps/s aggregate\)", out)
                if not m:
                    raise RuntimeError(f"no aggregate line in the CPU sweep's output:\n{out}")
                legs[num] = (int(m.group(1)), float(m.group(2)))
    else:
        kind = "port"
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        from sinksim_amd.data_paths import titanic
        ship, _ = titanic(sim.id)
        en, area, tOn = S.connection_tables(sample, sim)
        for num in ("portable", "portable32"):
            t0 = time.perf_counter()
            out = pyoracle.sweep(ship, sim, num, en, area, tOn, sw.tMax, sw.every, threads=cores)
            legs[num] = (sum(o["steps"] for o in out), time.perf_counter() - t0)
    steps, secs = legs[numerics]
    return {"value": round(steps / secs, 1), "unit": "sim-steps/s", "cores": cores, "kind": kind, "cpu": _cpu_model(),
            "legs": {k: round(v[0] / v[1], 1) for k, v in legs.items()},
            "sample": f"first {count} instances ({per_thread} per thread) of the same 1,024-instance sweep; {numerics}: "
                      f"{steps} steps in {secs:.1f} s (the sweep's run alone, no file IO)"}


def _device() -> str:
    from sinksim_amd.batch import device_info
    d = device_info()
    return f"{d['name'] or 'AMD GPU'} ({d['arch']}, {d['computeUnits']} CUs)"


def bench_uneven(args):
    """The uneven persistent batch: `count` seeded synthetic hulls, one instance each, in one refilling launch.
    Roofline: HBM, with 40 bytes per column (x, y, zlo, zhi, dxdy as doubles) and 20 per segment (bottom, top and
    column index) per simulation step as the algorithmic geometry traffic; the hulls together are far larger than
    the caches, so each step streams its hull."""
    import numpy as np
    from sinksim_amd.batch import Batch, BatchOptions
    from sinksim_amd.workloads import uneven_fleet
    count = args.count or 4096
    t_max = args.tmax or 2 * 3600.0
    mixed = args.numerics == "portable32"
    t0 = time.perf_counter()
    models, inst_model = uneven_fleet(count)
    print(f"{count} hulls generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    geo = np.array([40.0 * m[0].columns + (20.0 if not mixed else 12.0) * m[0].segments for m in models])
    opts = BatchOptions(count=count, mixed_precision=mixed, traced_instances=0, curve_every=30.0, t_max=t_max)
    # the launch shape, chosen once by the autotuner on the largest hulls (untimed)
    b = Batch.fleet(models, inst_model, opts)
    tune = b.autotune(sample=min(count, 2048), trial_steps=300)
    b.close()
    print(f"autotune: {tune}", file=sys.stderr, flush=True)

    def one_run():
        b = Batch.fleet(models, inst_model, opts)
        b.set_launch(tune["threads_per_sim"], tune["workgroups"])
        b.set_record_layout(tune["record_bytes"])
        secs = b.run()
        steps = b.results()["steps"]
        kms, launches = b.kernel_time()
        b.close()
        return secs, steps, kms, launches

    for w in range(args.warmup):
        secs, steps, _, _ = one_run()
        print(f"warmup {w}: {int(steps.sum())} steps in {secs:.3f} s", file=sys.stderr, flush=True)
    tot_s = tot_kms = 0.0
    tot_steps = tot_bytes = 0.0
    tot_launch = 0
    for k in range(args.steps):
        secs, steps, kms, launches = one_run()
        tot_s += secs; tot_kms += kms; tot_launch += launches
        tot_steps += float(steps.sum()); tot_bytes += float((steps * geo[inst_model]).sum())
        print(f"step {k}: {int(steps.sum())} steps in {secs:.3f} s ({steps.sum() / secs:.0f} steps/s)", file=sys.stderr, flush=True)
    value = tot_steps / tot_s
    achieved = tot_bytes / (tot_kms * 1e-3) / 1e9
    cols = np.array([m[0].columns for m in models]); sp = np.array([m[0].nodes for m in models])
    nc = np.array([m[1].connections for m in models])
    out = {
        "metric": "aggregate sim-steps/s, uneven persistent batch", "value": round(value, 1), "unit": "sim-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * tot_s / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64" if not mixed else "f32+f64",
        "data": "synthetic (seeded synthetic hulls)",
        "config": {"workload": f"{count} seeded synthetic hulls ({int(cols.min())}-{int(cols.max())} columns, {int(sp.min())}-{int(sp.max())} "
                               f"spaces, {int(nc.min())}-{int(nc.max())} connections), one instance each, {args.numerics} "
                               f"numerics, to founder or {t_max / 3600:g} h, curves every 30 s, refilled workgroups",
                   "instances": count, "numerics": args.numerics, "sim_steps_per_bench_step": int(tot_steps // args.steps),
                   "threads_per_sim": tune["threads_per_sim"], "workgroups": tune["workgroups"] or "one per instance or all resident",
                   "record_bytes": tune["record_bytes"],
                   "autotune_trials": [[t["threads_per_sim"], t["workgroups"], t["record_bytes"], round(t["rate"])] for t in tune["trials"]],
                   "device": _device()},
        # traffic: FETCH_SIZE + WRITE_SIZE of the timed launch on this workload at 4,096 hulls with the packed records
        # (FETCH_SIZE unscaled: profiles/run_batch_fp64_uneven4096.md); a constant here, the run itself is not counted
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 5),
                     "traffic": 43.20e12 if (count == 4096 and not mixed and tune["record_bytes"] == 40) else None,
                     "traffic_source": "profiles/run_batch_fp64_uneven4096.md" if (count == 4096 and not mixed and tune["record_bytes"] == 40) else None,
                     "kernel": "run_batch<%s, true>" % ("true" if mixed else "false"), "kernel_ms": round(tot_kms / max(tot_launch, 1), 3)},
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Gradients of the twin-experiment loss: `count` titanic64 instances (default tables), 12 scale factors (the area
    of breach, over, down and top; the discharge coefficient of every kind but open), loss = squared trim error at
    15, 30, 45 and 75 min and squared error of the times of overF, well and bridge, weighted as the reference's
    objective weighs them, against observations from seeded hidden parameters; the gradient at an independent seeded
    guess, relaxed level derivatives, a checkpoint every K steps, to founder or 260 min. One bench step = one gradient
    of every instance (ss_batch_gradient). value = gradients per second."""
    import numpy as np
    from sinksim_amd.batch import Batch, BatchOptions
    from sinksim_amd.calib import kind_params, titanic_loss
    from sinksim_amd.data_paths import titanic
    from sinksim_amd.workloads import twin_experiment
    count = args.count or 256
    if args.hull == "calibration":
        from sinksim_amd.calibrate import twin_problems
        from sinksim_amd.workloads import calibration_problems
        tp = twin_problems(count)
        ship, sim, pm, spec = tp.ship, tp.sim, tp.params, tp.spec
        t_max = tp.steps * sim.dt
        truth = np.exp(tp.truth)
        guess = np.exp(calibration_problems(count, seed=7, nparams=pm.count))
        label = (f"{count} calibration-hull twin problems (16,680 columns, 128 spaces, 600 connections), 24 kind-zone factors, "
                 f"trim at 20/40/60/80 min + 34 event times, {tp.steps} steps")
    else:
        t_max = args.tmax or 260 * 60.0
        ship, sim = titanic("titanic")
        pm = kind_params(sim)
        spec = titanic_loss(sim)
        truth, guess = twin_experiment(count, pm.count)
        label = (f"{count} titanic64 instances, twin experiment, 12 scale factors, trim at 15/30/45/75 min + overF/well/bridge "
                 f"times, to founder or 260 min")
    b = Batch(ship, sim, BatchOptions(count=count, mixed_precision=False, traced_instances=0, curve_every=0.0, t_max=t_max))
    twin = b.gradient(pm, truth, spec, t_max=t_max, want_grad=False)
    obs = twin["signals"]
    print(f"observations from the hidden parameters: forward {twin['ms']:.0f} ms, {int(twin['steps'].sum())} steps", file=sys.stderr, flush=True)
    fwd = b.gradient(pm, guess, spec, observations=obs, t_max=t_max, want_grad=False)
    print(f"forward only at the guess: {fwd['ms']:.0f} ms", file=sys.stderr, flush=True)

    def one():
        return b.gradient(pm, guess, spec, observations=obs, checkpoint_every=args.checkpoint, relaxed=not args.exact, t_max=t_max)

    for w in range(args.warmup):
        r = one()
        print(f"warmup {w}: {r['ms']:.0f} ms", file=sys.stderr, flush=True)
    tot_ms = tot_wall = 0.0
    for k in range(args.steps):
        t0 = time.perf_counter()
        r = one()
        wall = time.perf_counter() - t0
        tot_ms += r["ms"]
        tot_wall += wall
        if int(r["mismatches"].sum()) != 0:
            raise RuntimeError(f"{int(r['mismatches'].sum())} recomputed steps differ from the forward trajectory")
        print(f"step {k}: {count} gradients in {r['ms']:.0f} ms kernel, {wall:.2f} s wall; mean loss {r['loss'].mean():.3f}",
              file=sys.stderr, flush=True)
    b_launch = b.gradient_launch()
    b.close()
    steps = int(r["steps"].sum())
    secs = tot_ms * 1e-3 / args.steps
    out = {
        "metric": f"adjoint gradients/s (twin-experiment loss, {pm.count} parameters)", "value": round(count / secs, 3),
        "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * tot_wall / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": f"synthetic (seeded hidden parameters of the {args.hull} model)",
        "config": {"workload": f"{label}, relaxed={not args.exact}, checkpoint every {args.checkpoint} steps (0: automatic)",
                   "instances": count, "sim_steps": steps, "launch": b_launch,
                   "forward_only_ms": round(fwd["ms"], 1), "gradient_over_forward": round(tot_ms / args.steps / fwd["ms"], 2),
                   "device": _device()},
    }
    print(json.dumps(out), flush=True)


def bench_twin_calibration(args):
    """Calibration on many problems (twin experiments on the calibration hull, 24 kind-zone area factors, 20,000
    steps): the reference's search (5 rounds x 264 runs per problem) on `search-problems` of them against batched
    L-BFGS with adjoint gradients on all `problems` from scales 1, then forward-only sampling around each problem's
    best point, until every searched problem reaches its search's best objective or the rounds run out. value =
    search device time per problem / gradient-method device time per problem when half of the searched problems
    (the median problem) have reached their search's objective; the same ratio for all of them, and the share reached
    within a tenth of the search's time, are reported beside it."""
    import numpy as np
    from sinksim_amd.calibrate import twin_compare

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    r = twin_compare(problems=args.problems, search_problems=args.search_problems, iterations=args.iterations or 16,
                     polish_rounds=args.polish_rounds, log=log)
    sw, lb, tp = r["search"], r["lbfgs"], r["problems"]
    per_s, per_all, per_half = r["search_ms_per_problem"], r["ms_per_problem_all"], r["ms_per_problem_half"]
    calls = len(lb.device_ms) - 1
    tenth = [v for v, m in zip(r["share_reached"], r["device_ms"]) if m / args.problems <= per_s / 10]
    out = {
        "metric": "calibration device time per problem, search / adjoint L-BFGS at equal objective (median problem)",
        "value": round(per_s / per_half, 2) if per_half else None, "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * r["wall_s"], 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic: seeded hidden factors of twin experiments, observations from their runs",
        "config": {"workload": f"calibration hull (16,680 columns, 128 spaces, 600 connections), {tp.steps} steps, 24 kind-zone area "
                               f"factors, trim at 20/40/60/80 min + 34 event times, {args.problems} problems",
                   "search": {"problems": len(r["target"]), "runs_per_problem": sw["evaluations"], "device_s": round(sw["device_ms"] / 1e3, 3),
                              "device_s_per_problem": round(per_s / 1e3, 4), "sim_steps": sw["steps"],
                              "median_J_by_round": [float(f"{v:.4g}") for v in np.median(sw["J_by_round"], axis=1)],
                              "target_J": [float(f"{v:.4g}") for v in r["target"]]},
                   "gradient": {"problems": args.problems, "lbfgs_calls": calls,
                                "polish_rounds": r["phase"].count("polish"), "polish_runs_per_round": 16,
                                "device_s": round(r["device_ms"][-1] / 1e3, 3),
                                "device_s_per_problem_all_reached": round(per_all / 1e3, 4) if per_all else None,
                                "device_s_per_problem_half_reached": round(per_half / 1e3, 4) if per_half else None,
                                "lbfgs_sim_steps": lb.steps[-1],
                                "share_reached": [round(v, 3) for v in r["share_reached"]],
                                "phase": r["phase"],
                                "device_s_cumulative": [round(v / 1e3, 2) for v in r["device_ms"]],
                                "median_J": [float(f"{v:.4g}") for v in r["median_J"]]},
                   "ratio_all_reached": round(per_s / per_all, 2) if per_all else None,
                   "share_reached_within_a_tenth_of_search_time": tenth[-1] if tenth else 0.0,
                   "device": _device()},
    }
    print(json.dumps(out), flush=True)


def bench_starts_calibration(args):
    """One twin problem of the calibration hull (24 kind-zone factors, 20,000 steps): the reference's search (5 rounds x
    264 runs) against batched L-BFGS from `starts` starts (default 8,192: scales 1 and seeded points of the search's
    first box), the forward in portable32 numerics and the reverse in FP64, all starts in every adjoint call. value =
    the search's device time / the L-BFGS device time to the search's objective (null if not reached within the
    iterations)."""
    from sinksim_amd.calibrate import starts_compare

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    starts = args.starts or 8192
    r = starts_compare(starts=starts, iterations=args.iterations or 2, numerics=args.numerics if args.numerics != "portable" else "portable32",
                       log=log)
    sw, lb = r["search"], r["lbfgs"]
    reached = r["lbfgs_ms_to_target"]
    out = {
        "metric": "calibration device time, search / L-BFGS from many starts at equal objective (one problem)",
        "value": round(r["search_ms"] / reached, 4) if reached else None, "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * r["wall_s"], 1), "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic: seeded hidden factors of one twin problem, observations from its run",
        "config": {"workload": f"calibration hull (16,680 columns, 128 spaces, 600 connections), {r['problems'].steps} steps, "
                               f"24 kind-zone area factors, one problem, {starts} L-BFGS starts, portable32 forward, FP64 reverse",
                   "search": {"runs": sw["evaluations"], "device_s": round(sw["device_ms"] / 1e3, 3),
                              "J_by_round": [float(f"{v:.4g}") for v in sw["J_by_round"][:, 0]], "target_J": float(f"{r['target']:.4g}")},
                   "lbfgs": {"starts": starts, "calls": len(lb.best_J), "device_s_by_call": [round(v / 1e3, 2) for v in lb.device_ms],
                             "best_J_by_call": [float(f"{v:.4g}") for v in lb.best_J],
                             "device_s_to_target": round(reached / 1e3, 3) if reached else None},
                   "device": _device()},
    }
    print(json.dumps(out), flush=True)


def bench_calibration(args):
    """The reference's calibration (tools/py/calibrate_gpu.py: 5 rounds of 264 log-uniform points around the best,
    span 0.3 shrinking by half, seed 20261008, objective J against the 1912 observations) on the GPU batch, then
    batched L-BFGS with adjoint gradients of the same J from scales 1 and seeded points of the first box, until it
    reaches the search's best J. value = the search's device time / the L-BFGS device time to that objective."""
    import numpy as np
    from sinksim_amd.calibrate import compare
    from sinksim_amd.data_paths import titanic
    ship, sim = titanic("titanic")

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    r = compare(ship, sim, starts=args.starts or 256, iterations=args.iterations or 12, log=log)
    sw, lb = r["search"], r["lbfgs"]
    reached = r["lbfgs_ms_to_search_J"]
    out = {
        "metric": "calibration device time, reference search / adjoint L-BFGS at equal objective",
        "value": round(r["search_ms"] / reached, 3) if reached else None, "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * r["wall_s"], 1), "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "titanic64 model, 1912 observations (data/observations/titanic-1912.json)",
        "config": {"workload": "titanic64 scenario titanic, 4 kind scales (breach over down top), objective J of the reference",
                   "search": {"evaluations": sw.evaluations[-1], "best_J": round(sw.best_J[-1], 4), "device_s": round(sw.device_ms[-1] / 1e3, 3),
                              "sim_steps": sw.steps[-1], "best_scales": [round(float(np.exp(v)), 4) for v in sw.best_x[-1]],
                              "J_by_round": [round(j, 4) for j in sw.best_J]},
                   "lbfgs": {"starts": args.starts or 256, "calls": len(lb.best_J), "best_J": round(lb.best_J[-1], 4),
                             "device_s": round(lb.device_ms[-1] / 1e3, 3), "device_s_to_search_J": round(reached / 1e3, 3) if reached else None,
                             "sim_steps": lb.steps[-1], "best_scales": [round(float(np.exp(v)), 4) for v in lb.best_x[-1]],
                             "J_by_call": [round(j, 4) for j in lb.best_J]},
                   "device": _device()},
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); instances are independent, so N GPUs "
              "would each run count/N of them", file=sys.stderr)
        sys.exit(2)
    if args.workload == "uneven":
        bench_uneven(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "calibration":
        bench_twin_calibration(args)
        return
    if args.workload == "calibration-titanic":
        bench_calibration(args)
        return
    if args.workload == "calibration-starts":
        bench_starts_calibration(args)
        return
    args.count = args.count or 1024
    from sinksim_amd import sweep as S
    from sinksim_amd.batch import Batch, BatchOptions
    from sinksim_amd.data_paths import titanic
    from sinksim_amd.workloads import titanic_batch_sweep

    ship, sim = titanic("titanic")
    sw = titanic_batch_sweep(sim, count=args.count, t_max=args.tmax or 4 * 3600.0)
    en, area, tOn = S.connection_tables(sw, sim)
    mixed = args.numerics == "portable32"

    def one_run():
        b = Batch(ship, sim, BatchOptions(count=args.count, mixed_precision=mixed, traced_instances=0,
                                          curve_every=sw.every, t_max=sw.tMax))
        b.set_connections(en, area, tOn)
        secs = b.run()
        steps = b.total_steps()
        kms, launches = b.kernel_time()
        steps0 = b.result(0).steps
        b.close()
        return secs, steps, kms, launches, steps0

    if args.copy_probe > 0:
        import ctypes
        probe = ctypes.CDLL(os.path.join(ROOT, "tests", "probe", "libsinksim_probe.so"))
        probe.ssp_copy_probe.argtypes = [ctypes.c_longlong]
        probe.ssp_last_error.restype = ctypes.c_char_p
        if probe.ssp_copy_probe(args.copy_probe << 20) != 0:
            raise RuntimeError(probe.ssp_last_error().decode())
        print(f"copy probe: {args.copy_probe} MiB filled and copied", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        secs, steps, _, _, _ = one_run()
        print(f"warmup {w}: {steps} steps in {secs:.3f} s", file=sys.stderr, flush=True)
    tot_s, tot_steps, tot_kms, tot_launch = 0.0, 0, 0.0, 0
    for k in range(args.steps):
        secs, steps, kms, launches, steps0 = one_run()
        tot_s += secs
        tot_steps += steps
        tot_kms += kms
        tot_launch += launches
        print(f"step {k}: {steps} steps in {secs:.3f} s ({steps / secs:.0f} steps/s), kernel {kms:.1f} ms over "
              f"{launches} launch(es)", file=sys.stderr, flush=True)
    value = tot_steps / tot_s
    per_launch_ms = tot_kms / max(tot_launch, 1)
    flops_per_launch = FLOPS_PER_STEP * tot_steps / max(tot_launch, 1)
    achieved = flops_per_launch / (per_launch_ms * 1e-3) / 1e12
    out = {
        "metric": "aggregate sim-steps/s bit-exact at batch 1024",
        "value": round(value, 1),
        "unit": "sim-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * tot_s / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not mixed else "f32+f64",
        "data": "synthetic (seeded sweep of the titanic64 model)",
        "config": {"workload": f"titanic64 batch of {args.count} seeded instances, {args.numerics} numerics, to founder "
                               f"or 4 h, curves every 30 s", "instances": args.count, "numerics": args.numerics,
                   "sim_steps_per_bench_step": tot_steps // args.steps,
                   # full sinkings per minute at instance 0's length, as apps/sinksim_cuda_run.cpp:138-140 reports it
                   "sinkings_per_min": round(value / max(steps0, 1) * 60.0, 1), "device": _device()},
        # the FP64 vector ALU bounds this kernel (no dense contraction, so no MFMA); 78.6 TFLOP/s is MI355X's FP64
        # vector peak with an FMA counted as two operations
        "roofline": {"bound": "valu-fp64", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                     # HBM bytes per launch from the FETCH_SIZE and WRITE_SIZE counters of a profiled run of this
                     # workload, scaled by a copy kernel's reading (a constant here; the run itself is not counted)
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH if (args.count == 1024 and not mixed) else None,
                     "traffic_source": TRAFFIC_SOURCE if (args.count == 1024 and not mixed) else None,
                     # a launch is one epoch: run_batch at 256, 512 and 1,024 threads per simulation (all but one
                     # idle) and plan_epoch, one graph replay; kernel_ms is the launches' HIP-event time over their count
                     "kernel": f"run_batch<{'true' if mixed else 'false'}, false, 256|512|1024> + plan_epoch, per epoch",
                     "launches": tot_launch // max(args.steps, 1), "kernel_ms": round(per_launch_ms, 3)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sw, sim, args.cpu_sample, args.numerics)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
