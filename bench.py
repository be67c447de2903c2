"""Benchmark: FDTD cell-updates per second of batched plain SIRT on one MI355X.

A step is one full SIRT iteration over all shots of the workload: forward
propagation with residual record, checkpointed segment recomputation, the
adjoint sweep fused with the gradient correlation, the field update and the
error.  ``value`` counts every FDTD cell-update of the step (forward,
recompute and adjoint) per second of wall time, with all inputs resident in
HBM.

Default workload (--workload c3): 512 x 512 grid, 128 ring transducers fired
one at a time (128 shots, radius 0.39 n), 4000 time steps, forward wavefield
checkpointed every 200 steps, smooth random velocity (Gaussian-filtered noise,
sigma 8 cells, +-8 %) plus 4 sharp inclusions; measured traces from the FP64
solver.

Also reported: the roofline of the dominant stencil kernel (the average
launch time from HIP events on the engine's stream, with three labelled
fractions of the HBM peak: at the per-unit bytes of the one-step
formulation, at the bytes the blocked kernels must move, and at the HBM bytes
per launch measured by the committed counter summary), both stencil kernels'
figures, and the CPU baseline (the oracle's FP32 transcription built for the
host CPU, OpenMP over shots) on a bounded sample of the same workload.

No process that runs the engine imports torch: torch ships its own HIP and
HSA runtimes under the same sonames as ROCm's, and the engine would bind to
those instead of the ROCm runtime it is built against and the profiler
instruments.  The calibration copy uses the engine's own HIP runtime.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

from _pkg import tomo  # noqa: E402
import tomo_amd.workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per cell-update of one shot that the blocked stencil kernels must
# move (fractions.blocked_algorithmic): the coefficient image is shared by every shot of a
# launch (4/shots B, L2-resident); the two-step forward kernel reads u(k),
# u(k-1) and writes u(k+1), u(k+2) once per two steps (8 B), the one-step
# kernel 12 B; with three steps a recording pass writes only u(k+2), u(k+3)
# (16 B per three updates) and a recompute pass all three levels (20 B per
# three), 6 B on average over the two; the two-step adjoint reads z(k+1),
# z(k+2), u(k), u(k-1) and the per-shot gradient and writes z(k), z(k-1) and
# the gradient once per two steps (16 B), the one-step adjoint re-reads and
# re-writes the gradient each step (24 B).
FWD_BYTES = {1: 12.0, 2: 8.0, 3: 6.0}
ADJ_BYTES = {1: 24.0, 2: 16.0, 3: 12.0}
# The per-unit figures of the one-step formulation (roofline.achieved): a
# forward cell-update reads u(k), u(k-1) and the per-cell coefficient and
# writes u(k+1) (16 B); an adjoint cell-update fused with the correlation
# reads z(k+1), z(k+2), the coefficient and u(k) and writes z(k), with the
# gradient in registers (20 B).  With several steps per launch these exceed
# what the kernels must move (the forward's then exceeds the HBM peak).
FWD_UNIT_BYTES = 16.0
ADJ_UNIT_BYTES = 20.0
# Measured HBM bytes per launch of the stencil kernels (rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE of a --counter-pass run, scaled by the counters'
# reading for a 1 GiB device copy).
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "stencil_hbm_traffic_c3_512x512x128.md")

WORKLOADS = {
    # name: (n, steps, shots, checkpoint)
    "c3": (512, 4000, 128, 200),
    "c2": (256, 2000, 64, 0),
    "c1": (128, 1000, 32, 0),
}


def hip_runtime():
    """The HIP runtime the engine library is linked against (the process's
    libamdhip64.so.7, loaded with the library)."""
    tomo.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def device_copy(nbytes):
    """One device-to-device copy of nbytes (a plain copy of known size: its
    counter readings scale FETCH_SIZE / WRITE_SIZE)."""
    hip = hip_runtime()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipDeviceSynchronize() or hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize():
            raise RuntimeError("device copy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def native_oracle():
    """The oracle built for this host's CPU (-march=native, FMA contraction
    allowed): the timed CPU baseline.  Built here, on the machine that runs
    the bench, since the CPU that runs it is not known in advance."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("native oracle build failed:\n" + r.stdout + r.stderr)
    return os.path.join(ROOT, "oracle", "_build", "liboracle_native.so")


def make_workload(name):
    c, ii, jj, fo = W.problem(name)
    return c, ii, jj, fo


def cpu_baseline(c, ii, jj, fo, meas, threads, shots):
    """Oracle FP32 gradient of `shots` shots of the workload (same geometry,
    steps and checkpointing), OpenMP over shots, on the oracle built for this
    host (native_oracle)."""
    import oracle as orc
    orc.set_threads(threads)
    bbox = np.stack([ii, ii, jj, jj], axis=1).astype(np.int32).ravel()[: 4 * shots]
    cells = (ii + c.nx * jj).astype(np.int32)
    m32 = np.ascontiguousarray(meas[:shots], np.float32)
    f0 = np.zeros(fo.size, np.float32)
    t = time.perf_counter()
    orc.gradient(c, f0, bbox, cells, cells, m32, np.float32)
    return time.perf_counter() - t


def _speedup_summary(sp, seeds):
    """Speedups of one method to one target over the seeds (None: never
    reached, counted as 0 in the median)."""
    reached = [x for x in sp if x is not None]
    return {"speedups": sp, "reached": f"{len(reached)} of {seeds}",
            "at_least_3x": sum(1 for x in reached if x >= 3.0),
            "median_speedup": float(np.median([x if x is not None else 0.0 for x in sp])) if sp else None}


def ttt_summary(runs):
    """The time-to-target study over its seeds: per target, the learned run's
    speedups over plain SIRT and, when the runs carry momentum, those of
    plain SIRT with the same momentum and no network.  Seeds whose plain run
    never gets below 100 % are flagged and left out."""
    out = {"study": "ttt", "runs": runs,
           "flagged_never_below_100": [r["phantom_seed"] for r in runs if r["plain_never_below_100"]]}
    counted = [r for r in runs if not r["plain_never_below_100"]]
    for name in ("fixed50", "stop_rule", "fixed10"):
        out[name] = _speedup_summary([r["targets"][name]["speedup"] for r in counted], len(counted))
        if counted and counted[0]["momentum"]:
            out[name]["momentum_plain"] = _speedup_summary(
                [r["targets"][name]["momentum_plain_speedup"] for r in counted], len(counted))
    return out


def load_traffic(kernel, shape, tb):
    """HBM bytes per launch of `kernel` at this stencil shape and time
    blocking, from the committed counter summary (None when it has no row
    for them).  Rows read `| kernel | shape | steps per launch | ... |
    bytes per launch | ...`."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            lines = fh.read().splitlines()
    except OSError:
        return None
    want_shape = ",".join(str(v) for v in shape)
    for ln in lines:
        cols = [v.strip().strip("`") for v in ln.strip().strip("|").split("|")]
        if len(cols) < 7 or not cols[0].startswith(kernel + "<"):
            continue
        if cols[1] == want_shape and cols[2] == str(tb):
            try:
                return float(cols[6])
            except ValueError:
                return None
    return None


def kernel_line(kernel, unit_bytes, blocked_bytes, cells_launch, avg_ms, total_ms, traffic, traffic_of):
    """Roofline figures of one stencil kernel: `achieved` at the per-unit
    bytes of the one-step formulation (forward 16 B, adjoint with its
    correlation 20 B per cell-update) times the cell-updates of one launch,
    over the average launch time; beside it the same launch time at the
    bytes the blocked kernels must move, and at the HBM bytes per launch the
    counters measured (traffic; None without a counter row for this shape)."""
    sec = avg_ms * 1e-3
    unit = unit_bytes * cells_launch / sec / 1e9
    blocked = blocked_bytes * cells_launch / sec / 1e9
    counted = traffic / sec / 1e9 if traffic else None
    return {
        "kernel": kernel,
        "achieved": round(unit, 1), "frac": round(unit / HBM_PEAK_GBS, 4), "traffic": traffic,
        "traffic_of": traffic_of,
        "fractions": {"per_unit": round(unit / HBM_PEAK_GBS, 4),
                      "blocked_algorithmic": round(blocked / HBM_PEAK_GBS, 4),
                      "counter_measured": round(counted / HBM_PEAK_GBS, 4) if counted else None},
        "bytes_per_cell_update": {"per_unit": unit_bytes, "blocked_algorithmic": blocked_bytes,
                                  "counter_measured": round(traffic / cells_launch, 3) if traffic else None},
        "cells_per_launch": cells_launch, "avg_launch_ms": round(avg_ms, 5), "total_ms": round(total_ms, 2),
    }


def cpu_model():
    """The host CPU as /proc/cpuinfo names it (model name, family / model /
    stepping, sockets), with the logical CPU count and this process's
    affinity, or None."""
    info = {}
    sockets = set()
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ":" not in ln:
                    continue
                k, v = (x.strip() for x in ln.split(":", 1))
                if k == "physical id":
                    sockets.add(v)
                elif k in ("model name", "cpu family", "model", "stepping", "vendor_id") and k not in info:
                    info[k] = v
    except OSError:
        return None
    if not info:
        return None
    return (f"{info.get('model name', '?')} (vendor {info.get('vendor_id', '?')}, family {info.get('cpu family', '?')}, "
            f"model {info.get('model', '?')}, stepping {info.get('stepping', '?')}; {max(len(sockets), 1)} socket(s), "
            f"{os.cpu_count()} logical CPUs, {len(os.sched_getaffinity(0))} in this process's affinity)")


def counter_pass(args):
    """One FP32 forward (data) and one gradient (or, --forward-only, the
    recording forward pass of one) of the C3 workload, then a 1 GiB device
    copy: the program of the rocprofv3 --pmc runs that measure the stencil
    kernels' HBM bytes per dispatch (FETCH_SIZE, WRITE_SIZE; the copy's
    reading gives the counters' scale)."""
    log = (lambda m: print("[counters]", m, file=sys.stderr, flush=True))
    c, ii, jj, fo = make_workload("c3")
    e = tomo.Engine(c, device=0)
    if args.shape:
        e.set_shape(*[int(v) for v in args.shape.split(",")])
    e.set_time_blocking(args.time_blocking)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    e.set_field(fo)
    log("forward")
    meas = e.forward(fp64=False)
    e.set_field(np.zeros(fo.size, np.float32))
    e.set_measured(meas)
    if args.forward_only:
        log("misfit (recording forward pass)")
        e.misfit()
    else:
        log("gradient")
        e.gradient()
    shape = list(e.get_shape())
    tb, ab = e.time_blocking(), e.adjoint_blocking()
    e.close()
    log("copy 1 GiB")
    device_copy(1 << 30)
    print(json.dumps({"counter_pass": "c3", "stencil_shape": shape, "forward_steps_per_launch": tb,
                      "adjoint_steps_per_launch": ab, "copy_bytes": 2 << 30}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c4", "ref"],
                    help="c1-c4: the synthetic configurations; ref: the reference's own preset and Table 2 case "
                         "(192^2, 401 steps, 640 transducers in 64 groups, full history)")
    ap.add_argument("--c4-mode", default="stacked", choices=["stacked", "contexts"],
                    help="C4: one engine per side with that side's phantoms stacked along the shot axis (one launch "
                         "per level advances all of them), or one engine per phantom")
    ap.add_argument("--c4-threads", type=int, default=0,
                    help="host threads stepping the C4 engines concurrently (default: 2 stacked, 4 contexts)")
    ap.add_argument("--c4-passes", type=int, default=0,
                    help="C4 shots per pass (0: all shots in one pass, the phantoms of each host thread sharing "
                         "one wavefield workspace)")
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel from the host (default: each gradient pass is captured once into a "
                         "HIP graph and replayed)")
    ap.add_argument("--study", choices=["c2", "ttt", "train", "finetune", "relax"], default=None,
                    help="run a convergence study instead of the throughput benchmark (one JSON line)")
    ap.add_argument("--seeds", default="9100", help="held-out phantom seeds of the ttt study (comma separated)")
    ap.add_argument("--fp", type=float, default=0.0, help="--study c2: Ricker peak frequency (default 0.1)")
    ap.add_argument("--out", default="preconditioner_trained.npy", help="parameters written by --study train")
    ap.add_argument("--phantoms", type=int, default=48, help="training phantoms of --study train")
    ap.add_argument("--train-steps", type=int, default=4000, help="Adam steps of --study train")
    ap.add_argument("--train-target", default="best",
                    help="--study train targets: best (the trajectory's best iterate), ahead:K (K iterations on) "
                         "or truth (the true field)")
    ap.add_argument("--train-iters", type=int, default=50, help="--study train: plain iterations per phantom")
    ap.add_argument("--train-first-seed", type=int, default=7000,
                    help="--study train: first phantom seed (with --weights: continue training those parameters)")
    ap.add_argument("--stack", type=int, default=0,
                    help="--study train: phantoms per engine, stacked along the shot axis (0: one engine each)")
    ap.add_argument("--train-rounds", type=int, default=0,
                    help="--study train: rounds that add the learned iteration's own states (target: the truth) "
                         "and train on")
    ap.add_argument("--weights", default=None,
                    help="network parameters for --study ttt (default: the shipped ones); for --study train, the "
                         "parameters to continue from")
    ap.add_argument("--learned-iters", type=int, default=3,
                    help="--study ttt: learned iterations (with the plain fallback) before plain ones")
    ap.add_argument("--learned-alpha", type=float, default=1.0,
                    help="--study ttt: the learned iterations' mix, alpha * network step + (1 - alpha) * plain step")
    ap.add_argument("--momentum", default="",
                    help="--study relax: heavy-ball momenta to compare (comma separated); --study ttt: the momentum "
                         "of the learned run's plain iterations, with plain SIRT at the same momentum beside it; "
                         "--study train (with --stack): the momentum of the training trajectories")
    ap.add_argument("--autotune", action="store_true", help="choose the stencil shape with the autotuner")
    ap.add_argument("--time-blocking", type=int, default=3, choices=[1, 2, 3],
                    help="forward steps per launch (2, 3: the intermediate levels in registers)")
    ap.add_argument("--adjoint-blocking", type=int, default=3, choices=[1, 2, 3],
                    help="most adjoint steps per launch (3: z(k), z(k-1) in registers)")
    ap.add_argument("--shape", default="", help="stencil shape vec,fwd_rows,adj_rows,fwd_depth,adj_depth (default: engine choice)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--data", default="fp64", choices=["fp64", "fp32"],
                    help="solver of the measured traces (fp32 for counter-collection runs; the timed region is the same)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, copy 1 GiB device to device (a plain copy of known size, used to scale "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--counter-pass", action="store_true",
                    help="run one C3 forward and gradient plus a 1 GiB copy (the program for rocprofv3 --pmc)")
    ap.add_argument("--forward-only", action="store_true",
                    help="with --counter-pass: the recording forward pass (the misfit) instead of the gradient")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots would be sharded across GPUs with one all-reduce of the gradient image per iteration; "
              "that path is not built)", file=sys.stderr)
        sys.exit(2)

    if args.counter_pass:
        counter_pass(args)
        return
    if args.study:
        import tomo_amd.studies as S
        log = (lambda m: print("[study]", m, file=sys.stderr, flush=True))
        if args.study == "c2":
            out = S.c2_study(fp=args.fp or 0.1, log=log)
        elif args.study == "relax":
            out = S.relax_study(seed=int(args.seeds.split(",")[0]), log=log,
                                momenta=[float(b) for b in args.momentum.split(",")] if args.momentum else None)
        elif args.study == "train":
            out = S.train_study(args.out, phantoms=args.phantoms, train_steps=args.train_steps,
                                target=args.train_target, rounds=args.train_rounds, iters=args.train_iters,
                                stack=args.stack, first_seed=args.train_first_seed, base=args.weights,
                                momentum=float(args.momentum or 0.0), log=log)
        elif args.study == "finetune":
            out = S.finetune_study(args.out, base=args.weights, phantoms=args.phantoms,
                                   train_steps=args.train_steps, log=log)
        else:
            runs = []
            for v in args.seeds.split(","):
                runs.append(S.ttt_study(seed=int(v), weights=args.weights, learned_iters=args.learned_iters,
                                        momentum=float(args.momentum or 0.0), alpha=args.learned_alpha, log=log))
                # one line per seed as it finishes (the whole study runs for minutes)
                print(json.dumps({"ttt_seed": int(v), **{k: runs[-1][k] for k in (
                    "plain_never_below_100", "speedup_fixed50", "speedup_stop_rule", "speedup_fixed10")},
                                  **({"momentum_plain_speedup_fixed50":
                                      runs[-1]["targets"]["fixed50"]["momentum_plain_speedup"]}
                                     if runs[-1]["momentum"] else {})}),
                      file=sys.stderr, flush=True)
            out = ttt_summary(runs)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "ref":
        run_ref(args)
        return
    if args.workload == "c4":
        if args.c4_mode == "stacked":
            run_c4_stacked(args)
        else:
            run_c4(args)
        return
    c, ii, jj, fo = make_workload(args.workload)
    n, steps, shots, ckpt = WORKLOADS[args.workload]
    e = tomo.Engine(c, device=0)
    if args.shape:
        e.set_shape(*[int(v) for v in args.shape.split(",")])
    if not args.eager:
        e.set_graphs(True)
    e.set_time_blocking(args.time_blocking)
    e.set_adjoint_blocking(args.adjoint_blocking)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    e.set_truth(fo)
    e.set_field(fo)
    t = time.perf_counter()
    meas = e.forward(fp64=args.data == "fp64")
    t_data = time.perf_counter() - t
    # the boundary takes host arrays: the time to hand over this problem's
    # inputs (field and measured traces, host to HBM) is reported beside the
    # rate, which is measured with them resident
    meas = np.ascontiguousarray(meas, np.float32)
    f0 = np.zeros(fo.size, np.float32)
    t = time.perf_counter()
    e.set_field(f0)
    e.set_measured(meas)
    t_upload = time.perf_counter() - t
    upload_bytes = meas.nbytes + f0.nbytes
    print(f"[bench] workload {args.workload}: {n}x{n}, {shots} shots, {steps} steps; "
          f"{args.data.upper()} data in {t_data:.2f} s", file=sys.stderr, flush=True)

    if args.autotune:
        shape, _ = e.autotune()
        print(f"[bench] autotuned shape {shape}", file=sys.stderr, flush=True)
    # relaxation: a fixed scale that moves the field by at most 1e-3 in the first step
    e.gradient()
    g = np.stack([e.get_gradient(s) for s in range(shots)]).astype(np.float64).sum(0) / shots
    scale = float(1e-3 / max(np.abs(g).max(), 1e-30))

    for _ in range(args.warmup):
        e.sirt_iteration(scale, weighted=False, mode=tomo.UPDATE_MEAN)
    e.timing_reset()
    eps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, ep = e.sirt_iteration(scale, weighted=False, mode=tomo.UPDATE_MEAN)
        eps.append(ep)
        print(f"[bench] step {k}: eps {ep:.4f} %", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    tm = e.timing()
    tg = e.timing_graphs()
    if tg["untimed_cells"]:
        sys.exit("bench.py: the HIP runtime refused event-record nodes in the captured gradient graph, so its "
                 "stencil loops cannot be timed; run with --eager")
    cells = tm["fwd_cells"] + tm["adj_cells"]
    value = cells / elapsed / 1e9

    # a forward launch advances every shot of the pass by `tb` levels; the
    # engine's spans count levels and cover each launch with its edge passes
    tb, ab = e.time_blocking(), e.adjoint_blocking()
    shape = tuple(e.get_shape())
    fwd_cells_launch = shots * n * n * tb
    adj_cells_launch = shots * n * n * ab
    fwd_avg_ms = tm["fwd_ms"] / max(tm["fwd_launches"], 1) * tb
    adj_avg_ms = tm["adj_ms"] / max(tm["adj_launches"], 1) * ab
    fwd_kernel = {1: "forward_vec_kernel", 2: "forward2_kernel", 3: "forward3_kernel"}[tb]
    adj_kernel = {1: "adjoint_vec_kernel", 2: "adjoint2_kernel", 3: "adjoint3_kernel"}[ab]
    fwd_traffic = load_traffic(fwd_kernel, shape, tb) if args.workload == "c3" else None
    adj_traffic = load_traffic(adj_kernel, shape, ab) if args.workload == "c3" else None

    out = {
        "metric": "Gcell-updates/s of FDTD step",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded phantom; measured traces from the engine's {args.data.upper()} solver)",
        "config": {
            "workload": f"batched plain SIRT iteration: {n}x{n} grid, {shots} ring shots, {steps} steps, "
                        f"checkpoint every {ckpt or 'none'}",
            "grid": n, "shots": shots, "time_steps": steps, "checkpoint_every": ckpt,
            "cell_updates_per_step": cells / args.steps, "stencil_shape": list(shape),
            "forward_steps_per_launch": tb, "adjoint_steps_per_launch": ab,
            # each gradient pass (all 128 shots: fills, forward, misfit,
            # segment recomputes, adjoint) replayed as one captured HIP graph
            "launch": "eager" if args.eager else {"graph_replays": tg["replays"],
                                                  "graph_ms_per_step": round(tg["ms"] / args.steps, 3)},
            "eps_after_steps": eps[-1] if eps else None,
            # host-to-HBM hand-over of the inputs (once per problem), and the
            # rate of one iteration that includes it
            "input_upload": {"bytes": upload_bytes, "ms": round(t_upload * 1e3, 3),
                             "value_with_one_upload": round(cells / args.steps / (elapsed / args.steps + t_upload)
                                                            / 1e9, 3)},
        },
        "roofline": None,
        "forward": kernel_line(f"{fwd_kernel} (with its edge/receiver passes)", FWD_UNIT_BYTES, FWD_BYTES[tb],
                               fwd_cells_launch, fwd_avg_ms, tm["fwd_ms"], fwd_traffic,
                               "recording and replay launches as one gradient mixes them" if tb == 3
                               else "every launch"),
        "adjoint": kernel_line(f"{adj_kernel} (with its edge/receiver pass)", ADJ_UNIT_BYTES, ADJ_BYTES[ab],
                               adj_cells_launch, adj_avg_ms, tm["adj_ms"], adj_traffic, "every launch"),
        "cpu_baseline": None,
    }

    # the roofline line is the kernel with the largest share of the timed
    # region; with three steps per launch the forward's time splits over two
    # kernels of about equal share (recording and replay launches, 1333 and
    # 1320 per gradient at 4000 steps), the adjoint's is one kernel
    fwd_kernels = 2 if tb == 3 else 1
    dom = "adjoint" if tm["adj_ms"] > tm["fwd_ms"] / fwd_kernels else "forward"
    out["roofline"] = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", **out[dom]}
    if not args.no_cpu_baseline:
        import oracle as orc
        orc.use_library(native_oracle())
        # every OpenMP thread of this process (OMP_NUM_THREADS, or the CPUs
        # of its affinity), one shot each
        threads = args.cpu_threads or orc.num_threads()
        sample = min(threads, shots)
        t_cpu = cpu_baseline(c, ii, jj, fo, meas, threads, sample)
        cpu_cells = cells / args.steps / shots * sample
        out["cpu_baseline"] = {
            "value": round(cpu_cells / t_cpu / 1e9, 4), "unit": "Gcell-updates/s", "cores": orc.num_threads(),
            "kind": "port",
            "build": "oracle/tomo_oracle.cpp, g++ -O3 -march=native -ffp-contract=fast -fopenmp, built on this host",
            "sample": f"oracle FP32 gradient (forward + checkpoint recompute + adjoint) of {sample} of the "
                      f"{shots} shots, {n}x{n}, {steps} steps: {t_cpu:.2f} s (the whole iteration, "
                      f"{shots / sample:g}x the shots, would take {t_cpu * shots / sample:.0f} s at this rate)",
            "cpu_model": cpu_model(),
        }
    e.close()
    if args.copy_calibration:
        print("[bench] copy calibration: 1 GiB device copy", file=sys.stderr, flush=True)
        device_copy(1 << 30)
    print(json.dumps(out), flush=True)


def run_c4(args):
    """Uneven set: 64 phantoms of sides {128..512}, n/4 ring shots, 8n steps,
    stencil shapes from the autotuner per side; a step is one weighted SIRT
    iteration of every phantom.  The phantoms are dealt by cost (n^4) to
    --c4-threads host threads, each stepping its contexts one after the
    other on their own streams, so launches of different phantoms overlap
    on the GPU.  With --c4-passes 0 (the default) every phantom runs all its
    shots in one pass, the phantoms of a thread sharing one wavefield
    workspace.  (tomo_batch refuses a program of this size: about 260k
    launches per gradient.)"""
    from concurrent.futures import ThreadPoolExecutor
    log = (lambda m: print("[bench]", m, file=sys.stderr, flush=True))
    t = time.perf_counter()
    engines = W.c4_problems(log=log, passes=args.c4_passes)
    t_data = time.perf_counter() - t
    sides = W.c4_sides()

    # deal the phantoms to the threads, largest first, each to the least loaded
    nthreads = max(1, args.c4_threads or 4)
    groups = [[] for _ in range(nthreads)]
    load = [0.0] * nthreads
    for i in sorted(range(len(engines)), key=lambda i: -float(sides[i]) ** 4):
        t_min = min(range(nthreads), key=lambda t: load[t])
        groups[t_min].append(i)
        load[t_min] += float(sides[i]) ** 4
    # each thread steps its phantoms one after another: with all shots in one
    # pass they share one wavefield workspace, sized for the largest of them
    workspaces = []
    if args.c4_passes == 0:
        for g in groups:
            ws = tomo.Workspace(0)
            for i in g:
                ws.attach(engines[i])
            workspaces.append(ws)
    shapes = {}
    for e, n in zip(engines, sides):
        shapes[n] = e.autotune()[0]
    log(f"data for {len(engines)} phantoms in {t_data:.1f} s; shapes per side {shapes}")
    scales = []
    for e, n in zip(engines, sides):
        e.gradient()
        g = np.stack([e.get_gradient(s) for s in range(e.num_shots)]).astype(np.float64).sum(0) / e.num_shots
        scales.append(float(1e-3 / max(np.abs(g).max(), 1e-30)))

    def step_group(idx):
        for i in idx:
            engines[i].gradient()
            engines[i].update_weighted(scales[i])

    pool = ThreadPoolExecutor(max_workers=nthreads)

    def iteration():
        for f in [pool.submit(step_group, g) for g in groups if g]:
            f.result()

    for _ in range(args.warmup):
        iteration()
    for e in engines:
        e.timing_reset()
    t0 = time.perf_counter()
    for k in range(args.steps):
        iteration()
        log(f"step {k} done")
    elapsed = time.perf_counter() - t0
    pool.shutdown()
    cells = 0.0
    fwd_ms = adj_ms = 0.0
    for e in engines:
        tm = e.timing()
        cells += tm["fwd_cells"] + tm["adj_cells"]
        fwd_ms += tm["fwd_ms"]
        adj_ms += tm["adj_ms"]
    eps = float(np.mean([e.error() for e in engines]))
    out = {
        "metric": "Gcell-updates/s of FDTD step",
        "value": round(cells / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded phantoms; measured traces from the engine's FP64 solver)",
        "config": {
            "workload": "uneven set: 64 phantoms, sides from {128,192,256,384,512}, n/4 ring shots, 8n steps, "
                        "checkpoint every 128, "
                        + (f"passes of {args.c4_passes} shots" if args.c4_passes else
                           "all shots in one pass (one wavefield workspace per host thread)")
                        + f", one weighted SIRT iteration of every phantom per step, phantoms stepped by {nthreads} "
                        "host threads on their own streams",
            "shots_per_pass": args.c4_passes,
            "host_threads": nthreads,
            "sides": sides, "shapes": {str(k): list(v) for k, v in shapes.items()},
            "cell_updates_per_step": cells / args.steps, "mean_eps_after_steps": eps,
            # stencil spans summed over the phantoms (they overlap when threads > 1)
            "stencil_ms_share": round((fwd_ms + adj_ms) / (elapsed * 1e3), 4),
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    for e in engines:
        e.close()
    for ws in workspaces:
        ws.close()
    print(json.dumps(out), flush=True)


# The reference's only published timings (report.tex:187-199, Table 2): average
# SIRT execution time on a Tesla K20c, M = 64 and 32 sensor groups
REF_TABLE2_S = {64: 0.7625, 32: 0.23898}


def run_ref(args):
    """The reference's own case: its preset (sirt-mimo.cu:17-35: 192^2
    grid, 401 steps, sensor lines, one-way edges, the reference pulse), its
    640-transducer square ring in 64 groups of 10 (M = 64 of report.tex
    Table 2) and its phantom, full wavefield history as the reference keeps
    it; a step is one uniform SIRT iteration (sirt-mimo-io.cu:326-394).  The
    report does not say whether Table 2's times are per iteration or per run;
    the line gives the ratio per iteration."""
    c = tomo.config_reference()
    c.checkpoint_every = 0
    e = tomo.Engine(c, device=0)
    if not args.eager:
        e.set_graphs(True)
    ii, jj = tomo.transducers_reference(640)
    e.set_transducers(ii, jj, one_based=True)
    groups = e.set_groups(10)
    fo = tomo.phantom_reference()
    e.set_truth(fo)
    e.set_field(fo)
    meas = e.forward(fp64=True)
    e.set_field(np.zeros(fo.size, np.float32))
    e.set_measured(meas)
    e.gradient()
    g = np.stack([e.get_gradient(s) for s in range(groups)]).astype(np.float64).sum(0)
    scale = float(1e-3 / max(np.abs(g).max(), 1e-30))
    for _ in range(args.warmup):
        e.sirt_iteration(scale, weighted=False, mode=tomo.UPDATE_SUM)
    e.timing_reset()
    t0 = time.perf_counter()
    eps = None
    for _ in range(args.steps):
        _, eps = e.sirt_iteration(scale, weighted=False, mode=tomo.UPDATE_SUM)
    elapsed = time.perf_counter() - t0
    tm = e.timing()
    cells = tm["fwd_cells"] + tm["adj_cells"]
    e.close()
    sec = elapsed / args.steps
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle's FP32 gradient of every group on this host's threads
        # (an iteration less its update)
        import oracle as orc
        orc.use_library(native_oracle())
        orc.set_threads(args.cpu_threads or orc.num_threads())
        rc, ra = orc.receivers_lines(c)
        bbox = orc.groups(ii, jj, 10)
        m32 = np.ascontiguousarray(meas, np.float32)
        t = time.perf_counter()
        orc.gradient(c, np.zeros(fo.size, np.float32), bbox, rc, ra, m32, np.float32)
        t_cpu = time.perf_counter() - t
        cpu = {"value": round(cells / args.steps / t_cpu / 1e9, 4), "unit": "Gcell-updates/s",
               "cores": orc.num_threads(), "kind": "port",
               "build": "oracle/tomo_oracle.cpp, g++ -O3 -march=native -ffp-contract=fast -fopenmp, built on this host",
               "sample": f"oracle FP32 gradient of all {groups} groups, {c.nx}x{c.ny}, {c.nt} steps, full history: "
                         f"{t_cpu:.3f} s (the reference's Table 2 CPU time at M = 64: 46.87 s on 8 Xeon cores)",
               "cpu_model": cpu_model()}
    out = {
        "metric": "Gcell-updates/s of FDTD step",
        "value": round(cells / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "the reference's phantom; measured traces from the engine's FP64 solver",
        "config": {
            "workload": f"reference preset: {c.nx}x{c.ny} grid, {c.nt} steps, {groups} groups of 10 of the 640 "
                        "reference transducers, sensor lines, full wavefield history, one uniform SIRT iteration "
                        "per step",
            "launch": "eager" if args.eager else "graph replay of every gradient pass",
            "cell_updates_per_step": cells / args.steps, "eps_after_steps": eps,
            # Table 2 of the reference's report: K20c seconds per SIRT run or
            # iteration (not stated) at M = 64; ratio to this engine's iteration
            "reference_table2": {"k20c_avg_sirt_seconds": REF_TABLE2_S[groups] if groups in REF_TABLE2_S else None,
                                 "ratio_to_one_iteration": round(REF_TABLE2_S[groups] / sec, 1)
                                 if groups in REF_TABLE2_S else None},
        },
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_c4_stacked(args):
    """Uneven set, stacked: one engine per side (128, 192, 256, 384, 512)
    holding that side's phantoms along the shot axis, so each stencil launch
    advances every shot of every phantom of the side (in passes that keep
    the wavefield storage within 160 GB).  A step is one weighted SIRT
    iteration of every phantom, each phantom weighted over its own shots at
    its own relaxation.
    The engines are dealt by cost (n^4 x phantoms) to --c4-threads host
    threads (default 2), each stepping its engines one after the other on
    their own streams with one shared wavefield workspace; each gradient
    pass is replayed as a captured HIP graph unless --eager."""
    from concurrent.futures import ThreadPoolExecutor
    log = (lambda m: print("[bench]", m, file=sys.stderr, flush=True))
    t = time.perf_counter()
    stacks = W.c4_stacks(log=log)
    t_data = time.perf_counter() - t
    nthreads = max(1, args.c4_threads or 2)
    groups = [[] for _ in range(nthreads)]
    load = [0.0] * nthreads
    for i in sorted(range(len(stacks)), key=lambda i: -float(stacks[i][0]) ** 4 * len(stacks[i][2])):
        t_min = min(range(nthreads), key=lambda t: load[t])
        groups[t_min].append(i)
        load[t_min] += float(stacks[i][0]) ** 4 * len(stacks[i][2])
    workspaces = []
    for g in groups:
        ws = tomo.Workspace(0)
        for i in g:
            ws.attach(stacks[i][1])
        workspaces.append(ws)
    shapes = {}
    for n, e, _ in stacks:
        shapes[n] = e.autotune()[0]
        if not args.eager:
            e.set_graphs(True)
    log(f"data for {sum(len(x[2]) for x in stacks)} phantoms in {t_data:.1f} s; shapes per side {shapes}")
    # each phantom's own relaxation, as in a run of its own: 1e-3 over its
    # largest mean-gradient entry (tomo_update_weighted_each)
    scales = []
    for n, e, idx in stacks:
        e.gradient()
        S = e.num_shots // len(idx)
        sc = np.zeros(len(idx), np.float32)
        for p in range(len(idx)):
            g = np.zeros(n * n, np.float64)
            for s_ in range(p * S, (p + 1) * S):
                g += e.get_gradient(s_)
            sc[p] = 1e-3 / max(float(np.abs(g / S).max()), 1e-30)
        scales.append(sc)

    def step_group(idx):
        for i in idx:
            stacks[i][1].gradient()
            stacks[i][1].update_weighted(scales[i])

    pool = ThreadPoolExecutor(max_workers=nthreads)

    def iteration():
        for f in [pool.submit(step_group, g) for g in groups if g]:
            f.result()

    for _ in range(args.warmup):
        iteration()
    for _, e, _ in stacks:
        e.timing_reset()
    t0 = time.perf_counter()
    for k in range(args.steps):
        iteration()
        log(f"step {k} done")
    elapsed = time.perf_counter() - t0
    pool.shutdown()
    cells = 0.0
    per_side = {}
    for n, e, idx in stacks:
        tm = e.timing()
        c_ = tm["fwd_cells"] + tm["adj_cells"]
        if e.timing_graphs()["untimed_cells"]:
            sys.exit("bench.py: the HIP runtime refused event-record nodes in the captured gradient graph; "
                     "run with --eager")
        cells += c_
        per_side[str(n)] = {"phantoms": len(idx), "shots": e.num_shots, "shots_per_pass": e.cfg.shots_per_pass
                            or e.num_shots, "stencil_ms_per_step": round((tm["fwd_ms"] + tm["adj_ms"]) / args.steps, 1),
                            "cell_updates_per_step": c_ / args.steps}
    eps = float(np.mean(np.concatenate([np.atleast_1d(e.error()) for _, e, _ in stacks])))
    out = {
        "metric": "Gcell-updates/s of FDTD step",
        "value": round(cells / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded phantoms; measured traces from the engine's FP64 solver)",
        "config": {
            "workload": "uneven set: 64 phantoms, sides from {128,192,256,384,512}, n/4 ring shots, 8n steps, "
                        "checkpoint every 128; one engine per side with its phantoms stacked along the shot axis "
                        f"(one launch per level covers all of them), one weighted SIRT iteration of every phantom "
                        f"per step, engines stepped by {nthreads} host threads on their own streams",
            "launch": "eager" if args.eager else "graph replay of every gradient pass",
            "host_threads": nthreads, "sides": per_side, "shapes": {str(k): list(v) for k, v in shapes.items()},
            "cell_updates_per_step": cells / args.steps, "mean_eps_after_steps": eps,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    for _, e, _ in stacks:
        e.close()
    for ws in workspaces:
        ws.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
