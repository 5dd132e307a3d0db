#!/usr/bin/env python3
"""Benchmark: RHS cell-updates/s of the LSRK3 step at 512^3 FP32 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--mode parity|fast]

Workload: the bubble-cloud block (50 air bubbles, lognormal radii, in water at
p = 100; transmissive boundaries), binary32 SoA resident in HBM, full LSRK3
steps at CFL 0.3 (max speed of sound -> dt on the device -> 3 stages of
ghost fill + x/y/z sweeps + register update).  Inputs are resident before
the timed region; nothing crosses PCIe inside it.

value = N^3 cells x 3 stages x K steps / wall time of the K steps
(RHS cell-updates/s, whole job), for the FMA-contracted build (default) or the
bit-exact build (--mode parity); the other build is timed too and reported
under "other_build".  Per-kernel durations come from HIP events
recorded around every launch on the context's stream (mpcf_timing_*); the
roofline line is for the dominant kernel (the most expensive direction sweep)
with FLOPs from the library's own operation count of the scheme.

cpu_baseline: the CPU oracle (the restated reference scheme in the
reference's binary32 arithmetic, OpenMP over z-slices) timed on this host on
bounded samples of the bubble-cloud workload: 64^3 for 10 LSRK3 steps and
256^3 for one step, each the median of 3 repetitions after a warm-up step,
and the max-speed-of-sound reduction alone at 256^3; `value` is the 256^3
rate.  --cpu-full runs the whole plan instead (64^3 x 10, 256^3 x 5 and 512^3
x 2 steps, medians of 3).  The host's CPU model, nproc, the physical cores of
the GPU's NUMA node and the threads used are recorded with it.

ensemble / serving (the metric's second half: p50/p99 against the plain
solver): 64 uneven bubble-collapse cases batched vs one at a time, and
Poisson request streams served by the library's continuous-batching server
vs first-come first-served one at a time (cubism-jr-gpu-mpcf_amd/ensemble.py).
precision: the binary32 builds against a binary64 run of the same scheme
(cubism-jr-gpu-mpcf_amd/precision.py) on a smaller bubble cloud.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cubism-jr-gpu-mpcf_amd"))

PEAK_FP32_TFLOPS = 157.3      # MI355X dense FP32 vector, published
PEAK_HBM_GBS = 8000.0         # MI355X HBM3E, published
BYTES_PER_CELL_STAGE = 112.0  # 7 fields x 4 B x (read u, read R, write R, write u)
# algorithmic HBM bytes per cell of each sweep kernel as built here
KERNEL_BYTES_PER_CELL = {
    # bit-exact build: the x+y sweep(s) leave 10 accumulators, the z sweep does
    # the reference's single update of R and u
    "parity": {
        "k_sweep_xy": 28 + 40,                    # read u, write 10 accumulators
        "k_sweep_x": 28 + 40,                     # read u, write 10 accumulators
        "k_sweep_yz<1>": 28 + 40 + 40,            # read u, read + write accumulators
        "k_sweep_yz<2>": 28 + 40 + 28 + 28 + 28,  # read u, accumulators, R; write R, u'
    },
    # FAST build: the x+y part is folded into R where it is computed; 3 G/P
    # accumulators remain for the z sweep
    "fast": {
        # read u, R (not in a step's first stage: a = 0); write R, 3 accumulators;
        # the mean over the three stages
        "k_sweep_xy": 28 + 28 * 2 / 3 + 28 + 12,
        "k_sweep_x": 28 + 40,                     # read u, write 10 accumulators
        "k_sweep_yz<1>": 28 + 40 + 28 + 28 + 12,  # read u, accumulators, R; write R, 3 accumulators
        # read u, 3 accumulators, R; write R (not in the last stage of a step
        # that another step of the same call follows), u'
        "k_sweep_yz<2>": 28 + 12 + 28 + 28 * 2 / 3 + 28,
    },
}
# measured HBM bytes per cell of each kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE
# at 512^3, calibrated on the pack copy kernel), written by the profiling run
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_512.json")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=512)
    p.add_argument("--mode", choices=("parity", "fast"), default="fast")
    p.add_argument("--no-pipeline", action="store_true",
                   help="whole-block sweeps in every stage (MPCF_FLAG_NO_PIPELINE)")
    p.add_argument("--no-parity-line", action="store_true",
                   help="skip the extra timing of the bit-exact build reported under 'parity_build'")
    p.add_argument("--cfl", type=float, default=0.3)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-full", action="store_true",
                   help="CPU baseline over the whole plan: 64^3 x 10, 256^3 x 5 and 512^3 x 2 steps (minutes)")
    p.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble comparison")
    p.add_argument("--ensemble-cases", type=int, default=64)
    p.add_argument("--ensemble-steps", type=int, default=100)
    p.add_argument("--no-serving", action="store_true", help="skip the served-stream comparison")
    p.add_argument("--serving-requests", type=int, default=48)
    p.add_argument("--serving-policy", default="fifo_slack", choices=("all", "fifo_fill", "srpt_fill", "hybrid", "fifo_slack"))
    p.add_argument("--serving-fill", type=int, default=16 << 20, help="active-subset cell budget (floor) of the server's policy")
    p.add_argument("--no-precision", action="store_true", help="skip the binary32-vs-binary64 leg")
    p.add_argument("--precision-n", type=int, default=128)
    p.add_argument("--precision-steps", type=int, default=50)
    p.add_argument("--serving-load", type=float, nargs="+", default=[0.7, 1.0, 1.2],
                   help="offered load as a fraction of the plain solver's capacity on the request list")
    return p.parse_args()


def host_info():
    """CPU model, logical CPUs, and the physical cores of the NUMA node the
    GPU is attached to (sysfs; None where the host does not say)."""
    info = {"cpu_model": None, "nproc": os.cpu_count(), "gpu_numa_node": None, "socket_physical_cores": None,
            "node_physical_cores": None}
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass

    def cores_of(cpus):
        ids = set()
        for c in cpus:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            try:
                ids.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
            except OSError:
                return None
        return ids

    def cpulist(text):
        out = []
        for part in text.strip().split(","):
            if "-" in part:
                a_, b_ = part.split("-")
                out.extend(range(int(a_), int(b_) + 1))
            elif part:
                out.append(int(part))
        return out

    try:
        import glob
        for dev in sorted(glob.glob("/sys/class/drm/card*/device")):
            vendor = open(os.path.join(dev, "vendor")).read().strip()
            node = int(open(os.path.join(dev, "numa_node")).read().strip())
            if vendor == "0x1002" and node >= 0:
                info["gpu_numa_node"] = node
                cores = cores_of(cpulist(open(f"/sys/devices/system/node/node{node}/cpulist").read()))
                if cores is not None:
                    info["node_physical_cores"] = len(cores)
                    pkgs = {p for p, _ in cores}
                    allc = cores_of(range(os.cpu_count() or 0))
                    if allc is not None:
                        info["socket_physical_cores"] = len({c for c in allc if c[0] in pkgs})
                break
    except (OSError, ValueError):
        pass
    return info


def cpu_baseline(full: bool):
    """The restated reference scheme (the CPU oracle: binary32, the
    reference's operation order, OpenMP over z-slices) on this host's cores.
    Per size: one warm-up step, then the median of 3 timed repetitions of
    `steps` LSRK3 steps (3 stages + the max speed of sound each); the
    speed-of-sound reduction is also timed alone.  Rates are cell-updates/s
    (cells x stages / s), the unit of the GPU figure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import statistics
    import pyoracle
    import pympcf
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pyoracle.set_threads(nthreads)
    plan = [(64, 10), (256, 5), (512, 2)] if full else [(64, 10), (256, 1)]
    sizes = {}
    sos = None
    for n, steps in plan:
        f, _ = pympcf.ic_bubble_cloud(n, n, n, 50, 0)
        b = pyoracle.Block(n, n, n)
        b.data = f
        del f
        h = 1.0 / n
        t0 = time.perf_counter()
        b.step(h, 0.3, 1e30, [0] * 6)          # warm-up step
        log(f"CPU baseline {n}^3: warm-up step {time.perf_counter() - t0:.2f} s")
        reps = []
        for i in range(3):
            t0 = time.perf_counter()
            for _ in range(steps):
                b.step(h, 0.3, 1e30, [0] * 6)
            reps.append(time.perf_counter() - t0)
            log(f"CPU baseline {n}^3: repetition {i + 1} of 3, {steps} steps in {reps[-1]:.2f} s")
        med = statistics.median(reps)
        sizes[str(n)] = {"steps": steps, "repetitions_s": reps, "median_s_per_step": med / steps,
                         "cell_updates_per_s": n ** 3 * 3 * steps / med}
        log(f"CPU baseline {n}^3 x {steps} steps: median {med / steps:.3f} s/step "
            f"({n ** 3 * 3 * steps / med:.3e} cell-updates/s)")
        if n == 256:
            t = []
            for _ in range(3):
                t0 = time.perf_counter()
                b.max_sos()
                t.append(time.perf_counter() - t0)
            sos = {"n": n, "median_s": statistics.median(t), "cells_per_s": n ** 3 / statistics.median(t)}
        del b
    top = max(sizes, key=int)
    return {"value": sizes[top]["cell_updates_per_s"], "unit": "cell-updates/s", "cores": pyoracle.threads(),
            "kind": "port",
            "sample": (f"the oracle (restated reference scheme, binary32 in the reference's operation order, "
                       f"OpenMP over z-slices) on bubble-cloud blocks: "
                       + ", ".join(f"{n}^3 x {v['steps']} LSRK3 steps" for n, v in sizes.items())
                       + f", each the median of 3 repetitions after a warm-up step; value = the {top}^3 rate"),
            "sizes": sizes, "max_sos": sos, "threads": pyoracle.threads(), **host_info()}


def ensemble_leg(a):
    """64 bubble-collapse cases (edges {32..128}), 100 steps each: batched
    (all cases together, one launch per kernel kind, graph-captured steps,
    autotuned march length), the same batch kernels through the serving
    engine with least-remaining-work-first ordering, and the plain
    one-case-at-a-time solver."""
    import ensemble as E
    cases = E.ensemble_cases(a.ensemble_cases, a.ensemble_steps, seed=0)
    work = sum(c.work for c in cases)
    tb, mb, _ = E.run_batched(cases, cfl=a.cfl, fast=True, graph=True, autotune=True)
    log(f"ensemble batched: {mb:.2f} s")
    ts, ms, _ = E.run_served(cases, cfl=a.cfl, fast=True, policy="srpt_fill", fill_cells=16 << 20)
    log(f"ensemble served (least remaining work first): {ms:.2f} s")
    tp, mp, _ = E.run_plain(cases, cfl=a.cfl, fast=True)
    log(f"ensemble plain: {mp:.2f} s")
    return {
        "workload": f"{len(cases)} bubble-collapse cases, edges in {{32,48,64,96,128}}, {a.ensemble_steps} LSRK3 steps each, FP32",
        "cell_updates": work,
        "batched": dict(cell_updates_per_s=work / mb, makespan_s=mb, **E.latency_stats(list(tb.values()))),
        "served_srpt": dict(cell_updates_per_s=work / ms, makespan_s=ms, **E.latency_stats(list(ts.values()))),
        "plain": dict(cell_updates_per_s=work / mp, makespan_s=mp, **E.latency_stats(list(tp.values()))),
    }


def serving_leg(a):
    """Poisson request streams (edges 32..192 step 16, 50..500 steps) offered
    at fixed fractions of the plain solver's capacity, served by the library's
    continuous-batching server (mpcf_server) and by the plain sequential
    solver; time-to-solution = completion - arrival."""
    import ensemble as E
    # the plain solver's capacity on this request list: its requests run back
    # to back (request sizes do not depend on the arrival rate; only the
    # arrival times scale with it)
    base = E.stream_cases(a.serving_requests, seed=1, rate=1.0)
    _, mk, _ = E.run_plain(base, cfl=a.cfl, fast=True)
    plain_rate = len(base) / mk
    log(f"plain solver capacity on the request list: {plain_rate:.2f} requests/s")
    out = {"workload": f"{a.serving_requests} requests per load, edges 32..192 step 16, 50..500 steps, "
                       f"single-bubble collapse, FP32; Poisson arrivals",
           "plain_capacity_per_s": plain_rate, "policy": a.serving_policy, "fill_cells": a.serving_fill,
           "loads": []}
    for load in a.serving_load:
        rate = load * plain_rate
        reqs = E.stream_cases(a.serving_requests, seed=1, rate=rate)
        work = sum(c.work for c in reqs)
        ts, ms, stats, _ = E.serve_native(reqs, cfl=a.cfl, fast=True, policy=a.serving_policy,
                                          fill_cells=a.serving_fill)
        log(f"serving load {load:.2f}: continuous batching {ms:.2f} s")
        tp, mp = E.serve_plain(reqs, cfl=a.cfl, fast=True)
        log(f"serving load {load:.2f}: plain {mp:.2f} s")
        out["loads"].append({
            "offered_load": load, "rate_per_s": rate,
            "continuous_batching": dict(cell_updates_per_s=work / ms, makespan_s=ms,
                                        **E.latency_stats(list(ts.values())),
                                        **{k: stats[k] for k in ("steps", "captures", "rebuilds", "peak_active")}),
            "plain_fifo": dict(cell_updates_per_s=work / mp, makespan_s=mp, **E.latency_stats(list(tp.values()))),
        })
    return out


def precision_leg(a):
    """binary32 builds against a binary64 run of the same scheme on a smaller
    bubble cloud (the 512^3 x 200-step report is a separate run of
    cubism-jr-gpu-mpcf_amd/precision.py)."""
    import precision
    return precision.run(n=a.precision_n, steps=a.precision_steps, nbubbles=50, seed=0, cfl=a.cfl,
                         log=lambda *m: log("precision:", *m))


def log(*msg):
    print("[bench]", *msg, file=sys.stderr, flush=True)


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU path (halo exchange over RCCL) is not part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pympcf

    n = a.n
    fast = a.mode == "fast"
    t_ic = time.perf_counter()
    f, bubbles = pympcf.ic_bubble_cloud(n, n, n, 50, 0)
    t_ic = time.perf_counter() - t_ic
    log(f"initial condition {n}^3 built in {t_ic:.1f} s")

    def timed(fast_build, steps, warmup):
        blk = pympcf.Block(n, n, n, 1.0 / n, fast=fast_build, bc=[pympcf.BC_ABSORBING] * 6,
                           pipeline=not a.no_pipeline)
        blk.upload(pympcf.DATA, f)
        blk.upload(pympcf.TMP, np.zeros_like(f))
        blk.steps(warmup, a.cfl)
        blk.synchronize()
        log(f"{warmup} warm-up steps done; timing {steps} steps ({'fast' if fast_build else 'parity'} build)")
        blk.timing(True)
        blk.timing_read()                          # clear
        blk.synchronize()
        t0 = time.perf_counter()
        blk.steps(steps, a.cfl)              # K LSRK3 steps, queued at once
        blk.synchronize()
        wall = time.perf_counter() - t0
        ktimes = blk.timing_read()
        blk.timing(False)
        dt, sos = blk.last_dt()
        fused = blk.fused_xy
        blk.close()
        if not (np.isfinite(dt) and dt > 0):
            raise RuntimeError(f"non-finite dt after the timed steps: {dt}")
        return wall, ktimes, dt, sos, fused

    wall, ktimes, dt, sos, fused = timed(fast, a.steps, a.warmup)
    other = None
    if not a.no_parity_line:
        w2, k2, _, _, _ = timed(not fast, a.steps, 1)
        other = {"build": "fast" if not fast else "parity", "value": n ** 3 * 3 * a.steps / w2,
                 "ms_per_step": w2 / a.steps * 1e3,
                 "kernels": {k: {"avg_ms": v[0] / v[1]} for k, v in k2.items() if v[1]}}
    del f

    cells = n ** 3
    value = cells * 3 * a.steps / wall
    # algorithmic FLOPs: the operation count of the reference's arithmetic
    # (the bit-exact build); the FAST build executes fewer for the same work
    fl = pympcf.flop_count(False)
    fl_run = pympcf.flop_count(fast)
    # dominant kernel: the sweep kernel with the largest total time; the fused
    # x+y sweep is timed as "sweep_x" and does the x and the y work
    kflops = {"sweep_x": fl["sweep_x"] + (fl["sweep_y"] if fused else 0), "sweep_y": fl["sweep_y"],
              "sweep_z": fl["sweep_z"]}
    kflops_run = {"sweep_x": fl_run["sweep_x"] + (fl_run["sweep_y"] if fused else 0), "sweep_y": fl_run["sweep_y"],
                  "sweep_z": fl_run["sweep_z"]}
    kname = {"sweep_x": "k_sweep_xy" if fused else "k_sweep_x", "sweep_y": "k_sweep_yz<1>",
             "sweep_z": "k_sweep_yz<2>"}
    sweeps = {k: ktimes[k] for k in ("sweep_x", "sweep_y", "sweep_z") if ktimes[k][1]}
    dom = max(sweeps, key=lambda k: sweeps[k][0])
    dom_ms, dom_cnt = sweeps[dom]
    # a pipelined step launches each sweep once per z-slab: a launch covers
    # cells / (launches per stage)
    per_stage = dom_cnt / (3 * a.steps)
    cells_launch = cells / per_stage
    dom_avg_s = dom_ms / dom_cnt / 1e3
    dom_flops = kflops[dom] * cells_launch
    achieved = dom_flops / dom_avg_s / 1e12
    # per stage from the wall time (pipelined sweeps overlap)
    stage_s = wall / (3 * a.steps)
    hbm_stage = cells * BYTES_PER_CELL_STAGE / stage_s / 1e9

    kernels = {k: {"avg_ms": (v[0] / v[1] if v[1] else 0.0), "launches": v[1]} for k, v in ktimes.items() if v[1]}
    # measured bytes per cell come from the committed rocprofv3 summary of the
    # 512^3 workload (counters cannot be read inside this run); scaled to the
    # cells of one launch, and only for that size
    traffic = None
    traffic_source = None
    if n == 512 and os.path.exists(TRAFFIC_FILE):
        tr = json.load(open(TRAFFIC_FILE)).get(a.mode, {}).get(kname[dom])
        if tr is not None:
            traffic = tr * cells_launch
            traffic_source = ("profiles/traffic_512.json: rocprofv3 FETCH_SIZE + WRITE_SIZE per cell of this "
                              "kernel at 512^3, calibrated on a copy kernel, times the cells of one launch")
    alg_bytes = KERNEL_BYTES_PER_CELL[a.mode][kname[dom]] * cells_launch
    # the FAST z sweep moves its design bytes at the copy bandwidth: when it is
    # the longer sweep (768^3 and up) it is priced against HBM, every other
    # sweep against the VALU (all sweeps of the bit-exact build are VALU-bound)
    hbm_bound = fast and dom == "sweep_z"
    if hbm_bound:
        achieved, peak, unit = alg_bytes / dom_avg_s / 1e9, PEAK_HBM_GBS, "GB/s"
    else:
        peak, unit = PEAK_FP32_TFLOPS, "TFLOP/s"
    out = {
        "metric": ("RHS cell-updates/s at 512^3 FP32; ensemble p50/p99 latency vs plain solver" if n == 512
                   else f"RHS cell-updates/s at {n}^3 FP32 (the metric is quoted at 512^3)"),
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"bubble-cloud {n}^3 (50 air bubbles in water, p 100/1), LSRK3 CFL {a.cfl}, "
                        f"transmissive BCs, binary32 SoA resident in HBM",
            "build": ("FMA-contracted binary32 with hardware rcp/sqrt; within the bit-exact scheme's own "
                      "1-ulp noise floor (tests/test_gpu_fast.py)") if fast else "bit-exact reference arithmetic",
            "n": n, "mode": a.mode, "xy": "fused" if fused else "split", "parallelism": "single GPU",
            "pipeline": (f"{per_stage:g} z-slabs per stage, x+y and z sweeps of consecutive slabs on two streams"
                         if per_stage > 1 else "off"),
        },
        "roofline": {
            "bound": "hbm" if hbm_bound else "valu",
            "kernel": kname[dom],
            "achieved": achieved,
            "peak": peak,
            "unit": unit,
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "algorithmic_bytes": alg_bytes,
            "kernel_hbm_GBs": alg_bytes / dom_avg_s / 1e9,
            "flop_per_cell": kflops[dom],
            "cells_per_launch": cells_launch,
            "flop_per_cell_executed": kflops_run[dom],
            "avg_launch_ms": dom_ms / dom_cnt,
            "hbm_stage_GBs": hbm_stage,
            "hbm_stage_frac": hbm_stage / PEAK_HBM_GBS,
            # the published FP32 peak counts packed (two-wide) FMAs; every build
            # here is compiled without packed FP32, so one FMA per lane per
            # cycle, half of it, is the issue ceiling these kernels can reach
            "peak_unpacked_fp32": PEAK_FP32_TFLOPS / 2 if unit == "TFLOP/s" else None,
            "frac_unpacked_fp32": achieved / (PEAK_FP32_TFLOPS / 2) if unit == "TFLOP/s" else None,
        },
        "kernels": kernels,
        "flop_per_cell_stage": fl["cell_stage"],
        "flop_per_cell_stage_executed": fl_run["cell_stage"],
        "other_build": other,
        "ic_seconds": t_ic,
        "last_dt": dt,
        "last_sos": sos,
    }
    log(f"timed region {wall:.3f} s; {value:.3e} cell-updates/s")
    if not a.no_ensemble:
        out["ensemble"] = ensemble_leg(a)
    if not a.no_serving:
        out["serving"] = serving_leg(a)
    if not a.no_precision:
        out["precision"] = precision_leg(a)
    if not a.no_cpu:
        log("CPU baseline: the oracle on this host" + (" (whole plan)" if a.cpu_full else ""))
        out["cpu_baseline"] = cpu_baseline(a.cpu_full)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
