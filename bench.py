#!/usr/bin/env python3
"""Benchmark: RHS cell-updates/s of the LSRK3 step at 512^3 FP32 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--mode parity|fast]

Workload: the bubble-cloud block (50 air bubbles, lognormal radii, in water at
p = 100; transmissive boundaries), binary32 SoA resident in HBM, full LSRK3
steps at CFL 0.3 (max speed of sound -> dt on the device -> 3 stages of
ghost fill + x/y/z sweeps + register update).  Inputs are resident before
the timed region; nothing crosses PCIe inside it.

value = N^3 cells x 3 stages x K steps / wall time of the K steps
(RHS cell-updates/s, whole job).  Per-kernel durations come from HIP events
recorded around every launch on the context's stream (mpcf_timing_*); the
roofline line is for the dominant kernel (the most expensive direction sweep)
with FLOPs from the library's own operation count of the scheme.

cpu_baseline: the CPU oracle (the restated reference scheme, OpenMP over
z-slices) timed on this host on a bounded sample (a 128^3 block of the same
workload, one LSRK3 step), scaled to the same unit.
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
BYTES_PER_CELL_DIR = {        # algorithmic HBM bytes per cell of each sweep as built here
    "sweep_x": 28 + 40,                 # read u, write 10 accumulators
    "sweep_y": 28 + 40 + 40,            # read u, read + write accumulators
    "sweep_z": 28 + 40 + 28 + 28 + 28,  # read u, accumulators, R; write R, u'
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=512)
    p.add_argument("--mode", choices=("parity", "fast"), default="parity")
    p.add_argument("--cfl", type=float, default=0.3)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-n", type=int, default=160)
    return p.parse_args()


def cpu_baseline(n: int):
    """The restated reference scheme (oracle) on this host's cores: one LSRK3
    step of an n^3 bubble-cloud block.  Returns cell-updates/s (cell x stage)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    import pympcf
    ncores = os.cpu_count() or 1
    nthreads = int(os.environ.get("OMP_NUM_THREADS", ncores))
    pyoracle.set_threads(nthreads)
    f, _ = pympcf.ic_bubble_cloud(n, n, n, 50, 0)
    b = pyoracle.Block(n, n, n)
    b.data = f
    h = 1.0 / n
    b.step(h, 0.3, 1e30, [0] * 6)          # warm-up step
    t0 = time.perf_counter()
    b.step(h, 0.3, 1e30, [0] * 6)
    el = time.perf_counter() - t0
    return {"value": n ** 3 * 3 / el, "unit": "cell-updates/s", "cores": pyoracle.threads(), "kind": "port",
            "sample": f"one LSRK3 step (3 stages + max-SOS) of a {n}^3 bubble-cloud block, "
                      f"oracle (restated reference scheme, binary32, OpenMP) after one warm-up step; "
                      f"{el:.2f} s"}


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
    log(f"initial condition {n}^3 built in {t_ic:.1f} s; uploading")
    blk = pympcf.Block(n, n, n, 1.0 / n, fast=fast, bc=[pympcf.BC_ABSORBING] * 6)
    blk.upload(pympcf.DATA, f)
    blk.upload(pympcf.TMP, np.zeros_like(f))
    del f

    for _ in range(a.warmup):
        blk.step(a.cfl, sync=False)
    blk.synchronize()
    log(f"{a.warmup} warm-up steps done; timing {a.steps} steps ({a.mode})")

    blk.timing(True)
    blk.timing_read()                          # clear
    blk.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        blk.step(a.cfl, sync=False)
    blk.synchronize()
    wall = time.perf_counter() - t0
    ktimes = blk.timing_read()
    blk.timing(False)
    dt, sos = blk.last_dt()
    if not (np.isfinite(dt) and dt > 0):
        raise RuntimeError(f"non-finite dt after the timed steps: {dt}")

    cells = n ** 3
    value = cells * 3 * a.steps / wall
    fl = pympcf.flop_count(fast)
    # dominant kernel: the direction sweep with the largest total time
    sweeps = {k: ktimes[k] for k in ("sweep_x", "sweep_y", "sweep_z")}
    dom = max(sweeps, key=lambda k: sweeps[k][0])
    dom_ms, dom_cnt = sweeps[dom]
    dom_avg_s = dom_ms / dom_cnt / 1e3
    dom_flops = fl[dom] * cells
    achieved = dom_flops / dom_avg_s / 1e12
    stage_ms = sum(v[0] for v in sweeps.values()) / (3 * a.steps)
    hbm_stage = cells * BYTES_PER_CELL_STAGE / (stage_ms / 1e3) / 1e9

    kernels = {k: {"avg_ms": (v[0] / v[1] if v[1] else 0.0), "launches": v[1]} for k, v in ktimes.items() if v[1]}
    out = {
        "metric": "RHS cell-updates/s at 512^3 FP32",
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
            "n": n, "mode": a.mode, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": dom,
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
            "flop_per_cell": fl[dom],
            "avg_launch_ms": dom_ms / dom_cnt,
            "hbm_stage_GBs": hbm_stage,
            "hbm_stage_frac": hbm_stage / PEAK_HBM_GBS,
        },
        "kernels": kernels,
        "flop_per_cell_stage": fl["cell_stage"],
        "ic_seconds": t_ic,
        "last_dt": dt,
        "last_sos": sos,
    }
    log(f"timed region {wall:.3f} s; {value:.3e} cell-updates/s")
    if not a.no_cpu:
        log(f"CPU baseline: oracle on a {a.cpu_n}^3 sample")
        out["cpu_baseline"] = cpu_baseline(a.cpu_n)
    blk.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
