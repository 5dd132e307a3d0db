"""Ensembles and serving over the batch C-ABI (mpcf_batch_*).

ensemble_cases(): the batched-ensemble workload -- `ncases` independent
    bubble-collapse cases, cubic edges drawn from {32,48,64,96,128}, one air
    bubble of radius U[0.1,0.3] with centre jitter +-0.05, water/bubble
    pressure ratio log-uniform in [10, 1e4], `steps` LSRK3 steps each.
run_batched() / run_plain(): advance a case list together (one launch per
    kernel kind per stage, optional hipGraph) or one case at a time with its
    own context (the plain solver); both report per-case time-to-solution.
serve(): continuous batching of a Poisson request stream driven from Python
    over the batch API -- requests are admitted while their HBM footprint
    fits the budget, join the live batch between steps, leave it when done.
serve_native(): the same stream through the library's serving engine
    (mpcf_server_*): submission returns at once, an engine thread admits,
    steps and retires requests without synchronizing the device.
Both report p50/p99 time-to-solution and aggregate cell-updates/s;
serve_plain() is the sequential one-case-at-a-time baseline.
"""
from __future__ import annotations

import heapq
import time
from dataclasses import dataclass, field

import numpy as np

import pympcf


@dataclass
class Case:
    n: int
    steps: int
    radius: float
    pratio: float
    centre: tuple
    arrival: float = 0.0
    ident: int = 0

    def fields(self):
        return pympcf.ic_bubble_collapse(self.n, self.n, self.n, self.radius, self.pratio, self.centre)

    @property
    def work(self):
        """cell-updates (cells x stages) of the whole case"""
        return self.n ** 3 * 3 * self.steps


def draw_case(rng, n, steps, ident=0, arrival=0.0):
    r = rng.uniform(0.1, 0.3)
    pr = float(np.exp(rng.uniform(np.log(10.0), np.log(1e4))))
    c = tuple(0.5 + rng.uniform(-0.05, 0.05, size=3))
    return Case(n=int(n), steps=int(steps), radius=r, pratio=pr, centre=c, arrival=arrival, ident=ident)


def ensemble_cases(ncases=64, steps=100, seed=0, edges=(32, 48, 64, 96, 128)):
    rng = np.random.default_rng(seed)
    return [draw_case(rng, rng.choice(edges), steps, ident=i) for i in range(ncases)]


def stream_cases(nreq=64, seed=0, rate=1.0, edges=tuple(range(32, 193, 16)), steps=(50, 500)):
    """Poisson arrivals at `rate` requests/s; edge uniform in `edges`, steps
    uniform in [steps[0], steps[1]]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate)
        out.append(draw_case(rng, rng.choice(edges), rng.integers(steps[0], steps[1] + 1), ident=i, arrival=t))
    return out


def run_batched(cases, cfl=0.3, fast=True, graph=True, autotune=False, keep_results=False):
    """All cases start together and step together; a case finishes after its
    own number of steps.  Returns per-case time-to-solution (s), the makespan,
    and optionally the final states."""
    results = {}
    fields = {c.ident: c.fields() for c in cases}
    with pympcf.Batch(fast=fast, graph=graph) as B:
        tune = None
        if autotune:
            ids = {c.ident: B.add(fields[c.ident]) for c in cases}
            tune = B.autotune(steps=1, cfl=cfl)
            for i in ids.values():
                B.remove(i)
        t0 = time.perf_counter()
        ids = {c.ident: B.add(fields[c.ident]) for c in cases}
        left = {c.ident: c.steps for c in cases}
        done = {}
        while left:
            B.step(cfl)
            for k in list(left):
                left[k] -= 1
                if left[k] == 0:
                    st = B.download(ids[k])       # synchronises; the state is the answer
                    done[k] = time.perf_counter() - t0
                    if keep_results:
                        results[k] = st
                    B.remove(ids[k])
                    del left[k]
        makespan = time.perf_counter() - t0
    return done, makespan, results


def run_plain(cases, cfl=0.3, fast=True, keep_results=False):
    """One case at a time, each in its own context (the plain solver)."""
    done = {}
    results = {}
    fields = {c.ident: c.fields() for c in cases}
    t0 = time.perf_counter()
    for c in cases:
        with pympcf.Block(c.n, c.n, c.n, 1.0 / c.n, fast=fast) as b:
            b.data = fields[c.ident]
            for _ in range(c.steps):
                b.step(cfl, sync=False)
            st = b.data
            done[c.ident] = time.perf_counter() - t0
            if keep_results:
                results[c.ident] = st
    return done, time.perf_counter() - t0, results


def latency_stats(tts):
    v = np.array(sorted(tts))
    return {"p50_s": float(np.percentile(v, 50)), "p99_s": float(np.percentile(v, 99)),
            "mean_s": float(v.mean()), "max_s": float(v.max())}


def serve(requests, cfl=0.3, fast=True, graph=True, hbm_budget=None, max_live=256):
    """Continuous batching of a request stream (arrival times in seconds from
    start).  Admission: a request joins when its HBM footprint fits the
    remaining budget (default: 80 % of the device's free memory at start) and
    fewer than `max_live` cases are live.  Returns per-request
    time-to-solution (completion - arrival), the makespan and counters."""
    if hbm_budget is None:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        free_b, total_b = ctypes.c_size_t(), ctypes.c_size_t()
        hip.hipMemGetInfo(ctypes.byref(free_b), ctypes.byref(total_b))
        hbm_budget = int(0.8 * free_b.value)
    pending = sorted(requests, key=lambda r: r.arrival)
    queue = []          # arrived, not admitted (FIFO)
    live = {}           # ident -> [case id, steps left, bytes]
    used = 0
    tts = {}
    admitted_late = 0
    steps_run = 0
    with pympcf.Batch(fast=fast, graph=graph) as B:
        t0 = time.perf_counter()
        i = 0
        while i < len(pending) or queue or live:
            now = time.perf_counter() - t0
            while i < len(pending) and pending[i].arrival <= now:
                queue.append(pending[i])
                i += 1
            # admit in arrival order while the footprint fits
            while queue and len(live) < max_live:
                r = queue[0]
                nbytes = pympcf.case_bytes(r.n, r.n, r.n)
                if used + nbytes > hbm_budget and live:
                    break
                queue.pop(0)
                cid = B.add(r.fields())
                live[r.ident] = [cid, r.steps, nbytes, r]
                used += nbytes
                if time.perf_counter() - t0 > r.arrival + 1e-3:
                    admitted_late += 1
            if not live:
                if i < len(pending):
                    time.sleep(max(0.0, pending[i].arrival - (time.perf_counter() - t0)))
                continue
            B.step(cfl)
            steps_run += 1
            for k in list(live):
                live[k][1] -= 1
                if live[k][1] == 0:
                    cid, _, nbytes, r = live.pop(k)
                    B.download(cid)
                    tts[k] = time.perf_counter() - t0 - r.arrival
                    B.remove(cid)
                    used -= nbytes
        makespan = time.perf_counter() - t0
    return tts, makespan, {"batch_steps": steps_run, "queued_on_arrival": admitted_late}


def serve_plain(requests, cfl=0.3, fast=True):
    """The plain solver on the same stream: first come, first served, one case
    at a time in its own context."""
    pending = sorted(requests, key=lambda r: r.arrival)
    tts = {}
    t0 = time.perf_counter()
    for r in pending:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        f = r.fields()
        with pympcf.Block(r.n, r.n, r.n, 1.0 / r.n, fast=fast) as b:
            b.data = f
            for _ in range(r.steps):
                b.step(cfl, sync=False)
            _ = b.data
        tts[r.ident] = time.perf_counter() - t0 - r.arrival
    return tts, time.perf_counter() - t0


def default_budget(fraction=0.8):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free_b, total_b = ctypes.c_size_t(), ctypes.c_size_t()
    hip.hipMemGetInfo(ctypes.byref(free_b), ctypes.byref(total_b))
    return int(fraction * free_b.value)


def serve_native(requests, cfl=0.3, fast=True, policy="fifo_fill", fill_cells=0, hbm_budget=None, max_live=256,
                 graph=True, keep_results=False):
    """Submit each request at its arrival time to an mpcf_server and collect
    completion times.  Returns per-request time-to-solution, the makespan,
    the server's counters and (optionally) the results by request ident."""
    if hbm_budget is None:
        hbm_budget = default_budget()
    pending = sorted(requests, key=lambda r: r.arrival)
    with pympcf.Server(hbm_budget, policy=policy, fill_cells=fill_cells, max_live=max_live, cfl=cfl,
                       fast=fast, graph=graph) as S:
        t0 = pympcf.server_clock()
        by_id = {}
        for r in pending:
            now = pympcf.server_clock() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            rid = S.submit(r.fields(), r.steps, h=1.0 / r.n)
            by_id[rid] = r
        S.wait()
        makespan = pympcf.server_clock() - t0
        stats = S.stats()
        tts = {by_id[rid].ident: S.t_done[rid] - t0 - by_id[rid].arrival for rid in S.t_done}
        results = {by_id[rid].ident: out for rid, out in S.results.items()} if keep_results else {}
    return tts, makespan, stats, results


def run_served(cases, cfl=0.3, fast=True, policy="srpt_fill", fill_cells=8 << 20, hbm_budget=None, keep_results=False):
    """A closed ensemble through the serving engine: every case is submitted
    at time zero and the policy orders the work (least remaining work first
    finishes the small cases early while the cell budget keeps the GPU full).
    Returns per-case time-to-solution, the makespan and the results."""
    for c in cases:
        c.arrival = 0.0
    tts, makespan, _, results = serve_native(cases, cfl=cfl, fast=fast, policy=policy, fill_cells=fill_cells,
                                             hbm_budget=hbm_budget, keep_results=keep_results)
    return tts, makespan, results
