"""Batched ensembles: every case gives the same result alone (its own
context) and in a batch -- bit for bit, since the batch kernels run the same
per-block bodies -- with eager launches, with graph replay (checked after every
replay, with eager work and synchronisations in between), across additions
and removals between steps (continuous batching), and after autotuning; in
both launch modes (per stage, two persistent launches pulling (case, tile)
items from a device queue, in three workgroup sizes and with march lengths
short enough that every case has many items; or flat launches per case
group)."""
import sys

import numpy as np
import pytest

import states

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-batch]", *a, flush=True)
    sys.stdout.flush()


def cases(mpcf):
    out = [
        (states.random_smooth(20, 20, 20, seed=1), [0] * 6),
        (mpcf.ic_shock_tube(24, 16, 20, 0, 1e-3, 2), [0, 0, 1, 1, 0, 0]),
        (states.random_smooth(33, 17, 9, seed=3), [1] * 6),
        (mpcf.ic_bubble_collapse(48, 48, 48, 0.2, 100.0), [0] * 6),
        (mpcf.ic_shock_bubble(16, 16, 16, 4), [1, 0, 0, 1, 1, 0]),
        (states.random_smooth(70, 12, 10, seed=6), [0] * 6),
    ]
    return out


def alone(mpcf, s, bc, steps, fast=False):
    nz, ny, nx = s.shape[1:]
    with mpcf.Block(nx, ny, nz, 1.0 / max(nx, ny, nz), fast=fast, bc=bc, split_xy=False) as b:
        b.data = s
        dts = [b.step(0.3) for _ in range(steps)]
        return b.data, b.tmp, dts


MODES = {"persistent": (1, 256, 0), "persistent64": (1, 64, 8), "persistent128": (1, 128, 16), "groups": (0, 0, 0)}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("fast", [False, True], ids=["bitexact", "fast"])
@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_batch_equals_alone(mpcf, fast, graph, mode):
    cs = cases(mpcf)
    steps = 3
    m, threads, march = MODES[mode]
    say("batch", len(cs), "cases fast", fast, "graph", graph, "mode", mode)
    with mpcf.Batch(fast=fast, graph=graph, mode=m, threads=threads) as B:
        ids = [B.add(s, bc=bc) for s, bc in cs]
        if march:
            for i in ids:
                B.set_march(i, march, march)
        per_step = []
        for k in range(steps):
            B.step(0.3)
            # eager work and a synchronisation between replays
            per_step.append([B.last_dt(i)[0] for i in ids])
            _ = B.download(ids[k % len(ids)])
        q = B.query()
        if graph:
            assert q["captures"] == 1, q
        got = [(B.download(i), B.download(i, mpcf.TMP)) for i in ids]
    for j, (s, bc) in enumerate(cs):
        u, R, dts = alone(mpcf, s, bc, steps, fast=fast)
        assert [per_step[k][j] for k in range(steps)] == dts, j
        assert np.array_equal(got[j][0], u), j
        assert np.array_equal(got[j][1], R), j


@pytest.mark.parametrize("count", [1, 2, 3, 4])
def test_batch_group_splits_equal_alone(mpcf, count):
    """A batch steps its cases as up to four groups of about equal cells on as
    many streams: with one to four cases (one case a group; the six cases of
    test_batch_equals_alone split unevenly), every case still equals the same
    case stepped alone, through graph replays."""
    cs = cases(mpcf)[2:2 + count]
    steps = 3
    say("group split", count, "cases")
    with mpcf.Batch(fast=True, graph=True) as B:
        ids = [B.add(s, bc=bc) for s, bc in cs]
        for _ in range(steps):
            B.step(0.3)
        got = [(B.download(i), [B.last_dt(i)[0]]) for i in ids]
    for j, (s, bc) in enumerate(cs):
        u, _, dts = alone(mpcf, s, bc, steps, fast=True)
        assert got[j][1] == dts[-1:], j
        assert np.array_equal(got[j][0], u), j


def test_graph_replays_match_eager_each_step(mpcf):
    cs = cases(mpcf)[:4]
    say("graph replay vs eager, step by step")
    with mpcf.Batch(graph=True) as G, mpcf.Batch(graph=False) as E:
        gi = [G.add(s, bc=bc) for s, bc in cs]
        ei = [E.add(s, bc=bc) for s, bc in cs]
        for k in range(4):
            G.step(0.3)
            E.step(0.3)
            for a, b in zip(gi, ei):
                assert np.array_equal(G.download(a), E.download(b)), (k, a)


def test_continuous_batching_add_remove(mpcf):
    """Cases join and leave between steps; each finishes with exactly the
    state it has when run alone for its own number of steps."""
    cs = cases(mpcf)
    plan = [(0, 0, 4), (1, 0, 2), (2, 1, 3), (3, 2, 2), (4, 2, 4), (5, 3, 1)]   # (case, join step, steps)
    say("continuous batching")
    done = {}
    with mpcf.Batch(fast=True, graph=True) as B:
        live = {}
        for t in range(8):
            for c, join, n in plan:
                if join == t:
                    live[c] = [B.add(cs[c][0], bc=cs[c][1]), n]
            if not live:
                break
            B.step(0.3)
            for c in list(live):
                live[c][1] -= 1
                if live[c][1] == 0:
                    done[c] = B.download(live[c][0])
                    B.remove(live[c][0])
                    del live[c]
    assert sorted(done) == [c for c, _, _ in plan]
    for c, _, n in plan:
        u, _, _ = alone(mpcf, cs[c][0], cs[c][1], n, fast=True)
        assert np.array_equal(done[c], u), c


def test_autotune_keeps_state_and_results(mpcf):
    cs = cases(mpcf)
    say("autotune")
    with mpcf.Batch(fast=True) as B:
        ids = [B.add(s, bc=bc) for s, bc in cs]
        best = B.autotune(steps=1)
        ny_max = max(s.shape[2] for s, _ in cs)
        assert best in (0, 8, 16, 32, 64, ny_max), best
        mode, threads, lxy, lz = B.tuning(ny_max)
        say("  tuned: mode", mode, "threads", threads, "x+y march", lxy, "z march", lz)
        assert mode in (mpcf.BATCH_GROUPS, mpcf.BATCH_PERSISTENT) and threads in (64, 128, 256)
        assert lxy == best
        for i, (s, _) in zip(ids, cs):
            assert np.array_equal(B.download(i), s)     # tuning did not advance the cases
        B.step(0.3)
        got = [B.download(i) for i in ids]
    for j, (s, bc) in enumerate(cs):
        u, _, _ = alone(mpcf, s, bc, 1, fast=True)
        assert np.array_equal(got[j], u), j


@pytest.mark.parametrize("policy,fill,graph", [("all", 0, True), ("fifo_fill", 40000, True),
                                               ("srpt_fill", 40000, True), ("hybrid", 40000, True),
                                               ("fifo_slack", 8000, True), ("fifo_fill", 40000, False)],
                         ids=["all", "fifo_fill", "srpt_fill", "hybrid", "fifo_slack", "fifo_fill-eager"])
def test_server_results_equal_alone(mpcf, policy, fill, graph):
    """Requests submitted while others run, stepped in changing active subsets,
    each return exactly the state it reaches alone."""
    import time
    cs = cases(mpcf)
    steps = [3, 5, 2, 4, 1, 6]
    say("server", policy, "fill", fill, "graph", graph)
    with mpcf.Server(1 << 30, policy=policy, fill_cells=fill, max_live=4, fast=True, graph=graph) as S:
        ids = []
        for k, ((s, bc), n) in enumerate(zip(cs, steps)):
            h = 1.0 / max(s.shape[1:])
            ids.append(S.submit(s, n, h=h, bc=bc))
            if k % 2:
                time.sleep(0.002)       # later requests arrive while earlier ones run
        S.wait(60.0)
        st = S.stats()
        got = [S.results[i] for i in ids]
    assert st["completed"] == len(cs) and st["peak_active"] <= 4, st
    for j, ((s, bc), n) in enumerate(zip(cs, steps)):
        u, _, _ = alone(mpcf, s, bc, n, fast=True)
        assert np.array_equal(got[j], u), j


LONE_BUILDS = {"fast": dict(fast=True), "bitexact": dict(fast=False), "fast-soscpu": dict(fast=True, sos_cpu=True),
               "p64": dict(fast=True, p64=True)}


@pytest.mark.parametrize("build", list(LONE_BUILDS))
def test_server_lone_request_runs_whole_block_kernels(mpcf, build):
    """A lone request of 2^18 cells or more steps on the whole-block kernels,
    enqueued eagerly (no graph is captured); a smaller lone one replays the
    batch graph.  Both return exactly the state they reach alone, in the FAST,
    bit-exact and selective-binary64 builds and with the CPU form of the speed
    of sound."""
    kw = LONE_BUILDS[build]
    big = mpcf.ic_bubble_collapse(72, 64, 64, 0.2, 100.0)        # 294912 cells
    small = states.random_smooth(20, 20, 20, seed=1)
    for s, n, solo in ((big, 7, True), (small, 4, False)):
        say("server lone request", s.shape[1:], "solo" if solo else "batch graph", build)
        with mpcf.Server(1 << 30, policy="fifo_slack", fill_cells=1 << 20, max_live=4, graph=True, **kw) as S:
            i = S.submit(s, n, h=1.0 / max(s.shape[1:]))
            S.wait(60.0)
            st = S.stats()
            got = S.results[i]
        assert st["steps"] == n and st["captures"] == (0 if solo else 1), st
        nz, ny, nx = s.shape[1:]
        with mpcf.Block(nx, ny, nz, 1.0 / max(nx, ny, nz), split_xy=False, **kw) as b:
            b.data = s
            for _ in range(n):
                b.step(0.3)
            u = b.data
        assert np.array_equal(got, u)


@pytest.mark.parametrize("policy", ["fifo_slack", "all"])
def test_server_lone_request_joined_and_left(mpcf, policy):
    """A large request steps alone (whole-block kernels), is joined by small
    ones (batch graph), and finishes alone again; every result is exactly the
    state it reaches alone."""
    import time
    big = mpcf.ic_bubble_collapse(72, 64, 64, 0.2, 100.0)
    small = cases(mpcf)[:3]
    steps = [3, 5, 2]
    say("server lone request joined and left,", policy)
    with mpcf.Server(1 << 30, policy=policy, fill_cells=1 << 22, max_live=4, fast=True, graph=True) as S:
        ib = S.submit(big, 150, h=1.0 / 72)
        time.sleep(0.004)        # the large request runs alone for a while
        ids = [S.submit(s, n, h=1.0 / max(s.shape[1:]), bc=bc) for (s, bc), n in zip(small, steps)]
        S.wait(60.0)
        st = S.stats()
        got_big = S.results[ib]
        got = [S.results[i] for i in ids]
    say("  stats", st)
    assert st["completed"] == 4 and st["steps"] >= 150, st
    u, _, _ = alone(mpcf, big, [0] * 6, 150, fast=True)
    assert np.array_equal(got_big, u)
    for j, ((s, bc), n) in enumerate(zip(small, steps)):
        u, _, _ = alone(mpcf, s, bc, n, fast=True)
        assert np.array_equal(got[j], u), j


def test_server_budget_queues_requests(mpcf):
    """A budget that holds one case at a time still completes every request."""
    s = states.random_smooth(20, 20, 20, seed=1)
    one = mpcf.case_bytes(20, 20, 20)
    with mpcf.Server(one + 4096, policy="all", max_live=8, fast=True) as S:
        ids = [S.submit(s, 2, h=0.05) for _ in range(3)]
        S.wait(60.0)
        st = S.stats()
        outs = [S.results[i] for i in ids]
    assert st["peak_active"] == 1 and st["completed"] == 3, st
    u, _, _ = alone(mpcf, s, [0] * 6, 2, fast=True)
    for o in outs:
        assert np.array_equal(o, u)
