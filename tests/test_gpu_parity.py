"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bar (bit-exact build, the default context): every value equal to the oracle
(np.array_equal, so +0 == -0), dt bit-identical.  The FMA-contracted build
(MPCF_FLAG_FAST) is checked at a stated tolerance in test_gpu_fast.py.
"""
import sys

import numpy as np
import pytest

import states

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-parity]", *a, flush=True)
    sys.stdout.flush()


def assert_equal(got, ref, what):
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    if not np.array_equal(got, ref):
        bad = got != ref
        d = np.abs(got.astype(np.float64) - ref.astype(np.float64))
        idx = np.argwhere(bad)[:5]
        raise AssertionError(f"{what}: {bad.sum()} of {bad.size} differ, max |diff| {d.max():.3e}, "
                             f"first at {idx.tolist()}: got {got[bad][:5]}, ref {ref[bad][:5]}")


def weno_windows(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(5)[None, :]
    parts = [
        rng.uniform(-1, 1, (n, 5)),                                             # noise
        np.sin(rng.uniform(0, 1, (n, 1)) * t + rng.uniform(0, 6, (n, 1))),      # smooth
        1000.0 * (t >= rng.integers(1, 5, (n, 1))) + 1.0,                       # jumps
        4828.79 + 1e-3 * rng.uniform(-1, 1, (n, 5)),                            # near-constant, large
        np.full((n, 5), 2.5),                                                   # constant
        rng.uniform(0, 1, (n, 1)) * t + rng.uniform(-5, 5, (n, 1)),             # linear
    ]
    return np.ascontiguousarray(np.concatenate(parts).astype(np.float32))


def test_weno_bitexact(mpcf, oracle):
    w = weno_windows(40000, 1)
    say("weno windows", w.shape)
    m, p = mpcf.test_weno(w)
    rm, rp = oracle.weno(w)
    assert_equal(m, rm, "weno minus-state")
    assert_equal(p, rp, "weno plus-state")


@pytest.mark.parametrize("d", [0, 1, 2])
def test_face_bitexact(mpcf, oracle, d):
    st = states.random_stencils(60000, seed=10 + d)
    say("face batch dir", d, st.shape)
    got = mpcf.test_face(st, d)
    ref = oracle.face_batch(st, d)
    assert_equal(got, ref, f"face outputs dir {d}")


@pytest.mark.parametrize("cpu_form", [False, True])
def test_max_sos_bitexact(mpcf, oracle, cpu_form):
    nx, ny, nz = 24, 20, 28
    s = states.random_smooth(nx, ny, nz, seed=3)
    say("max_sos", (nx, ny, nz), "cpu_form", cpu_form)
    with mpcf.Block(nx, ny, nz, 1.0 / 28, sos_cpu=cpu_form) as b:
        b.data = s
        got = b.max_sos()
    ob = oracle.Block(nx, ny, nz)
    ob.data = s
    ref = ob.max_sos(cpu_form)
    assert np.float32(got) == np.float32(ref), (got, ref)


# the last shape spans three fused x+y tiles (58 cells each) with a partial one
BCS = [[0] * 6, [1] * 6, [0, 1, 1, 0, 1, 0], [1, 0, 0, 1, 0, 1]]
SHAPES = [(20, 24, 28), (33, 17, 9), (7, 66, 12), (131, 9, 6)]


@pytest.mark.parametrize("split", [False, True], ids=["fused_xy", "split_xy"])
@pytest.mark.parametrize("shape,bc", list(zip(SHAPES, BCS)))
def test_ghosts_and_convection_bitexact(mpcf, oracle, shape, bc, split):
    nx, ny, nz = shape
    s = states.random_smooth(nx, ny, nz, seed=nx + ny)
    R0 = states.random_smooth(nx, ny, nz, seed=99, two_phase=False) * np.float32(1e-3)
    a, dtinvh = np.float32(-17.0 / 32), np.float32(0.0123)
    say("convection", shape, "bc", bc, "split_xy", split)
    with mpcf.Block(nx, ny, nz, 1.0 / max(shape), bc=bc, split_xy=split) as b:
        b.data = s
        b.tmp = R0
        b.load_ghosts()
        b.convection(a, dtinvh)
        got_R = b.tmp
        got_u = b.data
    ob = oracle.Block(nx, ny, nz)
    ob.data = s
    ob.tmp = R0
    ob.apply_bc(bc)
    ob.rhs(a, dtinvh)
    assert_equal(got_u, s, "u untouched by convection")
    for q in range(7):
        assert_equal(got_R[q], ob.tmp[q], f"R' field {q}")


def test_process_all_matches_convection_update(mpcf, oracle):
    nx, ny, nz = 24, 16, 20
    s = states.random_smooth(nx, ny, nz, seed=5)
    R0 = states.random_smooth(nx, ny, nz, seed=6, two_phase=False) * np.float32(1e-3)
    a, b_, dtinvh = np.float32(-32.0 / 27), np.float32(0.75), np.float32(0.02)
    say("process_all vs convection+update", (nx, ny, nz))
    with mpcf.Block(nx, ny, nz, 1.0 / 24, bc=[0, 0, 1, 1, 0, 1]) as b1, \
         mpcf.Block(nx, ny, nz, 1.0 / 24, bc=[0, 0, 1, 1, 0, 1]) as b2:
        for b in (b1, b2):
            b.data = s
            b.tmp = R0
            b.load_ghosts()
        b1.process_all(a, b_, dtinvh)
        b2.convection(a, dtinvh)
        b2.update(b_)
        u1, R1, u2, R2 = b1.data, b1.tmp, b2.data, b2.tmp
    assert_equal(u1, u2, "u (process_all vs convection+update)")
    assert_equal(R1, R2, "R (process_all vs convection+update)")
    ob = oracle.Block(nx, ny, nz)
    ob.data = s
    ob.tmp = R0
    ob.apply_bc([0, 0, 1, 1, 0, 1])
    ob.rhs(a, dtinvh)
    ob.update(b_)
    assert_equal(u1, ob.data, "u vs oracle")
    assert_equal(R1, ob.tmp, "R vs oracle")


@pytest.mark.parametrize("cpu_sos,split", [(False, False), (True, False), (False, True)])
def test_lsrk3_steps_bitexact(mpcf, oracle, cpu_sos, split):
    nx, ny, nz = 28, 24, 20
    h = 1.0 / 28
    bc = [0, 1, 0, 0, 1, 0]
    s = states.random_smooth(nx, ny, nz, seed=11)
    say("lsrk3 3 steps", (nx, ny, nz), "cpu_sos", cpu_sos, "split_xy", split)
    ob = oracle.Block(nx, ny, nz)
    ob.data = s
    with mpcf.Block(nx, ny, nz, h, sos_cpu=cpu_sos, bc=bc, split_xy=split) as b:
        b.data = s
        for it in range(3):
            dt = b.step(0.3)
            dref, _ = ob.step(h, 0.3, 1e30, bc, cpu_sos)
            assert dt == dref, (it, dt, dref)
        u, R = b.data, b.tmp
    assert_equal(u, ob.data, "u after 3 steps")
    assert_equal(R, ob.tmp, "R after 3 steps")


@pytest.mark.parametrize("shape,bc", [((5, 7, 6), [1] * 6), ((6, 5, 9), [0, 1, 1, 0, 0, 1]), ((9, 6, 5), [1, 0, 0, 1, 1, 1])],
                         ids=["x5-reflect", "y5-mixed", "z5-mixed"])
def test_lsrk3_small_extents_bitexact(mpcf, oracle, shape, bc):
    """Extents of 5-6 cells: the stage outputs write the ghosts of both faces
    of an axis from overlapping boundary cells."""
    nx, ny, nz = shape
    h = 1.0 / max(shape)
    s = states.random_smooth(nx, ny, nz, seed=13)
    say("lsrk3 small extents", shape, bc)
    ob = oracle.Block(nx, ny, nz)
    ob.data = s
    for split in (False, True):
        ob.data = s
        ob.tmp = np.zeros_like(s)
        with mpcf.Block(nx, ny, nz, h, bc=bc, split_xy=split) as b:
            b.data = s
            for it in range(4):
                dt = b.step(0.3)
                dref, _ = ob.step(h, 0.3, 1e30, bc)
                assert dt == dref, (split, it, dt, dref)
            u, R = b.data, b.tmp
        assert_equal(u, ob.data, f"u after 4 steps (split {split})")
        assert_equal(R, ob.tmp, f"R after 4 steps (split {split})")


def test_dt_max_clamp(mpcf, oracle):
    nx = ny = nz = 16
    s = states.uniform(nx, ny, nz, rho=1.0, u=0.1, p=1.0)
    with mpcf.Block(nx, ny, nz, 1.0 / 16) as b:
        b.data = s
        dt = b.step(0.3, dt_max=1e-4)
    assert dt == 1e-4


def test_config1_shock_tube_10_steps_bitexact(mpcf, oracle):
    """64^3 air|water tube, CFL 0.3, transmissive, 10 LSRK3 steps."""
    n = 64
    s = states.shock_tube(n, n, n, axis=0, seed=2024)
    say("config-1 shock tube 64^3 x 10 steps")
    ob = oracle.Block(n, n, n)
    ob.data = s
    dts = []
    with mpcf.Block(n, n, n, 1.0 / n, bc=[0] * 6) as b:
        b.data = s
        for it in range(10):
            dts.append(b.step(0.3))
        u = b.data
    for it in range(10):
        dref, _ = ob.step(1.0 / n, 0.3, 1e30, [0] * 6)
        assert dts[it] == dref, (it, dts[it], dref)
    assert_equal(u, ob.data, "config-1 state after 10 steps")


def test_fresh_block_ignores_freed_memory(mpcf, oracle):
    """A block created over memory freed by another block (here one holding
    NaN everywhere) starts from zero: its steps match the oracle bit for bit.
    The zeroing must be ordered on the block's own stream before its first
    upload and stage."""
    n = 96
    bad = np.full((7, n, n, n), np.nan, dtype=np.float32)
    say("freed-memory reuse", n)
    with mpcf.Block(n, n, n, 1.0 / n) as b:
        b.data = bad
        b.tmp = bad
    s = states.random_smooth(n, n, n, seed=17)
    ob = oracle.Block(n, n, n)
    ob.data = s
    with mpcf.Block(n, n, n, 1.0 / n) as b:
        b.data = s
        for _ in range(2):
            assert b.step(0.3) == ob.step(1.0 / n, 0.3, 1e30, [0] * 6)[0]
        u, R = b.data, b.tmp
    assert_equal(u, ob.data, "u after reuse")
    assert_equal(R, ob.tmp, "R after reuse")


def test_folded_speed_of_sound_follows_every_write(mpcf, oracle):
    """The last stage of a step folds the next step's max speed of sound into
    its z sweep.  Anything else that writes u (upload, process_all, the full
    reduction of mpcf_max_sos) must make the next step reduce the state again:
    the dt sequence of a run interleaved with such calls equals the oracle's."""
    n = 40
    h = 1.0 / n
    s1 = states.random_smooth(n, n, n, seed=31)
    s2 = states.shock_tube(n, n, n, axis=1, seed=32)
    ob = oracle.Block(n, n, n)
    say("folded sos: steps, upload, max_sos, process_all", n)
    with mpcf.Block(n, n, n, h) as b:
        b.data = s1
        ob.data = s1
        for _ in range(2):
            assert b.step(0.3) == ob.step(h, 0.3, 1e30, [0] * 6)[0]
        # a new state: the folded maximum of the old one must not be used
        b.data = s2
        ob.data = s2
        for _ in range(2):
            assert b.step(0.3) == ob.step(h, 0.3, 1e30, [0] * 6)[0]
        # the full reduction in between gives the same maximum
        sos = b.max_sos()
        assert b.step(0.3) == ob.step(h, 0.3, 1e30, [0] * 6)[0]
        assert b.last_dt()[1] == sos
        # one stage driven from the host, then steps again
        dt = 0.3 * h / b.max_sos()
        b.process_all(0.0, 0.25, float(np.float32(dt / h)))
        ob.apply_bc([0] * 6)
        ob.rhs(0.0, float(np.float32(dt / h)))
        ob.update(0.25)
        for _ in range(2):
            assert b.step(0.3) == ob.step(h, 0.3, 1e30, [0] * 6)[0]
        assert_equal(b.data, ob.data, "u after interleaved calls")


@pytest.mark.parametrize("flags", [dict(fast=True), dict(p64=True), dict(fp64=True)])
def test_folded_speed_of_sound_equals_full_reduction(mpcf, flags):
    """In every build the folded maximum is the full reduction's: a run whose
    every step starts from a fresh reduction (mpcf_max_sos clears the folded
    value) takes the same dts to the same state."""
    n = 48
    s = states.random_smooth(n, n, n, seed=41)
    say("folded vs full sos", flags)
    with mpcf.Block(n, n, n, 1.0 / n, **flags) as a, mpcf.Block(n, n, n, 1.0 / n, **flags) as b:
        a.data = s
        b.data = s
        for _ in range(3):
            da = a.step(0.3)
            b.max_sos()
            assert da == b.step(0.3)
        ua = a.download_f64() if flags.get("fp64") else a.data
        ub = b.download_f64() if flags.get("fp64") else b.data
        assert np.array_equal(ua, ub)


@pytest.mark.parametrize("slabs,split", [(4, False), (5, False), (33, False), (4, True), (7, True)])
def test_pipelined_step_bit_exact(mpcf, oracle, slabs, split):
    """mpcf_step over z-slabs on two streams (fused or split x and y sweeps;
    even, uneven and 3-4-plane-thin slabs) steps exactly as the oracle: slab
    boundaries move no result."""
    nx, ny, nz = 24, 20, 130
    s = states.random_smooth(nx, ny, nz, seed=51)
    bc = [1, 0, 0, 1, 1, 0]
    ob = oracle.Block(nx, ny, nz)
    ob.data = s
    say("pipelined parity step", (nx, ny, nz), slabs, "split" if split else "fused")
    with mpcf.Block(nx, ny, nz, 1.0 / nz, bc=bc, split_xy=split, pipeline=slabs) as b:
        b.data = s
        for _ in range(3):
            assert b.step(0.3) == ob.step(1.0 / nz, 0.3, 1e30, bc)[0]
        assert_equal(b.data, ob.data, "u after pipelined steps")
        assert_equal(b.tmp, ob.tmp, "R after pipelined steps")


@pytest.mark.parametrize("flags,slabs", [(dict(fast=True), 3), (dict(fast=True), 8), (dict(p64=True), 4)])
def test_pipelined_step_equals_whole_block(mpcf, flags, slabs):
    n, nz = 48, 192
    s = states.random_smooth(n, n, nz, seed=52)
    say("pipelined vs whole-block", flags, slabs)
    with mpcf.Block(n, n, nz, 1.0 / nz, pipeline=slabs, **flags) as a, \
            mpcf.Block(n, n, nz, 1.0 / nz, pipeline=False, **flags) as b:
        a.data = s
        b.data = s
        for _ in range(3):
            assert a.step(0.3) == b.step(0.3)
        assert np.array_equal(a.data, b.data) and np.array_equal(a.tmp, b.tmp)


@pytest.mark.parametrize("flags,slabs", [(dict(fast=True), 4), (dict(fast=True), 8), (dict(p64=True), 3),
                                         (dict(), 4), (dict(fast=True), False), (dict(fp64=True), 2)])
def test_steps_equal_single_steps(mpcf, flags, slabs):
    """mpcf_steps (consecutive pipelined FAST steps overlapped across the step
    boundary, dt on the second stream) equals the same number of mpcf_step
    calls of a whole-block context bit for bit, dt included, also when
    single steps and multi-step calls alternate"""
    n, nz = 48, 192
    s = states.random_smooth(n, n, nz, seed=54)
    say("mpcf_steps vs mpcf_step", flags, slabs)
    with mpcf.Block(n, n, nz, 1.0 / nz, pipeline=slabs, **flags) as a, \
            mpcf.Block(n, n, nz, 1.0 / nz, pipeline=False, **flags) as b:
        a.data = s
        b.data = s
        for chunk in (5, 1, 3):
            if chunk == 1:
                da = a.step(0.3)
            else:
                a.steps(chunk, 0.3)
                da = a.last_dt()[0]
            for _ in range(chunk):
                db = b.step(0.3)
            assert da == db, chunk
            assert np.array_equal(a.data, b.data) and np.array_equal(a.tmp, b.tmp), chunk


def test_steps_arguments(mpcf):
    with mpcf.Block(16, 16, 16, 1.0 / 16, fast=True) as b:
        b.data = states.uniform(16, 16, 16)
        b.steps(0, 0.3)
        for bad in (dict(nsteps=-1, cfl=0.3), dict(nsteps=2, cfl=0.0), dict(nsteps=2, cfl=float("nan"))):
            with pytest.raises(mpcf.MpcfError):
                b.steps(**bad)


def sweep_launches(mpcf, b, steps=2):
    b.timing(True)
    b.timing_read()
    for _ in range(steps):
        b.step(0.3, sync=False)
    b.synchronize()
    t = b.timing_read()
    b.timing(False)
    return t["sweep_x"][1] // (3 * steps), t["sweep_z"][1] // (3 * steps)


def test_pipeline_slab_choice(mpcf):
    """The automatic choice for blocks of 2^24 cells or more in the FAST
    builds (fused x+y sweep): NZ/64 slabs for NZ < 320, NZ/100 for NZ < 512,
    min(8, NZ/64) above; whole-block sweeps below 2 slabs or 2^24 cells, in
    the bit-exact build (fused or split), for the split sweeps and under
    MPCF_FLAG_NO_PIPELINE; forced counts as asked, split sweeps included."""
    cases = [((128, 128, 128), dict(fast=True), 1), ((256, 256, 256), dict(fast=True), 4),
             ((256, 256, 384), dict(fast=True), 3), ((160, 160, 1024), dict(fast=True), 8),
             ((128, 128, 1280), dict(fast=True), 8), ((192, 192, 640), dict(fast=True), 8),
             ((512, 512, 64), dict(fast=True), 1), ((256, 256, 256), dict(split_xy=True), 1),
             ((256, 256, 256), dict(), 1), ((256, 256, 256), dict(pipeline=4), 4),
             ((64, 64, 96), dict(pipeline=3), 3),
             ((256, 256, 256), dict(fast=True, pipeline=False), 1), ((64, 64, 96), dict(fast=True, pipeline=6), 6)]
    for (nx, ny, nz), kw, want in cases:
        say("slab choice", (nx, ny, nz), kw)
        with mpcf.Block(nx, ny, nz, 1.0 / nx, **kw) as b:
            b.data = states.uniform(nx, ny, nz)
            assert sweep_launches(mpcf, b) == (want, want), (nx, ny, nz, kw)


def test_pipeline_arguments(mpcf):
    with mpcf.Block(16, 16, 40, 1.0 / 16, fast=True) as b:
        for bad in (-2, 1, 41, 65):
            with pytest.raises(mpcf.MpcfError, match="slabs"):
                b.set_pipeline(bad)
        for ok in (-1, 0, 2, 40):
            b.set_pipeline(ok)
    with mpcf.Block(16, 16, 40, 1.0 / 16) as b:          # split x, y sweeps (bit-exact build default)
        for ok in (4, 0, -1):
            b.set_pipeline(ok)


def test_pipeline_on_caller_stream(mpcf):
    """A pipelined step on a stream the caller made: the caller's stream is
    ordered after the step's last sweep (synchronizing only that stream
    finishes the step), and the results equal whole-block stepping."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    n, nz = 40, 96
    s = states.random_smooth(n, n, nz, seed=53)
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    say("pipelined step on a caller stream")
    try:
        with mpcf.Block(n, n, nz, 1.0 / nz, fast=True, pipeline=4) as a, \
                mpcf.Block(n, n, nz, 1.0 / nz, fast=True, pipeline=False) as b:
            own = a.stream
            a.set_stream(None)              # a fresh stream of its own, made before the old one goes
            assert a.stream and a.stream != own
            a.set_stream(st.value)
            assert a.stream == st.value
            a.data = s
            b.data = s
            for _ in range(4):
                a.step(0.3, sync=False)
                b.step(0.3)
            assert hip.hipStreamSynchronize(st) == 0
            assert np.array_equal(a.data, b.data) and np.array_equal(a.tmp, b.tmp)
            a.set_stream(None)
            assert a.stream != st.value
            a.step(0.3)
            b.step(0.3)
            assert np.array_equal(a.data, b.data)
    finally:
        hip.hipStreamDestroy(st)
