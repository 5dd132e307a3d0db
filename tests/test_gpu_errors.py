"""Error paths and the ghost band, through the C-ABI on the device.

* A cell that fails the reference's state assertions -- NaN, rho <= 0,
  E <= 0, G <= 0, c <= 0 (GPUkernels.cu:1171-1173,1218,
  MaxSpeedOfSound.cpp:31-48) -- is flagged by the speed-of-sound reduction
  and by the reduction folded into the last stage of a step, and surfaces as
  MPCF_ENUMERIC from mpcf_max_sos, mpcf_step and mpcf_last_dt.  The kernels
  finish normally: NaN arithmetic does not trap.  The maximum itself drops
  NaN operands, so without the flag a partly NaN block would keep a finite dt.
* A run made unstable on purpose (CFL 3 on a shock tube) is reported on the
  very step whose output first holds a non-physical cell; in the bit-exact
  build that step and the states are those of the CPU oracle.
* mpcf_load_ghosts fills the ghost band the way BoundaryConditions.h:60-109
  fills the reference's halo buffers: read back with mpcf_ghosts_read, every
  face equals the oracle's halo bit for bit (6 faces x absorbing /
  reflecting, four arithmetic builds).
"""
import sys

import numpy as np
import pytest

import states
from test_gpu_halo import DevBuf

pytestmark = pytest.mark.gpu

BUILDS = {
    "parity": {},
    "fast": {"fast": True},
    "p64": {"p64": True},
    "f64": {"fp64": True},
}


def say(*a):
    print("[gpu-errors]", *a, flush=True)
    sys.stdout.flush()


def bad_cells(s):
    """the flag's conditions evaluated in binary32 (GPU form of the pressure)"""
    r, mu, mv, mw, E, G, P = [s[q].astype(np.float32) for q in range(7)]
    with np.errstate(all="ignore"):
        p = (E - (mu * mu + mv * mv + mw * mw) * (np.float32(0.5) / r) - P) / G
        c = np.sqrt(((p + P) / G + p) / r)
        sp = c + np.maximum(np.maximum(np.abs(mu), np.abs(mv)), np.abs(mw)) / r
        ok = (r > 0) & (E > 0) & (G > 0) & (c > 0) & (sp <= np.finfo(np.float32).max)
    return ~ok


def poison(s, kind):
    out = s.copy()
    z, y, x = 3, 5, 7
    if kind == "nan_E":
        out[4, z, y, x] = np.nan
    elif kind == "nan_momentum":
        out[2, z, y, x] = np.nan
    elif kind == "inf_rho":
        out[0, z, y, x] = np.inf
    elif kind == "negative_rho":
        out[0, z, y, x] = -out[0, z, y, x]
    elif kind == "zero_G":
        out[5, z, y, x] = 0.0
    elif kind == "negative_P":
        # P below -G p (G + 1) / ... : negative enough that c^2 = ((p + P)/G + p)/rho < 0
        r, mu, mv, mw, E, G, P = (float(out[q, z, y, x]) for q in range(7))
        p = (E - 0.5 * (mu * mu + mv * mv + mw * mw) / r - P) / G
        out[6, z, y, x] = -2.0 * p * (G + 1.0) - 1.0
        out[4, z, y, x] = G * p + out[6, z, y, x] + 0.5 * (mu * mu + mv * mv + mw * mw) / r
    elif kind == "negative_c2":
        # p far below -pc: ((p + P)/G + p) < 0
        out[4, z, y, x] = out[6, z, y, x] - 50.0 * out[5, z, y, x] * 5000.0
    return out


KINDS = ["nan_E", "nan_momentum", "inf_rho", "negative_rho", "zero_G", "negative_P", "negative_c2"]


@pytest.mark.parametrize("build", list(BUILDS))
@pytest.mark.parametrize("kind", KINDS)
def test_bad_cell_is_reported(mpcf, build, kind):
    n = (16, 12, 10)
    s = poison(states.random_smooth(*n, seed=11), kind)
    assert bad_cells(s).sum() == 1
    say("bad cell", build, kind, n)
    with mpcf.Block(*n, 1.0 / 16, **BUILDS[build]) as b:
        b.data = s
        with pytest.raises(mpcf.MpcfError) as e:
            b.max_sos()
        assert e.value.code == mpcf.ENUMERIC
        with pytest.raises(mpcf.MpcfError) as e:
            b.step(0.3)
        assert e.value.code == mpcf.ENUMERIC
        # the asynchronous step queues normally; the flag waits in mpcf_last_dt
        b.data = s
        b.step(0.3, sync=False)
        with pytest.raises(mpcf.MpcfError) as e:
            b.last_dt()
        assert e.value.code == mpcf.ENUMERIC


@pytest.mark.parametrize("build", list(BUILDS))
def test_healthy_state_is_not_flagged(mpcf, build):
    n = (16, 12, 10)
    s = states.random_smooth(*n, seed=12)
    assert not bad_cells(s).any()
    say("healthy", build, n)
    with mpcf.Block(*n, 1.0 / 16, **BUILDS[build]) as b:
        b.data = s
        assert b.max_sos() > 0
        for _ in range(3):
            assert b.step(0.3) > 0
        b.last_dt()


def test_unstable_run_is_reported_on_its_first_bad_step(mpcf, oracle):
    """bit-exact build: CFL 3 on a 24x8x8 shock tube; steps 0..2 are
    physical, step 3 writes one cell with c^2 < 0 (as in the oracle)"""
    n = (24, 8, 8)
    h = 1.0 / 24
    s = states.shock_tube(*n, axis=0, seed=1)
    ob = oracle.Block(*n)
    ob.data = s
    say("unstable run, parity", n)
    with mpcf.Block(*n, h) as b:
        b.data = s
        first_bad = None
        for k in range(5):
            before = ob.data
            dref, _ = ob.step(h, 3.0, 1e30, [0] * 6)
            after = ob.data
            expect_bad = bad_cells(before).any() or bad_cells(after).any()
            if expect_bad:
                with pytest.raises(mpcf.MpcfError) as e:
                    b.step(3.0)
                assert e.value.code == mpcf.ENUMERIC, k
                d = e.value.dt
                first_bad = k if first_bad is None else first_bad
            else:
                d = b.step(3.0)
            assert d == dref, k
            if k <= 3:
                # up to and including the first bad step the states hold no
                # NaN and equal the oracle's bit for bit
                assert not np.isnan(after).any(), k
                assert np.array_equal(b.data, after), k
        assert first_bad == 3


@pytest.mark.parametrize("build", ["fast", "p64"])
def test_unstable_run_flags_follow_the_state(mpcf, build):
    """FAST builds: each step's flag agrees with the state before and after
    it (cells with c^2 within 1e-3 of zero would be ambiguous; none occur)"""
    n = (24, 8, 8)
    s = states.shock_tube(*n, axis=0, seed=1)
    say("unstable run", build, n)
    with mpcf.Block(*n, 1.0 / 24, **BUILDS[build]) as b:
        b.data = s
        seen_bad = False
        for k in range(6):
            before = b.data
            try:
                b.step(3.0)
                flagged = False
            except mpcf.MpcfError as e:
                assert e.code == mpcf.ENUMERIC
                flagged = True
            after = b.data
            expect = bad_cells(before).any() or bad_cells(after).any()
            assert flagged == expect, k
            seen_bad |= flagged
        assert seen_bad


@pytest.mark.parametrize("build", list(BUILDS))
@pytest.mark.parametrize("bc", [(0, 0, 0, 0, 0, 0), (1, 1, 1, 1, 1, 1), (0, 1, 1, 0, 0, 1), (1, 0, 0, 1, 1, 0)])
def test_load_ghosts_band_equals_oracle_halos(mpcf, oracle, build, bc):
    n = (13, 11, 9)
    s = states.random_smooth(*n, seed=21)
    fp64 = build == "f64"
    dtype = np.float64 if fp64 else np.float32
    ob = oracle.Block(*n, dtype=dtype)
    ob.data = s.astype(dtype)
    ob.apply_bc(list(bc))
    say("ghost band", build, bc, n)
    with mpcf.Block(*n, 1.0 / 13, bc=list(bc), **BUILDS[build]) as b:
        b.data = s
        b.load_ghosts()
        for face in range(6):
            count = b.halo_count(face)
            buf = DevBuf(count * np.dtype(dtype).itemsize)
            try:
                b.ghosts_read(face, buf.ptr)
                b.synchronize()
                got = buf.to_host(dtype).reshape(7, -1)
            finally:
                buf.free()
            ref = ob.get_halo(face)
            assert np.array_equal(got, ref), (build, bc, face)
