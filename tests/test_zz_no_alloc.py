"""The allocation-free entry point, mpcf_max_sos_noalloc, through the C-ABI.

It runs the speed-of-sound reduction and the dt of a step over a state the
caller keeps in its own device memory (host layout, seven fields at a
caller-chosen stride), with the reduction's working state in the library's
static device array instead of a context's allocation.  Checked against the
entry points it is a variant of, on the same state and flags, bit for bit:
the maximum against mpcf_max_sos (binary32) and mpcf_last_sos64 (the
binary64 reduction of the P64 / FP64 builds), dt against mpcf_step's, the
non-physical-cell flag against mpcf_max_sos's MPCF_ENUMERIC.  Two calls in a
row, the second on a smaller state, so the second call runs on the array the
first one left.
"""
import ctypes
import sys

import numpy as np
import pytest

import states
from test_gpu_halo import DevBuf

pytestmark = pytest.mark.gpu

SHAPES = [(40, 36, 33), (17, 12, 9)]   # (nx, ny, nz): the second call is smaller


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)
    sys.stdout.flush()


def builds(mpcf):
    return {
        "parity": 0,
        "parity_cpu_form": mpcf.FLAG_SOS_CPU,
        "fast": mpcf.FLAG_FAST,
        "p64": mpcf.FLAG_P64 | mpcf.FLAG_FAST,
        "f64": mpcf.FLAG_FP64,
    }


def block_kwargs(mpcf, flags):
    return {"fast": bool(flags & mpcf.FLAG_FAST), "sos_cpu": bool(flags & mpcf.FLAG_SOS_CPU),
            "p64": bool(flags & mpcf.FLAG_P64), "fp64": bool(flags & mpcf.FLAG_FP64)}


def noalloc(mpcf, s, h, cfl, dt_max, flags, pad):
    """one call on the state s (7, nz, ny, nx), fields `pad` elements apart
    beyond the cell count; returns (sos, dt, bad) read back from the device"""
    hip = DevBuf.hip()
    nz, ny, nx = s.shape[1:]
    cells = nx * ny * nz
    fs = cells + pad
    dtype = np.float64 if flags & mpcf.FLAG_FP64 else np.float32
    host = np.full((7, fs), np.nan, dtype)          # the gap between fields is never read
    host[:, :cells] = s.reshape(7, cells)
    ubuf = DevBuf(host.nbytes)
    obuf = DevBuf(32)
    try:
        assert hip.hipMemcpy(ubuf.ptr, host.ctypes.data, host.nbytes, 1) == 0   # host to device
        marker = np.full(4, -1.0)                    # outputs start as garbage
        assert hip.hipMemcpy(obuf.ptr, marker.ctypes.data, 32, 1) == 0
        rc = mpcf.lib().mpcf_max_sos_noalloc(ubuf.ptr, fs, nx, ny, nz, h, cfl, dt_max, flags,
                                             obuf.ptr, obuf.ptr + 8, obuf.ptr + 16, None)
        assert rc == 0, mpcf.lib().mpcf_last_error()
        assert hip.hipDeviceSynchronize() == 0
        raw = obuf.to_host(np.uint8)
    finally:
        ubuf.free()
        obuf.free()
    sos = raw[0:8].view(np.float64)[0]
    dt = raw[8:16].view(np.float64)[0]
    bad = int(raw[16:20].view(np.int32)[0])
    return sos, dt, bad


def reference(mpcf, s, h, cfl, dt_max, flags):
    """the context entry points on the same state: (max_sos binary32 or None,
    sos of the step's reduction in binary64, the step's dt)"""
    nz, ny, nx = s.shape[1:]
    with mpcf.Block(nx, ny, nz, h, **block_kwargs(mpcf, flags)) as b:
        b.data = s
        sos32 = None
        if not flags & (mpcf.FLAG_P64 | mpcf.FLAG_FP64):
            sos32 = b.max_sos()
        dt = b.step(cfl, dt_max)
        sos64 = b.last_sos64()
    return sos32, sos64, dt


@pytest.mark.parametrize("build", ["parity", "parity_cpu_form", "fast", "p64", "f64"])
def test_noalloc_equals_context_entry_points(mpcf, build):
    flags = builds(mpcf)[build]
    for k, (nx, ny, nz) in enumerate(SHAPES):
        s = states.random_smooth(nx, ny, nz, seed=31 + k)
        h, cfl = 1.0 / nx, 0.3
        for dt_max in (1e30, 1e-5):                  # dt from the speed of sound, then clamped
            say(build, "call", k, (nx, ny, nz), "dt_max", dt_max)
            sos, dt, bad = noalloc(mpcf, s, h, cfl, dt_max, flags, pad=5 * k)
            sos32, sos64, dref = reference(mpcf, s, h, cfl, dt_max, flags)
            assert bad == 0
            assert sos == sos64, (build, k, sos, sos64)
            if sos32 is not None:
                assert np.float32(sos) == np.float32(sos32) and float(np.float32(sos32)) == sos
            assert dt == dref, (build, k, dt, dref)


@pytest.mark.parametrize("build", ["parity", "fast", "p64", "f64"])
def test_noalloc_flags_a_bad_cell(mpcf, build):
    flags = builds(mpcf)[build]
    for k, (nx, ny, nz) in enumerate(SHAPES):
        s = states.random_smooth(nx, ny, nz, seed=41 + k)
        s[0, nz // 2, ny // 3, nx // 4] = -s[0, nz // 2, ny // 3, nx // 4]   # rho < 0 in one cell
        say(build, "bad cell, call", k, (nx, ny, nz))
        _, _, bad = noalloc(mpcf, s, 1.0 / nx, 0.3, 1e30, flags, pad=0)
        assert bad != 0
        # and the next call on a healthy state is not flagged (the array is cleared)
        s[0, nz // 2, ny // 3, nx // 4] = -s[0, nz // 2, ny // 3, nx // 4]
        _, _, bad = noalloc(mpcf, s, 1.0 / nx, 0.3, 1e30, flags, pad=0)
        assert bad == 0
        nz_, ny_, nx_ = s.shape[1:]
        with mpcf.Block(nx_, ny_, nz_, 1.0 / nx, **block_kwargs(mpcf, flags)) as b:
            s[0, nz // 2, ny // 3, nx // 4] = -s[0, nz // 2, ny // 3, nx // 4]
            b.data = s
            with pytest.raises(mpcf.MpcfError) as e:
                b.max_sos()
            assert e.value.code == mpcf.ENUMERIC
