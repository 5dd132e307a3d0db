"""The largest block of the benchmark's shape that one MI355X holds with room
to spare: 1024^3 cells, 143 GB of HBM, 4.6 GB per field (byte offsets past
2^32), element indices up to 1.15e9, stepped over eight z-slabs on two
streams.

A state that varies along one axis only steps, line for line and bit for
bit, like a thin block (that axis 1024 long, the other two 8 cells) holding
the same line: every cell of both sees the same operands in the same order
(the flux differences of the other two directions are computed from equal
values in both blocks), so any indexing slip at large offsets, in the x+y
tiles or across the slab boundaries, breaks the equality.  The thin block
runs whole-block, the large one pipelined.  Skipped where the device cannot
hold the block.
"""
import sys

import numpy as np
import pytest

import states

pytestmark = pytest.mark.gpu

N = 1024


def say(*a):
    print("[gpu-large]", *a, flush=True)
    sys.stdout.flush()


def line_state(nx, ny, nz, axis):
    """(7, nz, ny, nx) binary32: water | air interface at the middle of the
    axis, a pressure step, a velocity along the axis; uniform across it."""
    s = (np.arange(N) + 0.5) / N
    water = s < 0.5
    rho = np.where(water, 1000.0, 1.0) * (1.0 + 0.05 * np.sin(6 * np.pi * s))
    vel = 0.4 * np.cos(2 * np.pi * s)
    p = 1.0 + 4.0 * (s < 0.3)
    Gw, Pw = states.gp(*states.WATER)
    Ga, Pa = states.gp(*states.AIR)
    G = np.where(water, Gw, Ga)
    P = np.where(water, Pw, Pa)
    zero = np.zeros_like(s)
    vels = [zero, zero, zero]
    vels[axis] = vel
    shape = [1, 1, 1]                      # (z, y, x)
    shape[2 - axis] = N
    c = states.conserved(*(np.reshape(q, shape) for q in (rho, *vels, p, G, P)))
    return np.ascontiguousarray(np.broadcast_to(c, (7, nz, ny, nx)))


@pytest.mark.parametrize("axis", [0, 2], ids=["x", "z"])
def test_largest_block_line_states(mpcf, axis):
    thin = [8, 8, 8]
    thin[axis] = N
    h = 1.0 / N
    say("thin block", thin)
    with mpcf.Block(*thin, h, fast=True) as b:
        b.data = line_state(*thin, axis)
        dts = [b.step(0.3) for _ in range(2)]
        ref = b.data
    say(f"{N}^3 block, axis {axis}")
    try:
        big = mpcf.Block(N, N, N, h, fast=True)
    except mpcf.MpcfError as e:
        pytest.skip(f"{N}^3 does not fit on this device: {e}")
    try:
        big.data = line_state(N, N, N, axis)
        got_dts = [big.step(0.3) for _ in range(2)]
        say("dt", got_dts, "thin", dts)
        assert got_dts == dts
        got = big.data
    finally:
        big.close()
    line = ref[:, :1, :1, :] if axis == 0 else ref[:, :, :1, :1]
    for q in range(7):
        same = np.array_equal(got[q], np.broadcast_to(line[q], got[q].shape))
        assert same, f"field {q} differs from the thin block's line"
