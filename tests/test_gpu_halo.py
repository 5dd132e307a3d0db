"""Halo faces on one GPU: a domain split into blocks that exchange their face
halos before every stage (the device side of GPUlab's MPI halo exchange,
GPUlab.cpp:76-112,622-644) reproduces the undivided block bit for bit.

* mpcf_halo_export writes the three boundary cells in the reference's ghost
  map order (Types.h:31-69), checked against numpy slices of the block;
* blocks of 2x1x1, 2x2x2 and an uneven 3x1x2 split, with absorbing and
  reflecting outer faces, stepped through mpcf_max_sos / mpcf_process_all with
  an export -> import exchange per stage, equal the undivided block stepped
  the same way (same dt, same values) in the bit-exact, FMA-contracted and
  binary64 builds, and in the binary32 builds also the block stepped by
  mpcf_step;
* mpcf_step and mpcf_batch_add refuse halo faces.
"""
import ctypes
import sys

import numpy as np
import pytest

import states

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-halo]", *a, flush=True)
    sys.stdout.flush()


class DevBuf:
    """A plain device allocation through the HIP runtime (test plumbing)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            L = ctypes.CDLL("libamdhip64.so")
            L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            L.hipFree.argtypes = [ctypes.c_void_p]
            L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            cls._hip = L
        return cls._hip

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptr, self.nbytes = p.value, nbytes

    def to_host(self, dtype):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None


def face_cells(f, face):
    """The three boundary cells of `face` of f (7, NZ, NY, NX) in the ghost
    map order: x [NZ][NY][3], y [NZ][3][NX], z [3][NY][NX]."""
    d, side = divmod(face, 2)
    n = f.shape[3 - d]
    cells = [0, 1, 2] if side == 0 else [n - 3, n - 2, n - 1]
    if d == 0:
        return f[:, :, :, cells]
    if d == 1:
        return f[:, :, cells, :]
    return f[:, cells, :, :]


def test_halo_export_layout(mpcf):
    nx, ny, nz = 9, 7, 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.stack([q * 1000 + x + 10 * y + 100 * z for q in range(7)]).astype(np.float32)
    say("export layout", (nx, ny, nz))
    with mpcf.Block(nx, ny, nz, 1.0 / 9) as b:
        b.data = f
        for face in range(6):
            n = b.halo_count(face)
            d = face // 2
            area = [ny * nz, nx * nz, nx * ny][d]
            assert n == 7 * 3 * area
            buf = DevBuf(4 * n)
            try:
                b.halo_export(face, buf.ptr)
                b.synchronize()
                got = buf.to_host(np.float32)
            finally:
                buf.free()
            assert np.array_equal(got, face_cells(f, face).ravel()), face


def test_halo_refused_where_it_cannot_hold(mpcf):
    with mpcf.Block(8, 8, 8, 1.0 / 8, bc=[0, 2, 0, 0, 0, 0]) as b:
        b.data = states.uniform(8, 8, 8)
        with pytest.raises(mpcf.MpcfError, match="halo"):
            b.step(0.3)
        with pytest.raises(mpcf.MpcfError, match="face"):
            b.halo_count(6)
    bt = mpcf.Batch()
    try:
        with pytest.raises(mpcf.MpcfError, match="halo"):
            bt.add(states.uniform(8, 8, 8), 1.0 / 8, bc=[2, 0, 0, 0, 0, 0])
    finally:
        bt.close()


def split_edges(n, parts):
    return [round(n * i / parts) for i in range(parts + 1)]


def run_split(mpcf, f, parts, outer_bc, steps, cfl, kw, dtype):
    """Step f (7, N, N, N) as parts[0] x parts[1] x parts[2] blocks with a
    halo exchange before every stage; returns the assembled state and the dts."""
    N = f.shape[-1]
    h = 1.0 / N
    edges = [split_edges(N, p) for p in parts]
    blocks, bufs = {}, {}
    try:
        for i in range(parts[0]):
            for j in range(parts[1]):
                for k in range(parts[2]):
                    idx = (i, j, k)
                    lo = [edges[d][idx[d]] for d in range(3)]
                    hi = [edges[d][idx[d] + 1] for d in range(3)]
                    bc = []
                    for d in range(3):
                        for side in range(2):
                            edge = idx[d] == (0 if side == 0 else parts[d] - 1)
                            bc.append(outer_bc[2 * d + side] if edge else mpcf.BC_HALO)
                    b = mpcf.Block(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], h, bc=bc, **kw)
                    b.data = f[:, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
                    blocks[idx] = (b, lo, hi, bc)
                    for face in range(6):
                        if bc[face] == mpcf.BC_HALO:
                            bufs[idx, face] = DevBuf(np.dtype(dtype).itemsize * b.halo_count(face))
        dts = []
        for _ in range(steps):
            sos = max(b.max_sos() for b, *_ in blocks.values())
            dt = min(cfl * h / sos, 1e30)
            dts.append(dt)
            dtinvh = float(np.float32(dt / h))
            for s in range(3):
                for b, *_ in blocks.values():
                    b.synchronize()
                for idx, (b, lo, hi, bc) in blocks.items():
                    for face in range(6):
                        if bc[face] == mpcf.BC_HALO:
                            b.halo_export(face, bufs[idx, face].ptr)
                for b, *_ in blocks.values():
                    b.synchronize()
                for idx, (b, lo, hi, bc) in blocks.items():
                    for face in range(6):
                        if bc[face] == mpcf.BC_HALO:
                            d, side = divmod(face, 2)
                            nb = list(idx)
                            nb[d] += 1 if side else -1
                            # the neighbour's face towards this block
                            b.halo_import(face, bufs[tuple(nb), 2 * d + (1 - side)].ptr)
                    b.process_all(mpcf.LSRK3_A[s], mpcf.LSRK3_B[s], dtinvh)
        out = np.empty(f.shape, dtype)
        for b, lo, hi, bc in blocks.values():
            part = b.download_f64() if dtype == np.float64 else b.data
            out[:, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = part
        return out, dts
    finally:
        for b, *_ in blocks.values():
            b.close()
        for buf in bufs.values():
            buf.free()


CASES = [
    ("parity 2x1x1", (2, 1, 1), [0] * 6, {}),
    ("fast 2x2x2 mixed", (2, 2, 2), [1, 0, 0, 1, 1, 1], dict(fast=True)),
    ("parity 3x1x2 uneven reflecting", (3, 1, 2), [1, 1, 0, 0, 1, 0], {}),
    ("fp64 1x2x1", (1, 2, 1), [0, 1, 1, 0, 0, 0], dict(fp64=True)),
]


@pytest.mark.parametrize("name,parts,outer,kw", CASES, ids=[c[0] for c in CASES])
def test_split_domain_equals_single_block(mpcf, name, parts, outer, kw):
    N, steps, cfl = 32, 3, 0.3
    f = mpcf.ic_shock_bubble(N, N, N, 3)
    dtype = np.float64 if kw.get("fp64") else np.float32
    say(name, "undivided block, same host loop")
    ref, dts = run_split(mpcf, f, (1, 1, 1), outer, steps, cfl, kw, dtype)
    if dtype == np.float32:
        # binary32 builds: the host loop is mpcf_step itself (same dt, same
        # stage parameters).  A binary64 mpcf_step keeps sos and dt/h in
        # binary64 on the device, which mpcf_max_sos / mpcf_process_all (the
        # reference's binary32 signatures) do not carry.
        say(name, "undivided block, mpcf_step")
        with mpcf.Block(N, N, N, 1.0 / N, bc=outer, **kw) as b:
            b.data = f
            assert [b.step(cfl) for _ in range(steps)] == dts
            assert np.array_equal(b.data, ref)
    say(name, "split", parts)
    got, dts_split = run_split(mpcf, f, parts, outer, steps, cfl, kw, dtype)
    assert dts_split == dts
    if not np.array_equal(got, ref):
        bad = np.argwhere(got != ref)
        raise AssertionError(f"{name}: {len(bad)} values differ, first at {bad[:5].tolist()}")
