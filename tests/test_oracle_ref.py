"""Pins the oracle's ghost fills and CPU speed of sound against the
reference's own sources compiled by oracle/Makefile (target `ref`):
source/BoundaryConditions.h (applyBC_absorbing / applyBC_reflecting with the
reference's ghost maps, Types.h:42-69) and source/MaxSpeedOfSound.cpp
(MaxSpeedOfSound_CPP::compute).  Skipped where oracle/_ref was not built."""
import ctypes as C

import numpy as np
import pytest

import states


@pytest.fixture(scope="module")
def ref(oracle):
    r = oracle.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref/libref.so not built (needs the reference tree)")
    return r


def ref_sizes(ref):
    n = (C.c_int * 3)()
    ref.ref_sizes(C.cast(n, C.c_void_p))
    return tuple(n)


def test_ghost_fills_match_reference(oracle, ref):
    nx, ny, nz = ref_sizes(ref)
    s = states.random_smooth(nx, ny, nz, seed=21)
    for kind in (0, 1):
        b = oracle.Block(nx, ny, nz)
        b.data = s
        b.apply_bc([kind] * 6)
        for face in range(6):
            want = b.get_halo(face)
            got = np.zeros_like(want)
            arr = np.ascontiguousarray(s)
            ref.ref_apply_bc(face, kind, arr.ctypes.data_as(C.c_void_p), got.ctypes.data_as(C.c_void_p))
            assert np.array_equal(got, want), (kind, face)


def test_reflecting_flips_only_the_normal_momentum(oracle, ref):
    nx, ny, nz = ref_sizes(ref)
    s = states.random_smooth(nx, ny, nz, seed=22)
    b = oracle.Block(nx, ny, nz)
    b.data = s
    b.apply_bc([0] * 6)
    absorbing = [b.get_halo(f) for f in range(6)]
    b.apply_bc([1] * 6)
    for face in range(6):
        h = b.get_halo(face)
        d = face // 2
        # the layer next to the boundary mirrors the boundary cell: same |value|
        assert np.array_equal(np.abs(h[1 + d]).sum() > 0, True)
        for q in range(7):
            if q == 1 + d:
                continue
            assert np.all(np.isfinite(h[q]))
        assert not np.array_equal(h, absorbing[face])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cpu_speed_of_sound_matches_reference(oracle, ref, seed):
    nx, ny, nz = ref_sizes(ref)
    s = np.ascontiguousarray(states.random_smooth(nx, ny, nz, seed=seed))
    b = oracle.Block(nx, ny, nz)
    b.data = s
    got = np.float32(b.max_sos(cpu_form=True))
    want = np.float32(ref.ref_max_sos_cpp(s.ctypes.data_as(C.c_void_p)))
    assert got == want


def test_speed_of_sound_closed_form(oracle):
    # uniform state: c + max|u_i| with c^2 = gamma (p + pc) / rho
    for (gamma, pc), rho, p, vel in (((1.4, 0.0), 1.0, 1.0, (0.3, -0.7, 0.1)),
                                     ((6.59, 4096.0), 1000.0, 100.0, (2.0, 0.5, -3.0))):
        s = states.uniform(8, 8, 8, rho=rho, u=vel[0], v=vel[1], w=vel[2], p=p, phase=(gamma, pc))
        b = oracle.Block(8, 8, 8)
        b.data = s
        expect = np.sqrt(gamma * (p + pc) / rho) + max(abs(x) for x in vel)
        for form in (False, True):
            assert abs(b.max_sos(form) - expect) <= 2e-5 * expect
