"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/gen_golden.py): the oracle must keep reproducing them on the CPU,
and the HIP path (bit-exact build) must reproduce them on the GPU."""
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_oracle_reproduces_rhs_fixture(oracle):
    f = load("rhs_mixed_bc.npz")
    nz, ny, nx = f["u"].shape[1:]
    b = oracle.Block(nx, ny, nz)
    b.data = f["u"]
    b.tmp = f["R0"]
    b.apply_bc(list(f["bc"]))
    b.rhs(f["a"], f["dtinvh"])
    assert np.array_equal(b.tmp, f["R"])


@pytest.mark.parametrize("name", ["steps_tube.npz", "steps_smooth_reflecting.npz"])
def test_oracle_reproduces_step_fixtures(oracle, name):
    f = load(name)
    nz, ny, nx = f["u0"].shape[1:]
    b = oracle.Block(nx, ny, nz)
    b.data = f["u0"]
    dts = [b.step(float(f["h"]), float(f["cfl"]), 1e30, list(f["bc"]))[0] for _ in range(len(f["dts"]))]
    assert dts == list(f["dts"])
    assert np.array_equal(b.data, f["u"]) and np.array_equal(b.tmp, f["R"])


def test_oracle_reproduces_pointwise_fixture(oracle):
    f = load("pointwise.npz")
    m, p = oracle.weno(f["weno_in"])
    assert np.array_equal(m, f["weno_minus"]) and np.array_equal(p, f["weno_plus"])
    for d in range(3):
        assert np.array_equal(oracle.face_batch(f["face_in"], d), f["face_out"][d])


def test_host_scheme_reproduces_pointwise_fixture(mpcf):
    f = load("pointwise.npz")
    m, p = mpcf.host_weno(f["weno_in"])
    assert np.array_equal(m, f["weno_minus"]) and np.array_equal(p, f["weno_plus"])
    for d in range(3):
        assert np.array_equal(mpcf.host_face(f["face_in"], d), f["face_out"][d])


@pytest.mark.gpu
def test_gpu_reproduces_rhs_fixture(mpcf):
    f = load("rhs_mixed_bc.npz")
    nz, ny, nx = f["u"].shape[1:]
    print("[golden] rhs fixture", (nx, ny, nz), flush=True)
    with mpcf.Block(nx, ny, nz, 1.0 / max(nx, ny, nz), bc=f["bc"]) as b:
        b.data = f["u"]
        b.tmp = f["R0"]
        b.load_ghosts()
        b.convection(f["a"], f["dtinvh"])
        assert np.array_equal(b.tmp, f["R"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["steps_tube.npz", "steps_smooth_reflecting.npz"])
def test_gpu_reproduces_step_fixtures(mpcf, name):
    f = load(name)
    nz, ny, nx = f["u0"].shape[1:]
    print("[golden] step fixture", name, flush=True)
    with mpcf.Block(nx, ny, nz, float(f["h"]), bc=f["bc"]) as b:
        b.data = f["u0"]
        dts = [b.step(float(f["cfl"])) for _ in range(len(f["dts"]))]
        assert dts == list(f["dts"])
        assert np.array_equal(b.data, f["u"]) and np.array_equal(b.tmp, f["R"])


@pytest.mark.gpu
def test_gpu_reproduces_pointwise_fixture(mpcf):
    f = load("pointwise.npz")
    m, p = mpcf.test_weno(f["weno_in"])
    assert np.array_equal(m, f["weno_minus"]) and np.array_equal(p, f["weno_plus"])
    for d in range(3):
        assert np.array_equal(mpcf.test_face(f["face_in"], d), f["face_out"][d])
