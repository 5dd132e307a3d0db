"""CPU checks of the product's per-cell scheme (mpcf_scheme.h compiled for the
host by the library): the cell-organised reconstruction (shared smoothness
indicators, one reconstruction per cell and direction) and the shared-term
HLLC must equal the oracle's face-organised restatement of the reference bit
for bit.  No GPU involved."""
import numpy as np
import pytest

import states
from test_gpu_parity import weno_windows


def test_host_weno_equals_oracle(mpcf, oracle):
    w = weno_windows(20000, 3)
    m, p = mpcf.host_weno(w)
    rm, rp = oracle.weno(w)
    assert np.array_equal(m, rm)
    assert np.array_equal(p, rp)


def test_host_weno_fast_close_to_oracle(mpcf, oracle):
    """The FAST reconstruction (indicators on first differences, weights
    applied to third differences about the central candidate, median-of-three
    clip), compiled for the host with exact division, against the oracle's
    reference-arithmetic WENO on noise, smooth, jump, near-constant, constant
    and linear windows: every state within 1e-5 of its window's largest
    magnitude.  Rounding alone stays near 1e-6 (noise windows; near-constant
    large windows differ by a few ulp, where the reference's expanded
    quadratics cancel), while a wrong weight, candidate or clip moves a state
    by a sizeable fraction of the window."""
    w = weno_windows(20000, 5)
    m, p = mpcf.host_weno(w, fast=True)
    rm, rp = oracle.weno(w)
    scale = np.abs(w.astype(np.float64)).max(axis=1)
    for got, ref in ((m, rm), (p, rp)):
        err = np.abs(got.astype(np.float64) - ref.astype(np.float64)) / scale
        assert err.max() < 1e-5, err.max()
    # constant windows reconstruct exactly
    k = np.all(w == w[:, :1], axis=1)
    assert k.any() and np.array_equal(m[k], w[k, 2]) and np.array_equal(p[k], w[k, 2])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_host_face_equals_oracle(mpcf, oracle, d):
    st = states.random_stencils(30000, seed=40 + d)
    assert np.array_equal(mpcf.host_face(st, d), oracle.face_batch(st, d))


def test_flop_count_is_consistent(mpcf):
    f = mpcf.flop_count()
    # the three sweeps differ only in the emit arithmetic of their pass
    assert 900 < f["sweep_x"] < 1200 and 900 < f["sweep_y"] < 1200 and 900 < f["sweep_z"] < 1200
    assert f["cell_stage"] == f["sweep_x"] + f["sweep_y"] + f["sweep_z"]
    assert f["cell_step"] == 3 * f["cell_stage"] + f["sos_cell"]
    total = f["add"] + f["mul"] + f["div"] + f["sqrt"] + f["dadd"] + f["dmul"] + f["ddiv"]
    assert total == f["cell_stage"]
    # per cell-stage: 3 directions x 7 quantities x 3 binary64 reciprocals
    assert f["ddiv"] == 63
    ff = mpcf.flop_count(fast=True)
    assert ff["ddiv"] == 0 and ff["dmul"] == 0
