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
