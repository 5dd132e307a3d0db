"""Dumps and checkpoints (HDF5Dumper_MPI.h:30-186, Types.h:188-280,
Sim_SteadyStateMPI.cpp:138-179): the XDMF wrapper, save.info, the device
streamers against a numpy restatement, and save -> restart continuing
bit for bit."""
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import states


def test_xdmf_wrapper(mpcf, tmp_path):
    p = tmp_path / "data_0007.xmf"
    mpcf.write_xdmf(p, "data_0007.raw", 12, 10, 8, 9, 0.125, 7, "Tensor")
    root = ET.parse(p).getroot()
    grid = root.find("Domain/Grid")
    assert grid.find("Time").get("Value") == "00007"
    topo = grid.find("Topology")
    assert topo.get("TopologyType") == "3DCORECTMesh" and topo.get("Dimensions") == "8 10 12"
    spacing = [d for d in grid.findall("Geometry/DataItem") if d.get("Name") == "Spacing"][0]
    assert [float(v) for v in spacing.text.split()] == [0.125] * 3
    att = grid.find("Attribute")
    assert att.get("Name") == "data" and att.get("AttributeType") == "Tensor"
    item = att.find("DataItem")
    assert item.get("Dimensions") == "8 10 12 9" and item.get("Format") == "Binary"
    assert item.get("Precision") == "4" and item.text.strip() == "data_0007.raw"


def test_save_info_round_trip(mpcf, tmp_path):
    p = tmp_path / "save.info"
    mpcf.write_save_info(p, 0.123456789012345678, 42, 3)
    lines = p.read_text().split()
    assert lines == ["0.123456789012346", "42", "3"]       # 15 significant digits
    t, st, fc = mpcf.read_save_info(p)
    assert (st, fc) == (42, 3) and t == pytest.approx(0.123456789012346, rel=1e-15)
    with pytest.raises(mpcf.MpcfError):
        mpcf.read_save_info(tmp_path / "none.info")


def say(*a):
    print("[gpu-io]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
def test_streamers_match_restatement(mpcf):
    s = states.random_smooth(13, 11, 9, seed=21)
    nz, ny, nx = s.shape[1:]
    say("streamers", (nx, ny, nz))
    with mpcf.Block(nx, ny, nz, 1.0 / 13) as b:
        b.data = s
        ten = b.stream_out(mpcf.STREAM_TENSOR)
        sav = b.stream_out(mpcf.STREAM_SAVE)
        sca = b.stream_out(mpcf.STREAM_SCALAR)
    r, ru, rv, rw, E, G, P = s
    assert np.array_equal(sav, np.moveaxis(s, 0, -1))
    assert np.array_equal(sca[..., 0], r)
    ss = (ru * ru + rv * rv) + rw * rw                      # binary32, the kernel's order
    p = ((E.astype(np.float64) - 0.5 * ss.astype(np.float64) / r.astype(np.float64) - P) / G).astype(np.float32)
    expect = np.stack([r, ru / r, rv / r, rw / r, p, G, P, 0 * r, 0 * r], axis=-1)
    assert np.array_equal(ten, expect)


@pytest.mark.gpu
def test_dump_files(mpcf, tmp_path):
    s = states.random_smooth(10, 9, 8, seed=22)
    with mpcf.Block(10, 9, 8, 0.1) as b:
        b.data = s
        b.dump(mpcf.STREAM_TENSOR, tmp_path, "data_0001", step=5)
        ten = b.stream_out(mpcf.STREAM_TENSOR)
    raw = np.fromfile(tmp_path / "data_0001.raw", dtype="<f4").reshape(8, 9, 10, 9)
    assert np.array_equal(raw, ten)
    root = ET.parse(tmp_path / "data_0001.xmf").getroot()
    assert root.find("Domain/Grid/Attribute/DataItem").get("Dimensions") == "8 9 10 9"


@pytest.mark.gpu
@pytest.mark.parametrize("fast", [False, True], ids=["bitexact", "fast"])
def test_save_restart_continues_bit_for_bit(mpcf, tmp_path, fast):
    s = states.random_smooth(18, 14, 12, seed=23)
    nz, ny, nx = s.shape[1:]
    h, bc = 1.0 / 18, [0, 1, 0, 0, 1, 1]
    say("save/restart", fast)
    with mpcf.Block(nx, ny, nz, h, bc=bc, fast=fast) as a:
        a.data = s
        t = sum(a.step(0.3) for _ in range(3))
        a.save(tmp_path, t, 3, 1)
        saved = a.data
        for _ in range(2):
            a.step(0.3)
        ua = a.data
    with mpcf.Block(nx, ny, nz, h, bc=bc, fast=fast) as b:
        t2, st, fc = b.restart(tmp_path)
        assert (st, fc) == (3, 1) and t2 == pytest.approx(t, rel=1e-14)
        assert np.array_equal(b.data, saved)
        for _ in range(2):
            b.step(0.3)                                    # stage 1 has a = 0: R restarts from zero
        assert np.array_equal(b.data, ua)
