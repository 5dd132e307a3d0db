"""The C++ driver (csrc/mpcf_run.cpp over the facades in csrc/mpcf.hpp):
argument errors on the CPU, and on the GPU a full run of the config-1 tube
checked bit-for-bit against the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG

EXE = os.path.join(PKG, "mpcf_run")


def run(args, **kw):
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=120, **kw)


@pytest.mark.parametrize("args", [["-foo", "1"], ["-sim", "nope"], ["-n", "abc"], ["-mode", "quick"], ["-n"]])
def test_driver_rejects_bad_arguments(args):
    r = run(args)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)


def read_dump(path):
    with open(path, "rb") as f:
        assert f.read(8) == b"MPCF0001"
        nx, ny, nz, steps = np.frombuffer(f.read(16), dtype=np.int32)
        t = np.frombuffer(f.read(8), dtype=np.float64)[0]
        data = np.frombuffer(f.read(), dtype=np.float32).reshape(7, nz, ny, nx)
    return int(steps), float(t), data


@pytest.mark.gpu
def test_driver_tube_bitexact(tmp_path, mpcf, oracle):
    n, steps = 32, 6
    out = tmp_path / "tube.bin"
    print("[driver] mpcf_run tube", n, steps, flush=True)
    r = run(["-sim", "tube", "-n", str(n), "-steps", str(steps), "-seed", "5", "-dump", str(out)])
    assert r.returncode == 0, r.stderr
    got_steps, t, u = read_dump(out)
    assert got_steps == steps
    s = mpcf.ic_shock_tube(n, n, n, 0, 1e-3, 5)
    ob = oracle.Block(n, n, n)
    ob.data = s
    tref = 0.0
    for _ in range(steps):
        tref += ob.step(1.0 / n, 0.3, 1e30, [0] * 6)[0]
    assert t == tref
    assert np.array_equal(u, ob.data)


def write_cloud(d):
    (d / "cloud_config.dat").write_text("2 0\n0.05 0.12\n0.2 0.2 0.2\n0.8 0.8 0.8\n")
    (d / "cloud.dat").write_text("0 0.45 0.5 0.5 0.12\n1 0.7 0.45 0.55 0.07\n")


def test_driver_sic_needs_cloud_files(tmp_path):
    r = run(["-sim", "sic", "-n", "8", "-steps", "1"], cwd=tmp_path)
    assert r.returncode == 1 and "cloud_config.dat" in r.stderr


@pytest.mark.gpu
def test_driver_sic_dump_save_restart(tmp_path, mpcf):
    """-sim sic from cloud files (initial state = mpcf_ic_sic_cloud), tensorial
    dumps, and a checkpoint/restart that continues exactly like the
    uninterrupted run."""
    write_cloud(tmp_path)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    n = 24
    common = ["-sim", "sic", "-n", str(n), "-pressureratio", "20", "-shockSx", "0.15"]
    print("[driver] sic 6 steps with dumps", flush=True)
    r = run(common + ["-steps", "6", "-dumpperiod", "3", "-fpath", "a", "-dump", "full.bin"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "a" / "data_0000.raw").stat().st_size == n ** 3 * 9 * 4
    assert (tmp_path / "a" / "data_0001.xmf").exists()
    print("[driver] sic 4 steps + save, then restart for 2", flush=True)
    r = run(common + ["-steps", "4", "-saveperiod", "4", "-fpath", "b"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    r = run(common + ["-steps", "2", "-restart", "1", "-fpath", "b", "-dump", "resumed.bin"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    s_full, t_full, u_full = read_dump(tmp_path / "full.bin")
    s_res, t_res, u_res = read_dump(tmp_path / "resumed.bin")
    assert np.array_equal(u_res, u_full)
    assert t_res == pytest.approx(t_full, rel=1e-14)
    # the first dump is the state after step 3 of the same initial condition
    prm = mpcf.sic_params(pressure_ratio=20, shock_s=(0.15, 0, 0))
    f0 = mpcf.ic_sic_cloud(n, n, n, params=prm, shapes=mpcf.read_cloud(tmp_path / "cloud.dat"))
    with mpcf.Block(n, n, n, 1.0 / n) as b:
        b.data = f0
        for _ in range(3):
            b.step(0.3)
        ten = b.stream_out(mpcf.STREAM_TENSOR)
    raw = np.fromfile(tmp_path / "a" / "data_0000.raw", dtype="<f4").reshape(n, n, n, 9)
    assert np.array_equal(raw, ten)
