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


@pytest.mark.parametrize("args,code,msg", [
    (["-sim", "sod", "-config", "w"], 2, "-config x|y|z"),
    (["-sim", "2DSBIMPI", "-config", "xy"], 1, "xybubbles.dat"),
    (["-sim", "2dsbi"], 1, "-config"),
    (["-sim", "StaticIC"], 1, "-config"),
    (["-sim", "steady", "-nsteps", "0"], 2, "needs -tend"),
])
def test_driver_case_argument_errors(tmp_path, args, code, msg):
    r = run(args + ["-n", "8"], cwd=tmp_path)
    assert r.returncode == code and msg in r.stderr, (r.returncode, r.stderr)


def oracle_run(oracle, f, n, bc, steps=None, tend=1e30, dumpinterval=0.0, nsteps=0):
    """The driver's loop (Sim_SteadyStateMPI.cpp:182-221) on the oracle: dt cut
    to land on the dump times, (float) t == (float) tnextdump."""
    ob = oracle.Block(n, n, n)
    ob.data = f
    t, done, tnext, dumps = 0.0, 0, dumpinterval, 1 if dumpinterval > 0 else 0
    while (steps is None or done < steps) and t < tend:
        dt_max = tend - t
        if dumpinterval > 0 and tnext - t < dt_max:
            dt_max = tnext - t
        t += ob.step(1.0 / n, 0.3, dt_max, bc)[0]
        done += 1
        if dumpinterval > 0 and np.float32(t) == np.float32(tnext):
            tnext += dumpinterval
            dumps += 1
        if nsteps and done == nsteps:
            break
    return t, done, ob.data, dumps + (1 if dumpinterval > 0 else 0)


@pytest.mark.gpu
def test_driver_sod_bitexact(tmp_path, mpcf, oracle):
    n = 32
    print("[driver] SodMPI -config y, 5 steps", flush=True)
    r = run(["-sim", "SodMPI", "-config", "y", "-rho1", "2", "-u1", "0.5", "-n", str(n), "-steps", "5",
             "-dump", "sod.bin"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    steps, t, u = read_dump(tmp_path / "sod.bin")
    f = mpcf.ic_sod(n, n, n, params=mpcf.sod_params(config="y", rho1=2.0, u1=0.5))
    tr, done, ur, _ = oracle_run(oracle, f, n, [0] * 6, steps=5)
    assert steps == done == 5 and t == tr
    assert np.array_equal(u, ur)


@pytest.mark.gpu
def test_driver_2dsbi_bitexact_reflecting(tmp_path, mpcf, oracle):
    n = 32
    (tmp_path / "yxbubbles.dat").write_text("2\n0.4 0.3 0.5 0.15\n0.6 0.7 0.5 0.1\n")
    print("[driver] 2DSBIMPI -config yx, 6 steps", flush=True)
    r = run(["-sim", "2DSBIMPI", "-config", "yx", "-n", str(n), "-steps", "6", "-dump", "sbi.bin"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    steps, t, u = read_dump(tmp_path / "sbi.bin")
    prm = mpcf.sbi_params(config="yx")
    f = mpcf.ic_sbi(n, n, n, params=prm, bubbles=mpcf.read_bubbles(tmp_path / "yxbubbles.dat"))
    tr, done, ur, _ = oracle_run(oracle, f, n, mpcf.sbi_bc(tuple(prm.dims)), steps=6)
    assert steps == done == 6 and t == tr
    assert np.array_equal(u, ur)


@pytest.mark.gpu
def test_driver_dumpinterval_loop(tmp_path, mpcf, oracle):
    """-tend / -dumpinterval / -nsteps 0: the reference's loop, dumps at the
    initial state, every interval and the end."""
    n = 32
    (tmp_path / "d").mkdir()
    print("[driver] SodMPI tend 0.05, dumpinterval 0.0125", flush=True)
    r = run(["-sim", "sod", "-n", str(n), "-tend", "0.05", "-dumpinterval", "0.0125", "-nsteps", "0",
             "-fpath", "d", "-dump", "end.bin"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    steps, t, u = read_dump(tmp_path / "end.bin")
    f = mpcf.ic_sod(n, n, n)
    tr, done, ur, ndumps = oracle_run(oracle, f, n, [0] * 6, tend=0.05, dumpinterval=0.0125)
    assert steps == done and t == tr and np.float32(t) == np.float32(0.05)
    assert np.array_equal(u, ur)
    assert ndumps == 6
    for k in range(ndumps):
        assert (tmp_path / "d" / f"data_{k:04d}.raw").stat().st_size == n ** 3 * 9 * 4
    assert not (tmp_path / "d" / f"data_{ndumps:04d}.raw").exists()


@pytest.mark.gpu
def test_driver_static_probe(tmp_path, mpcf):
    n = 8
    print("[driver] StaticIC -config zy", flush=True)
    r = run(["-sim", "StaticIC", "-config", "zy", "-n", str(n), "-fpath", ".", "-fname", "probe"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "probe_0000.raw", dtype="<f4").reshape(n, n, n, 9)
    assert np.array_equal(raw[..., 0], mpcf.ic_static(n, n, n, dims=(2, 1, 0))[0])
