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
