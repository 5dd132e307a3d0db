"""Generate tests/golden/*.npz: seeded inputs and the CPU oracle's outputs.

    python tests/golden/gen_golden.py

The reference ships no golden vectors or fixtures, so these are produced by
the oracle (oracle/), whose functions are pinned independently by
tests/test_oracle_*.py.  Each file holds inputs and expected outputs only.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import pyoracle  # noqa: E402
import states    # noqa: E402


def rhs_case(name, shape, seed, bc, a, dtinvh):
    nx, ny, nz = shape
    s = states.random_smooth(nx, ny, nz, seed=seed)
    R0 = (states.random_smooth(nx, ny, nz, seed=seed + 100, two_phase=False) * np.float32(1e-3)).astype(np.float32)
    b = pyoracle.Block(nx, ny, nz)
    b.data = s
    b.tmp = R0
    b.apply_bc(bc)
    b.rhs(np.float32(a), np.float32(dtinvh))
    np.savez_compressed(os.path.join(HERE, name), u=s, R0=R0, bc=np.array(bc, np.int32),
                        a=np.float32(a), dtinvh=np.float32(dtinvh), R=b.tmp)


def step_case(name, s, bc, steps, cfl=0.3):
    nz, ny, nx = s.shape[1:]
    h = 1.0 / max(nx, ny, nz)
    b = pyoracle.Block(nx, ny, nz)
    b.data = s
    dts = [b.step(h, cfl, 1e30, bc)[0] for _ in range(steps)]
    np.savez_compressed(os.path.join(HERE, name), u0=s, bc=np.array(bc, np.int32), h=h, cfl=cfl,
                        dts=np.array(dts), u=b.data, R=b.tmp)


def main():
    rhs_case("rhs_mixed_bc.npz", (16, 12, 10), 1, [0, 1, 1, 0, 0, 1], -17.0 / 32, 0.021)
    step_case("steps_tube.npz", states.shock_tube(24, 8, 8, axis=0, seed=3), [0] * 6, 3)
    step_case("steps_smooth_reflecting.npz", states.random_smooth(12, 14, 10, seed=4), [1] * 6, 2)
    rng = np.random.default_rng(5)
    w = np.concatenate([rng.uniform(-1, 1, (600, 5)), 1000.0 * (np.arange(5)[None, :] >= rng.integers(1, 5, (600, 1))) + 1,
                        4828.79 + 1e-3 * rng.uniform(-1, 1, (600, 5))]).astype(np.float32)
    m, p = pyoracle.weno(w)
    st = states.random_stencils(900, seed=6)
    faces = np.stack([pyoracle.face_batch(st, d) for d in range(3)])
    np.savez_compressed(os.path.join(HERE, "pointwise.npz"), weno_in=w, weno_minus=m, weno_plus=p,
                        face_in=st, face_out=faces)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
