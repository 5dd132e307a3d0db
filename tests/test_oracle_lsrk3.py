"""Pins the oracle's LSRK3 step (source/LSRK3_IntegratorMPI.cpp:27-62):
third-order convergence in time of the 2N-storage scheme on a smooth
advection problem at fixed grid (binary64 instantiation, so round-off does
not mask the temporal error).  A wrong coefficient or stage order drops the
observed order to 1 or 2."""
import numpy as np

import states


def run(oracle, s, T, m):
    nz, ny, nx = s.shape[1:]
    b = oracle.Block(nx, ny, nz, dtype=np.float64)
    b.data = s.astype(np.float64)
    t = 0.0
    for _ in range(m):
        dt, _ = b.step(1.0 / nx, 50.0, T / m, [0] * 6)   # CFL large: dt_max = T/m is used
        t += dt
    assert abs(t - T) < 1e-12
    return b.data


def test_third_order_in_time(oracle):
    nx, ny, nz = 32, 5, 5
    x = (np.arange(nx) + 0.5) / nx
    rho = (1.0 + 0.2 * np.sin(2 * np.pi * x))[None, None, :] * np.ones((nz, ny, nx))
    one = np.ones_like(rho)
    G, P = states.gp(*states.AIR)
    s = states.conserved(rho, 1.0 * one, 0 * one, 0 * one, 1.0 * one, G * one, P * one, dtype=np.float64)
    T = 0.02
    ref = run(oracle, s, T, 256)
    errs = [np.abs(run(oracle, s, T, m) - ref).max() for m in (4, 8, 16)]
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(2.6 < o < 3.4 for o in orders), (errs, orders)


def test_linear_amplification_factor():
    """The 2N-storage coefficients (a, b) = ((0,-17/32,-32/27), (1/4,8/9,3/4))
    applied to du/dt = lambda u reproduce 1 + z + z^2/2 + z^3/6 (z = lambda dt):
    the stability polynomial of every three-stage third-order RK method."""
    A = (0.0, -17.0 / 32, -32.0 / 27)
    B = (1.0 / 4, 8.0 / 9, 3.0 / 4)
    for z in (-0.1, 0.3, -1.7, 0.05 + 0.0j, 0.8j, -1.0 + 0.5j):
        u, R = 1.0, 0.0
        for a, b in zip(A, B):
            R = a * R + z * u
            u = u + b * R
        assert abs(u - (1 + z + z * z / 2 + z ** 3 / 6)) < 1e-12
