"""Seeded test states for the 7-field stiffened-gas system (test inputs only).

Arrays are float32 (7, NZ, NY, NX) with fields rho, rho u, rho v, rho w, E,
G = 1/(gamma-1), P = gamma pc/(gamma-1) (source/NodeBlock.h:26-29), and
E = G p + P + rho |u|^2 / 2 (source/Sim/Sim_SteadyStateMPI.cpp:115-117).
"""
from __future__ import annotations

import numpy as np

AIR = (1.4, 0.0)
WATER = (6.59, 4096.0)


def conserved(rho, u, v, w, p, G, P, dtype=np.float32):
    E = G * p + P + 0.5 * rho * (u * u + v * v + w * w)
    out = np.stack([rho, rho * u, rho * v, rho * w, E, G, P]).astype(dtype)
    return np.ascontiguousarray(out)


def gp(gamma, pc):
    return 1.0 / (gamma - 1.0), gamma * pc / (gamma - 1.0)


def centres(nx, ny, nz):
    h = 1.0 / max(nx, ny, nz)
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) * h, (np.arange(ny) + 0.5) * h,
                          (np.arange(nx) + 0.5) * h, indexing="ij")
    return x, y, z, h


def random_smooth(nx, ny, nz, seed=0, amp=0.2, vel=0.5, two_phase=True):
    """Two-phase field with smooth random modes, a tanh bubble and noise."""
    rng = np.random.default_rng(seed)
    x, y, z, h = centres(nx, ny, nz)
    def modes(a):
        f = np.zeros_like(x)
        for _ in range(3):
            k = rng.integers(1, 4, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            f += a * np.sin(2 * np.pi * k[0] * x + ph[0]) * np.sin(2 * np.pi * k[1] * y + ph[1]) \
                * np.cos(2 * np.pi * k[2] * z + ph[2]) / 3
        return f
    if two_phase:
        c = rng.uniform(0.35, 0.65, size=3)
        r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
        phi = 0.5 * (1 + np.tanh((r - 0.25) / (1.5 * h)))   # 0 in bubble (air), 1 outside (water)
    else:
        phi = np.zeros_like(x)
    Ga, Pa = gp(*AIR)
    Gw, Pw = gp(*WATER)
    G = Ga + (Gw - Ga) * phi
    P = Pa + (Pw - Pa) * phi
    rho = (1.0 + modes(amp)) * (1 - phi) + 1000.0 * (1.0 + modes(amp)) * phi
    p = 1.0 + modes(amp) + 9.0 * phi
    u = vel * modes(1.0)
    v = vel * modes(1.0)
    w = vel * modes(1.0)
    rho = rho * (1 + rng.uniform(-1e-3, 1e-3, size=rho.shape))
    return conserved(rho, u, v, w, p, G, P)


def shock_tube(nx, ny, nz, axis=0, seed=0, noise=1e-3):
    """Air | water tube along `axis` (tanh over 2 cells, config-1 states)."""
    rng = np.random.default_rng(seed)
    x, y, z, h = centres(nx, ny, nz)
    s = (x, y, z)[axis]
    phi = 0.5 * (1 + np.tanh((s - 0.5) / (2 * h)))
    Ga, Pa = gp(*AIR)
    Gw, Pw = gp(*WATER)
    G = Ga + (Gw - Ga) * phi
    P = Pa + (Pw - Pa) * phi
    rho = 1.0 + (1000.0 - 1.0) * phi
    p = 1.0 + (0.1 - 1.0) * phi
    if noise:
        rho = rho * (1 + rng.uniform(-noise, noise, size=rho.shape))
    zero = np.zeros_like(rho)
    return conserved(rho, zero, zero, zero, p, G, P)


def uniform(nx, ny, nz, rho=1.0, u=0.0, v=0.0, w=0.0, p=1.0, phase=AIR):
    G, P = gp(*phase)
    one = np.ones((nz, ny, nx))
    return conserved(rho * one, u * one, v * one, w * one, p * one, G * one, P * one)


def random_stencils(n, seed=0, two_phase=True):
    """n x 7 x 6 conserved stencils (face i-1/2 of points i-3..i+2) spanning
    smooth, discontinuous and material-interface cases."""
    rng = np.random.default_rng(seed)
    kind = rng.integers(0, 4, size=n)
    t = np.arange(6)[None, :]
    base_rho = np.where(rng.random(n) < 0.5, 1.0, 1000.0)[:, None]
    jump = (t >= rng.integers(1, 6, size=n)[:, None]).astype(float)
    smooth = np.sin(rng.uniform(0, 1, (n, 1)) * t + rng.uniform(0, 6, (n, 1)))
    noise = rng.uniform(-1, 1, (n, 6))
    def pick(a_s, a_j, a_n):
        return np.where(kind[:, None] == 0, a_s * smooth,
               np.where(kind[:, None] == 1, a_j * jump,
               np.where(kind[:, None] == 2, a_n * noise, 0.0)))
    rho = base_rho * (1 + 0.3 * pick(1, 1, 1)) + (1000 - base_rho) * 0
    if two_phase:
        wphase = np.where(kind[:, None] == 3, jump, (base_rho > 10).astype(float) * np.ones((1, 6)))
        rho = np.where(kind[:, None] == 3, 1.0 + 999.0 * wphase, rho)
    else:
        wphase = np.zeros((n, 6))
    Ga, Pa = gp(*AIR)
    Gw, Pw = gp(*WATER)
    G = Ga + (Gw - Ga) * wphase
    P = Pa + (Pw - Pa) * wphase
    p = 1.0 + 0.5 * pick(1, 1, 1) + 5.0 * wphase
    u = rng.uniform(-1, 1, (n, 1)) + 0.3 * pick(1, 1, 1)
    v = rng.uniform(-1, 1, (n, 1)) + 0.3 * pick(1, 1, 1)
    w = rng.uniform(-1, 1, (n, 1)) + 0.3 * pick(1, 1, 1)
    st = conserved(rho, u, v, w, p, G, P)          # (7, n, 6)
    return np.ascontiguousarray(np.transpose(st, (1, 0, 2)))


def shock_bubble(nx, ny, nz):
    """Shock-bubble without noise (the config-2 state): water (rho 1000,
    p 1), an air bubble of radius 0.1 at the centre (rho 1, p 1), post-shock
    water (rho 1000, p 1e3, u 10) at x < 0.3, tanh over 1.5 cells.  The
    squared radius adds the y and z terms first, so the field is exactly
    symmetric under y <-> z (and under y -> 1 - y, z -> 1 - z)."""
    x, y, z, h = centres(nx, ny, nz)
    r = np.sqrt((x - 0.5) ** 2 + ((y - 0.5) ** 2 + (z - 0.5) ** 2))
    bubble = 0.5 * (1 - np.tanh((r - 0.1) / (1.5 * h)))
    shock = 0.5 * (1 - np.tanh((x - 0.3) / (1.5 * h)))
    Ga, Pa = gp(*AIR)
    Gw, Pw = gp(*WATER)
    rho = shock * 1000.0 + (1 - shock) * (1.0 * bubble + 1000.0 * (1 - bubble))
    u = shock * 10.0
    p = shock * 1000.0 + (1 - shock) * 1.0
    G = Gw * (1 - bubble) + Ga * bubble
    P = Pw * (1 - bubble) + Pa * bubble
    zero = np.zeros_like(rho)
    return conserved(rho, u, zero, zero, p, G, P)
