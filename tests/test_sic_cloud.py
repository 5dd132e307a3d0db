"""The shock-induced cloud collapse input (Sim_SICCloudMPI.cpp:43-357,
Sim_SICCloudMPI.h:148-181), pinned on the CPU:

* the post-shock state satisfies the Rankine-Hugoniot jump conditions of a
  stiffened gas (mass, momentum, energy) across a shock moving at u0 + c0 M --
  a wrong term in any of Johnsen's relations breaks one of them;
* the 27-point Simpson initial condition equals an independent numpy
  restatement of set_IC + _ic_quad;
* a sphere's volume is recovered from the averaged indicator;
* cloud_config.dat / cloud.dat are read as the reference reads them.
"""
import numpy as np
import pytest


def enthalpy(p, rho, g, pc):
    # stiffened gas: e = (p + g pc) / ((g - 1) rho), h = e + p / rho
    return g * (p + pc) / ((g - 1.0) * rho)


@pytest.mark.parametrize("psi,p0,u0,g,pc", [(40, 1, 0, 6.59, 4096), (353, 1, 0, 6.59, 4096),
                                            (10, 100, 2.5, 6.59, 4096), (3, 1, 0, 1.4, 0), (25, 0.5, -1, 4.4, 600)])
def test_post_shock_satisfies_rankine_hugoniot(mpcf, psi, p0, u0, g, pc):
    prm = mpcf.sic_params(p0=p0, u0=u0, pressure_ratio=psi, g1=g, pc1=pc)
    rho1, u1, p1, mach, c0 = mpcf.sic_post_shock(prm)
    rho0 = prm.rho0
    assert p1 == pytest.approx(psi * p0, rel=1e-15)
    assert c0 == pytest.approx(np.sqrt(g * (p0 + pc) / rho0), rel=1e-15)
    W = u0 + c0 * mach                          # shock speed
    m0, m1 = rho0 * (u0 - W), rho1 * (u1 - W)   # mass flux through the shock
    assert m1 == pytest.approx(m0, rel=1e-12)
    assert rho1 * (u1 - W) ** 2 + p1 == pytest.approx(rho0 * (u0 - W) ** 2 + p0, rel=1e-12)
    h0, h1 = enthalpy(p0, rho0, g, pc), enthalpy(p1, rho1, g, pc)
    assert h1 + 0.5 * (u1 - W) ** 2 == pytest.approx(h0 + 0.5 * (u0 - W) ** 2, rel=1e-12)
    if psi > 1:
        assert rho1 > rho0 and u1 > u0 and mach > 1


def restated_ic(n3, h, prm, shapes):
    """numpy restatement of set_IC (Sim_SICCloudMPI.cpp:247-276) averaged
    with the 27-point Simpson rule of integral() (:303-331)."""
    nx, ny, nz = n3
    rho0, u0, p0 = prm.rho0, prm.u0, prm.p0
    g1, g2, pc1, pc2 = prm.g1, prm.g2, prm.pc1, prm.pc2
    psi = prm.pressure_ratio
    c0 = np.sqrt(g1 * (p0 + pc1) / rho0)
    M = np.sqrt((g1 + 1) / (2 * g1) * (psi - 1) * p0 / (p0 + pc1) + 1)
    u1 = u0 + c0 * (psi - 1) * p0 / (g1 * (p0 + pc1) * M)
    t1, t2 = (g1 + 1) / (g1 - 1), (psi * p0 + pc1) / (p0 + pc1)
    rho1, p1 = rho0 * (t1 * t2 + 1) / (t1 + t2), psi * p0
    n = np.array(prm.shock_n[:]) / np.linalg.norm(prm.shock_n[:])
    S = np.array(prm.shock_s[:])
    z, y, x = np.meshgrid(h * (np.arange(nz) + 0.5), h * (np.arange(ny) + 0.5), h * (np.arange(nx) + 0.5),
                          indexing="ij")
    w = {-1: 1 / 6, 0: 2 / 3, 1: 1 / 6}
    out = np.zeros((7,) + x.shape)
    for jz in (-1, 0, 1):
        for jy in (-1, 0, 1):
            for jx in (-1, 0, 1):
                X, Y, Z = x + jx * h / 2, y + jy * h / 2, z + jz * h / 2
                shock = (n[0] * (X - S[0]) + n[1] * (Y - S[1]) + n[2] * (Z - S[2]) <= 0).astype(float)
                d = np.full(X.shape, np.inf)
                for cx, cy, cz, r in shapes:
                    d = np.minimum(d, np.sqrt((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) - r)
                if prm.eps > 0:
                    a = np.pi * np.clip((d + prm.eps) / (2 * prm.eps), 0, 1)
                    b = 0.5 + 0.5 * np.cos(a)
                else:
                    b = (d <= 0).astype(float)
                rho = shock * rho1 + (1 - shock) * (prm.rhoB * b + rho0 * (1 - b))
                u = shock * u1 + (1 - shock) * (prm.uB * b + u0 * (1 - b))
                p = shock * p1 + (1 - shock) * (prm.pB * b + p0 * (1 - b))
                G = 1 / (g1 - 1) * (1 - b) + 1 / (g2 - 1) * b
                P = g1 * pc1 / (g1 - 1) * (1 - b) + g2 * pc2 / (g2 - 1) * b
                wt = w[jx] * w[jy] * w[jz]
                out += wt * np.stack([rho, rho * u, 0 * rho, 0 * rho, p * G + P + 0.5 * rho * u * u, G, P])
    return out


@pytest.mark.parametrize("eps", [0.0, 0.06])
def test_ic_matches_numpy_restatement(mpcf, eps):
    n3, h = (20, 16, 12), 1.0 / 20
    shapes = np.array([[0.45, 0.4, 0.3, 0.17], [0.75, 0.5, 0.3, 0.09]])
    prm = mpcf.sic_params(shock_s=(0.2, 0, 0), shock_n=(1.0, 0.3, 0.0), eps=eps, pressure_ratio=20)
    f = mpcf.ic_sic_cloud(*n3, h=h, params=prm, shapes=shapes)
    ref = restated_ic(n3, h, prm, shapes)
    for q in range(7):
        scale = np.abs(ref[q]).max() or 1.0
        np.testing.assert_allclose(f[q], ref[q], rtol=0, atol=2e-7 * scale, err_msg=f"field {q}")


def test_uniform_liquid_without_shapes_or_shock(mpcf):
    prm = mpcf.sic_params(shock_s=(-1.0, 0, 0))          # the whole block is ahead of the shock
    f = mpcf.ic_sic_cloud(8, 6, 5, h=0.1, params=prm)
    G, P = 1 / 5.59, 6.59 * 4096 / 5.59
    expect = np.array([1000.0, 0, 0, 0, 1.0 * G + P, G, P], np.float32)
    for q in range(7):
        assert np.all(f[q] == expect[q]), q


def test_sphere_volume_from_averaged_indicator(mpcf):
    n, R = 48, 0.25
    prm = mpcf.sic_params(shock_s=(-1.0, 0, 0))
    f = mpcf.ic_sic_cloud(n, n, n, params=prm, shapes=np.array([[0.5, 0.5, 0.5, R]]))
    g_liq, g_gas = 1 / 5.59, 1 / 0.4
    frac = (f[5].astype(np.float64) - g_liq) / (g_gas - g_liq)
    vol = frac.sum() / n ** 3
    assert vol == pytest.approx(4 / 3 * np.pi * R ** 3, rel=5e-3)


def test_read_cloud_files(mpcf, tmp_path):
    cfg = tmp_path / "cloud_config.dat"
    cfg.write_text("3 1\n0.01 0.05\n0.1 0.2 0.3\n0.9 0.8 0.7\n")
    d = mpcf.read_cloud_config(cfg)
    assert d["n_shapes"] == 3 and d["n_small"] == 1 and d["min_rad"] == 0.01 and d["max_rad"] == 0.05
    assert d["seed_s"] == (0.1, 0.2, 0.3) and d["seed_e"] == (0.9, 0.8, 0.7) and d["n_sensors"] == 0
    cfg.write_text("3 1\n0.01 0.05\n0.1 0.2 0.3\n0.9 0.8 0.7\n4\n")
    assert mpcf.read_cloud_config(cfg)["n_sensors"] == 4
    cl = tmp_path / "cloud.dat"
    cl.write_text("0 0.5 0.5 0.5 0.1\n1 0.25 0.3 0.35 0.05\n\n2 0.7 0.7 0.7 0.02\n")
    sh = mpcf.read_cloud(cl)
    np.testing.assert_array_equal(sh, [[0.5, 0.5, 0.5, 0.1], [0.25, 0.3, 0.35, 0.05], [0.7, 0.7, 0.7, 0.02]])
    with pytest.raises(mpcf.MpcfError):
        mpcf.read_cloud(tmp_path / "missing.dat")
    cfg.write_text("3 1\n0.01\n")
    with pytest.raises(mpcf.MpcfError):
        mpcf.read_cloud_config(cfg)
