"""Initial conditions of the reference's other drivers, pinned on the CPU:
Sim_SodMPI (Sim_SodMPI.cpp:19-95), Sim_2DSBIMPI (Sim_2DSBIMPI.cpp:20-233)
with SimTools::getPostShockRatio (Types.h:101-109), Sim_SteadyStateMPI
(Sim_SteadyStateMPI.cpp:95-134) and Sim_StaticIC (Sim_StaticIC.cpp:76-97).

What pins them:
* the post-shock state satisfies the Rankine-Hugoniot jump conditions of an
  ideal gas across a shock running into gas at rest at M c0 (a wrong ratio,
  sign or sound speed breaks mass, momentum or energy);
* the stored states give back the primitive state they were set from
  (p = (E - |m|^2 / (2 rho) - P) / G, u = m / rho), so the energy and the
  momentum (rho u, not u) are the reference's;
* -config permutations, -reverse and the cylinder geometry are exact
  symmetries of the fields, checked by index permutation;
* the first-positive-bubble rule and the mollified disc area;
* a numpy float32 restatement of the 2D shock-bubble roundings (Real is
  binary32 in the reference's default build, Makefile.config:14).
"""
import ctypes
import math

import numpy as np
import pytest

# the binary32 cosine the C++ std::cos(float) resolves to (glibc cosf): the
# reference's heaviside_smooth takes the cosine of a Real angle
_LIBM = ctypes.CDLL("libm.so.6")
_LIBM.cosf.restype, _LIBM.cosf.argtypes = ctypes.c_float, [ctypes.c_float]
cosf = np.vectorize(lambda a: _LIBM.cosf(float(a)), otypes=[np.float32])


def ideal_h(p, rho, g):
    return g * p / ((g - 1.0) * rho)


@pytest.mark.parametrize("mach", [1.05, 1.22, 2.0, 5.0])
@pytest.mark.parametrize("g", [1.4, 1.67, 3.0])
def test_post_shock_rankine_hugoniot(mpcf, mach, g):
    rho0, p0 = 1.3, 0.7
    rho1, u1, p1 = mpcf.post_shock_ratio((rho0, 0.0, p0), mach, g)
    W = mach * math.sqrt(g * p0 / rho0)          # shock speed into gas at rest
    assert rho1 * (W - u1) == pytest.approx(rho0 * W, rel=1e-13)
    assert p1 + rho1 * (W - u1) ** 2 == pytest.approx(p0 + rho0 * W ** 2, rel=1e-13)
    assert ideal_h(p1, rho1, g) + 0.5 * (W - u1) ** 2 == pytest.approx(ideal_h(p0, rho0, g) + 0.5 * W ** 2,
                                                                        rel=1e-13)
    assert rho1 > rho0 and p1 > p0 and u1 > 0


def test_post_shock_binary32_and_quirks(mpcf):
    d = mpcf.post_shock_ratio((1.0, 0.0, 1.0), 1.22, 1.4)
    f = mpcf.post_shock_ratio((1.0, 0.0, 1.0), 1.22, 1.4, real32=True)
    for a, b in zip(d, f):
        assert np.float64(np.float32(b)) == b        # a binary32 value
        assert b == pytest.approx(a, rel=1e-6)
    # the pre-shock velocity does not enter (the reference adds no frame shift)
    assert mpcf.post_shock_ratio((1.0, 5.0, 1.0), 1.22, 1.4) == d
    # pc enters the speeds of sound only: the density and pressure ratios stay
    s = mpcf.post_shock_ratio((1.0, 0.0, 1.0), 1.22, 1.4, pc=3.0)
    assert s[0] == d[0] and s[2] == d[2] and s[1] != d[1]


def primitive(f):
    r, m, E, G, P = f[0].astype(np.float64), f[1:4].astype(np.float64), f[4], f[5], f[6]
    ke = 0.5 * (m ** 2).sum(0) / r
    return r, m / r, (E - ke - P) / G


def test_sod_defaults(mpcf):
    n = 16
    f = mpcf.ic_sod(n, 4, 3)
    left, right = f[:, :, :, : n // 2], f[:, :, :, n // 2:]    # centres h(i + 1/2) vs x0 = 0.5
    assert np.all(left[0] == 1.0) and np.all(right[0] == 0.125)
    assert np.all(left[4] == 2.5) and np.all(right[4] == np.float32(0.25))
    assert np.all(f[1:4] == 0) and np.all(f[5] == 2.5) and np.all(f[6] == 0)


@pytest.mark.parametrize("conf", ["x", "y", "z"])
def test_sod_states_config_reverse(mpcf, conf):
    kw = dict(rho1=2.0, u1=3.0, p1=5.0, rho2=0.5, u2=-1.5, p2=0.2, g1=4.4, pc1=6.0, g2=1.4, pc2=0.0, x0=0.3)
    n = (12, 10, 8)
    d = "xyz".index(conf)
    for rev in (0, 1):
        f = mpcf.ic_sod(*n, h=0.1, params=mpcf.sod_params(config=conf, reverse=rev, **kw))
        r, u, p = primitive(f)
        pos = 0.1 * (np.arange(n[d]) + 0.5)
        side1 = (pos > 0.3) if rev else (pos < 0.3)
        shape = [1, 1, 1]
        shape[2 - d] = n[d]
        s1 = np.broadcast_to(side1.reshape(shape), r.shape)
        np.testing.assert_allclose(r, np.where(s1, 2.0, 0.5), rtol=1e-7)
        np.testing.assert_allclose(u[d], np.where(s1, 3.0, -1.5), rtol=1e-6)
        for o in range(3):
            if o != d:
                assert np.all(u[o] == 0)
        np.testing.assert_allclose(p, np.where(s1, 5.0, 0.2), rtol=2e-5)
        np.testing.assert_allclose(f[5], np.where(s1, 1 / 3.4, 2.5), rtol=1e-7)
        np.testing.assert_allclose(f[6], np.where(s1, 4.4 * 6.0 / 3.4, 0.0), rtol=1e-7)


def test_sod_bad_arguments(mpcf):
    with pytest.raises(ValueError):
        mpcf.sod_params(config="w")
    p = mpcf.sod_params()
    p.dir = 3
    with pytest.raises(mpcf.MpcfError):
        mpcf.ic_sod(4, 4, 4, params=p)


def test_sbi_config_table(mpcf):
    want = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yx": (1, 0, 2), "yz": (1, 2, 0), "zx": (2, 0, 1), "zy": (2, 1, 0)}
    for k, v in want.items():
        assert mpcf.sbi_config(k) == v
        assert mpcf.sbi_bubble_file(v) == k + "bubbles.dat"
        bc = mpcf.sbi_bc(v)
        assert [bc[2 * a] for a in range(3)] == [int(a == v[1]) for a in range(3)]
        assert all(bc[2 * a] == bc[2 * a + 1] for a in range(3))
    with pytest.raises(mpcf.MpcfError):
        mpcf.sbi_config("xx")


def sbi_numpy(ext, h, prm, bubbles):
    """numpy float32 restatement of Sim_2DSBIMPI::_ic (Sim_2DSBIMPI.cpp:160-208)
    with Real = binary32, before the optional mirror."""
    f32 = np.float32
    d0, d1 = prm.dims[0], prm.dims[1]
    g, m = f32(prm.g1), f32(prm.mach)
    pre = [f32(prm.rho), f32(prm.u), f32(prm.p)]
    m2 = f32(m * m)
    gm1, gp1 = float(g) - 1.0, float(g) + 1.0
    Mpost = math.sqrt((float(m2) * gm1 + 2.0) / (2.0 * float(g) * float(m2) - gm1))
    rho1 = f32(gp1 * float(m2) / (gm1 * float(m2) + 2.0) * float(pre[0]))
    p1 = f32(1.0 / gp1 * (2.0 * float(g) * float(m2) - gm1) * float(pre[2]))
    c0 = float(m * np.sqrt(f32(g * (f32(prm.pc1) + pre[2]) / pre[0])))
    c1 = Mpost * float(np.sqrt(f32(g * (f32(prm.pc1) + p1) / rho1)))
    u1 = f32(c0 - c1)
    eps = prm.mollfactor * math.sqrt(3.0) * h
    nz, ny, nx = ext[2], ext[1], ext[0]
    pos = np.meshgrid(*(h * (np.arange(e) + 0.5) for e in (nz, ny, nx)), indexing="ij")[::-1]
    shock = np.where((pos[d0] - prm.x0).astype(np.float32) > 0, f32(0), f32(1))
    bubble = np.zeros(pos[0].shape, np.float32)
    done = np.zeros(pos[0].shape, bool)
    bub = bubbles.astype(np.float32)
    for b in bub:
        r = np.sqrt((pos[d0] - float(b[d0])) ** 2 + (pos[d1] - float(b[d1])) ** 2).astype(np.float32)
        phi = (r - b[3]).astype(np.float32)
        alpha = (math.pi * np.minimum(1.0, np.maximum(0.0, 0.5 * (phi.astype(np.float64) + eps) / eps))).astype(np.float32)
        val = (0.5 + 0.5 * cosf(alpha).astype(np.float64)).astype(np.float32)
        take = ~done
        bubble[take] = val[take]
        done |= val > 0
    nb, ns = (f32(1) - bubble), (f32(1) - shock)
    s, b64, nb64, ns64 = shock.astype(float), bubble.astype(float), nb.astype(float), ns.astype(float)
    R = (s * float(rho1) + ns64 * (prm.rhobubble * b64 + prm.rho * nb64)).astype(np.float32)
    M = ((s * float(u1) + ns64 * prm.u) * R.astype(float)).astype(np.float32)
    G1, G2 = prm.g1 - 1, prm.g2 - 1
    G = (1.0 / G1 * nb64 + 1.0 / G2 * b64).astype(np.float32)
    P = (prm.g1 * prm.pc1 / G1 * nb64 + prm.g2 * prm.pc2 / G2 * b64).astype(np.float32)
    pr = (s * float(p1) + ns64 * (prm.pbubble * b64 + prm.p * nb64)).astype(np.float32)
    ke = (0.5 * (M.astype(float) ** 2) / R.astype(float)).astype(np.float32)
    out = np.zeros((7,) + pos[0].shape, np.float32)
    out[0], out[1 + d0], out[5], out[6] = R, M, G, P
    out[4] = pr * G + P + ke
    return out


BUBBLES = np.array([[0.45, 0.5, 0.1, 0.2], [0.9, 0.25, 0.7, 0.1]])


@pytest.mark.parametrize("conf", ["xy", "zx"])
@pytest.mark.parametrize("extra", [{}, dict(mach=2.2, pc1=0.5, pc2=2.0, g2=2.5, mollfactor=2, u=0.3)])
def test_sbi_matches_float32_restatement(mpcf, conf, extra):
    prm = mpcf.sbi_params(config=conf, **extra)
    ext = [0, 0, 0]
    ext[prm.dims[0]], ext[prm.dims[1]], ext[prm.dims[2]] = 40, 36, 3
    h = 1.0 / 40
    b = BUBBLES.copy()
    # bubble coordinates (x, y, z) are read per axis: put the plane ones on dims
    b[:, [prm.dims[0], prm.dims[1], prm.dims[2]]] = BUBBLES[:, :3]
    got = mpcf.ic_sbi(*ext, h=h, params=prm, bubbles=b)
    want = sbi_numpy(ext, h, prm, b)
    assert np.array_equal(got, want), [int((got[q] != want[q]).sum()) for q in range(7)]


def test_sbi_plateaus_exact(mpcf):
    prm = mpcf.sbi_params(config="xy")
    n, h = 64, 1.0 / 64
    f = mpcf.ic_sbi(n, n, 2, h=h, params=prm, bubbles=np.array([[0.5, 0.5, 0.0, 0.2]]))
    rho1, u1, p1 = mpcf.post_shock_ratio((1.0, 0.0, 1.0), 1.22, 1.4, real32=True)
    R1 = np.float32(rho1)
    M1 = np.float32(u1 * float(R1))
    assert np.all(f[0, :, :, :6] == R1) and np.all(f[1, :, :, :6] == M1)          # x < x0 = 0.1
    np.testing.assert_allclose(primitive(f[:, :, :, :6])[2], p1, rtol=1e-6)
    assert np.all(f[2:4] == 0)
    far = f[:, :, :4, 10:]                                                        # y < 0.07: outside the disc
    assert np.all(far[0] == 1) and np.all(far[1] == 0) and np.all(far[4] == 2.5) and np.all(far[5] == 2.5)
    core = f[:, :, 28:36, 28:36]                                                  # |r| < 0.07 inside r = 0.2
    gb = np.float32(1.0 / 0.67)
    assert np.all(core[0] == np.float32(0.138)) and np.all(core[5] == gb) and np.all(core[4] == gb)
    assert np.all(f[6] == 0)
    assert np.array_equal(f[:, 0], f[:, 1])                                       # cylinders along z
    # the mollified indicator keeps the disc's area (symmetric smoothing)
    frac = (f[5, 0].astype(float) - 2.5) / (1 / 0.67 - 2.5)
    assert frac.sum() * h * h == pytest.approx(math.pi * 0.04, rel=5e-3)


def test_sbi_first_positive_bubble(mpcf):
    n, h = 32, 1.0 / 32
    prm = mpcf.sbi_params(config="xy", x0=-1.0)                   # no shocked cells
    # cell (16, 16) has its centre at (0.515625, 0.515625): in the smoothing
    # band of bubble A, deep inside bubble B
    A = [0.3, 0.515625, 0.0, 0.215625 - 0.01]
    B = [0.515625, 0.515625, 0.0, 0.3]
    fa = mpcf.ic_sbi(n, n, 1, h=h, params=prm, bubbles=np.array([A, B]))
    fb = mpcf.ic_sbi(n, n, 1, h=h, params=prm, bubbles=np.array([B, A]))
    eps = math.sqrt(3) * h
    ind = 0.5 + 0.5 * math.cos(math.pi * min(1, max(0, 0.5 * (0.01 + eps) / eps)))
    assert 0 < ind < 1
    want_g = 2.5 * (1 - ind) + ind / 0.67
    assert fa[5, 0, 16, 16] == pytest.approx(want_g, rel=1e-6)    # first positive: A's value
    assert fb[5, 0, 16, 16] == np.float32(1 / 0.67)               # B first: fully inside


@pytest.mark.parametrize("conf", ["xz", "yx", "yz", "zx", "zy"])
def test_sbi_config_is_a_permutation(mpcf, conf):
    A, B, Cn = 14, 12, 3
    h = 1.0 / 14
    base = mpcf.ic_sbi(A, B, Cn, h=h, params=mpcf.sbi_params(config="xy", x0=0.2),
                       bubbles=np.array([[0.55, 0.45, 0.0, 0.25]]))
    prm = mpcf.sbi_params(config=conf, x0=0.2)
    d = prm.dims
    ext, b = [0, 0, 0], [0.0, 0.0, 0.0, 0.25]
    ext[d[0]], ext[d[1]], ext[d[2]] = A, B, Cn
    b[d[0]], b[d[1]], b[d[2]] = 0.55, 0.45, 0.0
    got = mpcf.ic_sbi(*ext, h=h, params=prm, bubbles=np.array([b]))
    qmap = [0, 1 + d[0], 1 + d[1], 1 + d[2], 4, 5, 6]
    for i in range(A):
        for j in range(B):
            for k in range(Cn):
                idx = [0, 0, 0]
                idx[d[0]], idx[d[1]], idx[d[2]] = i, j, k
                assert np.array_equal(got[qmap, idx[2], idx[1], idx[0]], base[:, k, j, i])


@pytest.mark.parametrize("nx", [16, 15])
def test_sbi_reverse_mirrors(mpcf, nx):
    b = np.array([[0.3, 0.4, 0.0, 0.15]])
    f = mpcf.ic_sbi(nx, 10, 2, h=1 / 16, params=mpcf.sbi_params(config="xy", x0=0.12), bubbles=b)
    r = mpcf.ic_sbi(nx, 10, 2, h=1 / 16, params=mpcf.sbi_params(config="xy", x0=0.12, reverse=1), bubbles=b)
    mir = f[:, ::-1, ::-1, ::-1].copy()
    mir[1:4] *= -1
    half = nx // 2
    assert np.array_equal(r[:, :, :, :half], mir[:, :, :, :half])
    assert np.array_equal(r[:, :, :, nx - half:], mir[:, :, :, nx - half:])
    if nx % 2:                                    # the middle slab is left as it was
        assert np.array_equal(r[:, :, :, half], f[:, :, :, half])


def test_read_bubbles(mpcf, tmp_path):
    p = tmp_path / "xybubbles.dat"
    p.write_text("3\n0.1 0.2 0.3 0.04\n0.5 0.5 0.5 0.1\n  0.9 0.8\n0.7 0.05\n")
    b = mpcf.read_bubbles(p)
    assert b.shape == (3, 4) and b[2].tolist() == [0.9, 0.8, 0.7, 0.05]
    assert mpcf.read_bubbles(p, cap=1).shape == (1, 4)
    (tmp_path / "short.dat").write_text("2\n0.1 0.2 0.3 0.04\n")
    with pytest.raises(mpcf.MpcfError, match="short"):
        mpcf.read_bubbles(tmp_path / "short.dat")
    with pytest.raises(mpcf.MpcfError, match="cannot read"):
        mpcf.read_bubbles(tmp_path / "none.dat")


@pytest.mark.parametrize("dims", [(0, 1, 2), (2, 1, 0), (1, 0, 2)])
def test_static_index_order(mpcf, dims):
    ext = (5, 4, 3)
    f = mpcf.ic_static(*ext, dims=dims)
    idx = np.meshgrid(*(np.arange(e) for e in ext[::-1]), indexing="ij")[::-1]   # idx[a] on (z, y, x)
    want = idx[dims[0]] + ext[dims[0]] * (idx[dims[1]] + ext[dims[1]] * idx[dims[2]])
    for q in range(7):
        assert np.array_equal(f[q], want.astype(np.float32))
    with pytest.raises(mpcf.MpcfError):
        mpcf.ic_static(3, 3, 3, dims=(0, 0, 1))


def test_steady_state(mpcf):
    f = mpcf.ic_steady(3, 2, 2, rho=2.0, u=1.0, v=2.0, w=3.0, p=3.0, g=4.0, pc=5.0)
    r, u, p = primitive(f)
    np.testing.assert_allclose(r, 2.0)
    np.testing.assert_allclose(u, np.array([1.0, 2.0, 3.0])[:, None, None, None] * np.ones_like(r), rtol=1e-7)
    np.testing.assert_allclose(p, 3.0, rtol=1e-6)
    np.testing.assert_allclose(f[5], 1 / 3, rtol=1e-7)
    np.testing.assert_allclose(f[6], 4 * 5 / 3, rtol=1e-7)
    d = mpcf.ic_steady(2, 2, 2)
    assert d[:, 0, 0, 0].tolist() == [1, 0, 0, 0, 2.5, 2.5, 0]
