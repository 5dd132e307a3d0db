"""Pins the oracle's per-face computation (oracle face_flux: primitive
recovery, WENO5 x 14, Einfeldt/Batten speeds, HLLC, extraterm velocity;
reference source/GPU/GPUonly.cuh:802-1042, GPUkernels.cu:203-287):

* against an independent float64 implementation written from the textbook
  (Jiang-Shu WENO5 in the 13/12-1/4 form; Einfeldt's HLLE speeds; Batten's
  s*; Toro's branch form of HLLC, F*_K = F_K + S_K (U*_K - U_K), with G and P
  as density-weighted scalars and the face velocity as the HLLC flux of a unit
  scalar).  The reference's sigma-weighted form is algebraically the branch
  form, so any wrong sign, index, argument order or term fails this;
* closed forms: equal states give the physical flux (exactly when u_n = 0).
"""
import numpy as np
import pytest

import states
from test_oracle_weno import weno_js_textbook


def prim64(st):
    """st: (n,7,6) conserved -> dict of (n,6) float64 primitives, like the
    reference's per-point conversion."""
    s = st.astype(np.float64)
    r = s[:, 0]
    u, v, w = s[:, 1] / r, s[:, 2] / r, s[:, 3] / r
    G, P = s[:, 5], s[:, 6]
    p = (s[:, 4] - 0.5 * r * (u * u + v * v + w * w) - P) / G
    return dict(r=r, u=u, v=v, w=w, p=p, G=G, P=P)


def clip(val, win):
    lo = win[:, 1:4].min(axis=1)
    hi = win[:, 1:4].max(axis=1)
    return np.minimum(np.maximum(val, lo), hi)


def recon(q):
    """q: (n,6) at i-3..i+2 -> (left state, right state) at face i-1/2."""
    wl = q[:, 0:5]
    wr = q[:, 1:6]
    left = clip(weno_js_textbook(*wl.T), wl)
    right = clip(weno_js_textbook(*wr[:, ::-1].T), wr[:, ::-1])
    return left, right


def face_textbook(st, d):
    pr = prim64(st)
    L, R = {}, {}
    for k in pr:
        L[k], R[k] = recon(pr[k])
    comp = ["u", "v", "w"]
    n_ = comp[d]
    t1, t2 = [c for c in comp if c != n_]
    cL = np.sqrt(((L["p"] + L["P"]) / L["G"] + L["p"]) / L["r"])
    cR = np.sqrt(((R["p"] + R["P"]) / R["G"] + R["p"]) / R["r"])
    uL, uR = L[n_], R[n_]
    Rr = np.sqrt(R["r"] / L["r"])
    ubar = (uL + Rr * uR) / (1 + Rr)
    dbar = np.sqrt((cL ** 2 + Rr * cR ** 2) / (1 + Rr) + 0.5 * Rr / (1 + Rr) ** 2 * (uR - uL) ** 2)
    sL = np.minimum(ubar - dbar, uL - cL)
    sR = np.maximum(ubar + dbar, uR + cR)
    ss = (R["p"] - L["p"] + L["r"] * uL * (sL - uL) - R["r"] * uR * (sR - uR)) / (L["r"] * (sL - uL) - R["r"] * (sR - uR))

    def side(S, K):
        un = K[n_]
        E = K["G"] * K["p"] + K["P"] + 0.5 * K["r"] * (K["u"] ** 2 + K["v"] ** 2 + K["w"] ** 2)
        U = {"r": K["r"], n_: K["r"] * un, t1: K["r"] * K[t1], t2: K["r"] * K[t2], "E": E, "G": K["G"], "P": K["P"], "one": np.ones_like(un)}
        F = {"r": K["r"] * un, n_: K["r"] * un * un + K["p"], t1: K["r"] * un * K[t1], t2: K["r"] * un * K[t2],
             "E": un * (E + K["p"]), "G": K["G"] * un, "P": K["P"] * un, "one": un}
        chi = (S - un) / (S - ss)
        Us = {"r": K["r"] * chi, n_: K["r"] * chi * ss, t1: K["r"] * chi * K[t1], t2: K["r"] * chi * K[t2],
              "E": chi * (E + (ss - un) * (K["r"] * ss + K["p"] / (S - un))), "G": K["G"] * chi, "P": K["P"] * chi,
              "one": chi}
        Fs = {k: F[k] + S * (Us[k] - U[k]) for k in U}
        return F, Fs

    FL, FsL = side(sL, L)
    FR, FsR = side(sR, R)
    out = np.zeros((st.shape[0], 12))
    keys = ["r", "u", "v", "w", "E", "G", "P", "one"]
    for j, k in enumerate(keys):
        f = np.where(sL >= 0, FL[k], np.where(ss >= 0, FsL[k], np.where(sR > 0, FsR[k], FR[k])))
        out[:, j] = f
    out[:, 8], out[:, 9], out[:, 10], out[:, 11] = L["G"], R["G"], L["P"], R["P"]
    return out


def supersonic_stencils(n, seed):
    st = states.random_stencils(n, seed=seed, two_phase=False)
    # add a large uniform normal velocity (Mach ~ 5) in all three directions
    rng = np.random.default_rng(seed)
    sgn = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    out = st.astype(np.float64)
    for comp in (1, 2, 3):
        r = out[:, 0]
        u_old = out[:, comp] / r
        u_new = u_old + sgn[:, None] * 8.0
        out[:, 4] += 0.5 * r * (u_new ** 2 - u_old ** 2)
        out[:, comp] = r * u_new
    return out.astype(np.float32)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_face_matches_textbook(oracle, d):
    st = np.concatenate([states.random_stencils(8000, seed=d), supersonic_stencils(2000, seed=7 + d)])
    got = oracle.face_batch(st, d).astype(np.float64)
    ref = face_textbook(st, d)
    # scale per face: magnitude of the physical fluxes involved
    pr = prim64(st)
    rho_c = pr["r"].max(axis=1) * (np.abs(pr["u"]).max(axis=1) + np.abs(pr["v"]).max(axis=1) +
                                   np.abs(pr["w"]).max(axis=1) + 10.0)
    E = np.abs(st[:, 4].astype(np.float64)).max(axis=1)
    scale = {0: rho_c, 1: rho_c * 10 + pr["p"].max(axis=1), 4: (E + np.abs(pr["p"]).max(axis=1)) * 20}
    for j in range(12):
        sc = scale.get(j, scale[1] if j in (2, 3) else None)
        if sc is None:
            sc = np.abs(ref[:, j]) + np.abs(st[:, 5:7].astype(np.float64)).max(axis=(1, 2)) * 20 + 20
        err = np.abs(got[:, j] - ref[:, j]) / sc
        assert np.max(err) < 1e-4, (d, j, np.max(err), np.argmax(err))


def test_textbook_face_detects_planted_error(oracle):
    """A plausible slip (energy flux with p dropped from (E+p)) is caught."""
    st = states.random_stencils(2000, seed=3)
    got = oracle.face_batch(st, 0).astype(np.float64)
    ref = face_textbook(st, 0)
    pr = prim64(st)
    planted = ref[:, 4] - np.clip(pr["u"][:, 2], -1, 1) * pr["p"][:, 2]   # drop u*p from F_E
    E = np.abs(st[:, 4].astype(np.float64)).max(axis=1)
    sc = (E + np.abs(pr["p"]).max(axis=1)) * 20
    assert np.max(np.abs(got[:, 4] - planted) / sc) > 1e-3


def const_stencils(n, seed, un_zero):
    rng = np.random.default_rng(seed)
    rho = rng.choice([1.0, 1000.0], n) * rng.uniform(0.8, 1.2, n)
    water = rho > 10
    G = np.where(water, 1 / 5.59, 2.5)
    P = np.where(water, 6.59 * 4096 / 5.59, 0.0)
    p = rng.uniform(0.1, 10, n)
    vel = rng.uniform(-2, 2, (3, n))
    return rho, vel, p, G, P


@pytest.mark.parametrize("d", [0, 1, 2])
def test_equal_states_at_rest_give_exact_physical_flux(oracle, d):
    n = 3000
    rho, vel, p, G, P = const_stencils(n, d, True)
    vel[:] = 0.0
    cons = states.conserved(rho, vel[0], vel[1], vel[2], p, G, P)        # (7, n)
    st = np.ascontiguousarray(np.repeat(cons.T[:, :, None], 6, axis=2))   # (n,7,6)
    out = oracle.face_batch(st, d)
    # at rest the binary32 pressure is the correctly rounded (E - P)/G of the
    # stored values (E keeps only ~1e-3 of the liquid p: the stored-E effect)
    p32 = (st[:, 4, 0] - st[:, 6, 0]) / st[:, 5, 0]
    zero = np.zeros(n, np.float32)
    assert np.array_equal(out[:, 0], zero)                   # rho u_n
    assert np.array_equal(out[:, 1 + d], p32)                # rho u_n^2 + p = p
    for j in (1, 2, 3):
        if j != 1 + d:
            assert np.array_equal(out[:, j], zero)           # rho u_n u_t
    assert np.array_equal(out[:, 4], zero)                   # u_n (E + p)
    assert np.array_equal(out[:, 5], zero) and np.array_equal(out[:, 6], zero)
    assert np.array_equal(out[:, 7], zero)                   # face velocity


@pytest.mark.parametrize("d", [0, 1, 2])
def test_equal_states_moving_give_physical_flux(oracle, d):
    n = 3000
    rho, vel, p, G, P = const_stencils(n, 10 + d, False)
    cons = states.conserved(rho, vel[0], vel[1], vel[2], p, G, P)
    st = np.ascontiguousarray(np.repeat(cons.T[:, :, None], 6, axis=2))
    out = oracle.face_batch(st, d).astype(np.float64)
    pr = prim64(st)
    un = pr["uvw"[d]][:, 0]
    E = st[:, 4, 0].astype(np.float64)
    phys = [pr["r"][:, 0] * un] + [pr["r"][:, 0] * un * pr[c][:, 0] + (pr["p"][:, 0] if i == d else 0)
                                    for i, c in enumerate("uvw")] + \
           [un * (E + pr["p"][:, 0]), un * pr["G"][:, 0], un * pr["P"][:, 0], un]
    for j, f in enumerate(phys):
        sc = np.abs(f) + np.abs(E) * 1e-3 + pr["r"][:, 0] + 1
        assert np.max(np.abs(out[:, j] - f) / sc) < 2e-5, j
