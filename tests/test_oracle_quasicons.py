"""Pins the oracle's quasi-conservative G/P assembly (oracle rhs(); reference
source/GPU/GPUkernels.cu:57-59,439-441,812-814 for the cell sums and
1104-1136 for the divergence): for the two material fields,

    R'_q = a R_q - (dt/h) ( sum_d [F_q(i+1/2) - F_q(i-1/2)] - q_cell div u* ),

the discrete form of dq/dt + div(q u) - q div u = 0 (Johnsen & Colonius 2006:
G and P are advected, not conserved), with q_cell the mean of the six face
values the cell itself reconstructs (its plus-state at each low face, its
minus-state at each high face) and div u* the sum over directions of the HLLC
face-velocity differences.  Three checks, none of which retypes the oracle:

(i)   uniform G and P under a compressive flow with varying rho and p:
      advection of a constant leaves it constant, so R'_G = a R_G and
      R'_P = a R_P to round-off of (dt/h) G |u*|;
(ii)  G, P linear along one axis with a linear velocity along it (rho and p
      uniform): the scheme is exact for linear data, so away from the
      boundaries R'_G = a R_G - dt u dG/dx exactly (closed form), in each of
      the three directions;
(iii) random two-phase states with div u != 0 and noisy G and P: the oracle
      equals an independent float64 assembly written from the form above on
      top of the textbook face solver of test_oracle_riemann (its own WENO,
      Einfeldt/Batten speeds and HLLC in Toro's branch form).

Each check is sharp enough: the oracle with a planted slip -- the
coefficient 1/3 instead of 1/6, the sign of the term flipped, or the
neighbours' face values paired into the cell sum instead of the cell's own --
fails it (the slips are planted in a float64 model of the oracle's assembly
that is first shown to reproduce the oracle's own result).
"""
import numpy as np
import pytest

import states
from test_oracle_riemann import face_textbook

A_COEF = np.float32(-17.0 / 32)
DTINVH = np.float32(0.05)


# ---------------------------------------------------------------------------
# helpers: ghost padding per BoundaryConditions.h:60-109, face stencils
# ---------------------------------------------------------------------------
def padded(s, bc):
    """(7, nz, ny, nx) -> (7, nz+6, ny+6, nx+6); ghosts of face f = 2 d + side
    copy the boundary cell (0, absorbing) or mirror the block with the normal
    momentum negated (1, reflecting).  Edges and corners stay zero: the
    one-directional stencils never read them."""
    nz, ny, nx = s.shape[1:]
    out = np.zeros((7, nz + 6, ny + 6, nx + 6), s.dtype)
    out[:, 3:-3, 3:-3, 3:-3] = s
    for d in range(3):
        ax = 3 - d
        n = s.shape[ax]
        for side in range(2):
            kind = bc[2 * d + side]
            for g in range(3):               # ghost layer g, 0 = outermost on the left
                if side == 0:
                    dst, src = g, (0 if kind == 0 else 2 - g)
                else:
                    dst, src = n + 3 + g, (n - 1 if kind == 0 else n - 1 - g)
                val = np.take(s, src, axis=ax).copy()
                if kind == 1:
                    val[1 + d] = -val[1 + d]
                sl = [slice(None), slice(3, -3), slice(3, -3), slice(3, -3)]
                sl[ax] = dst
                out[tuple(sl)] = val
    return out


def face_stencils(sp, d):
    """(n_faces, 7, 6) stencils of the faces of direction d, in the oracle's
    face-array order (fz, fy, fx)."""
    ax = 3 - d
    sl = [slice(None), slice(3, -3), slice(3, -3), slice(3, -3)]
    sl[ax] = slice(None)
    w = np.lib.stride_tricks.sliding_window_view(sp[tuple(sl)], 6, axis=ax)
    return np.ascontiguousarray(np.moveaxis(w, 0, -2)).reshape(-1, 7, 6)


def textbook_faces(s, bc, d):
    """12 face arrays of direction d from the textbook solver, shaped like
    the oracle's Block.faces(d): 7 fluxes, u*, G-, G+, P-, P+."""
    nz, ny, nx = s.shape[1:]
    shape = (nz + (d == 2), ny + (d == 1), nx + (d == 0))
    out = face_textbook(face_stencils(padded(s, bc), d), d)
    return np.moveaxis(out.reshape(*shape, 12), -1, 0)


def lo_hi(arr, d):
    """values at each cell's low and high face of direction d."""
    ax = arr.ndim - 1 - d
    n = arr.shape[ax] - 1
    return np.take(arr, np.arange(n), axis=ax), np.take(arr, np.arange(1, n + 1), axis=ax)


def assemble(faces, pairing="own"):
    """flux differences (7 fields), q_cell sums (G, P) and div u* of every
    cell from the face arrays of the three directions (float64).  pairing
    'own': the cell's plus-state at its low face and minus-state at its high
    face (correct); 'neighbour': the neighbours' states at those faces."""
    dF = dU = sG = sP = 0.0
    for d in range(3):
        f = faces[d].astype(np.float64)
        lo, hi = lo_hi(f, d)
        dF = dF + (hi[:7] - lo[:7])
        dU = dU + (hi[7] - lo[7])
        if pairing == "own":
            sG = sG + lo[9] + hi[8]
            sP = sP + lo[11] + hi[10]
        else:
            sG = sG + lo[8] + hi[9]
            sP = sP + lo[10] + hi[11]
    return dF, sG, sP, dU


def rhs_textbook(s, bc, a, R0, dtinvh):
    """R' from the quasi-conservative form on textbook faces (float64)."""
    dF, sG, sP, dU = assemble([textbook_faces(s, bc, d) for d in range(3)])
    src = np.zeros_like(dF)
    src[5], src[6] = sG / 6 * dU, sP / 6 * dU
    return float(a) * R0.astype(np.float64) - float(dtinvh) * (dF - src)


SLIPS = {
    "coefficient 1/3": dict(coef=1.0 / 3.0, sign=1.0, pairing="own"),
    "sign flipped": dict(coef=1.0 / 6.0, sign=-1.0, pairing="own"),
    "neighbour pairing": dict(coef=1.0 / 6.0, sign=1.0, pairing="neighbour"),
}


def oracle_rhs(oracle, s, a, R0, dtinvh, bc):
    nz, ny, nx = s.shape[1:]
    b = oracle.Block(nx, ny, nz)
    b.data = s
    b.tmp = R0
    b.apply_bc(list(bc))
    b.rhs(a, dtinvh)
    return b.tmp.astype(np.float64), [b.faces(d) for d in range(3)]


def planted(R, faces, dtinvh, coef, sign, pairing):
    """the oracle's R' with its G/P term replaced by a slipped one: R' is
    a R - dtinvh (dF - term), so adding dtinvh (term_slip - term) gives
    a R - dtinvh (dF - term_slip)"""
    _, sG, sP, dU = assemble(faces)
    _, sG2, sP2, dU2 = assemble(faces, pairing)
    out = R.copy()
    out[5] += float(dtinvh) * (sign * coef * sG2 * dU2 - sG / 6 * dU)
    out[6] += float(dtinvh) * (sign * coef * sP2 * dU2 - sP / 6 * dU)
    return out


# ---------------------------------------------------------------------------
# the three checks; each returns its worst normalised error
# ---------------------------------------------------------------------------
def compressive_uniform_state(n=(14, 12, 10), seed=0, phase=states.WATER):
    """uniform G, P; varying rho, p; a velocity field with div u = O(|u|)"""
    rng = np.random.default_rng(seed)
    x, y, z, h = states.centres(*n)
    G, P = states.gp(*phase)
    rho = (1000.0 if phase == states.WATER else 1.0) * (1 + 0.2 * np.sin(2 * np.pi * (x + 2 * y)) * np.cos(2 * np.pi * z))
    p = (100.0 if phase == states.WATER else 1.0) * (1 + 0.3 * np.cos(2 * np.pi * (x - z)) * np.sin(2 * np.pi * y))
    amp = 3.0 if phase == states.WATER else 0.5
    u = -amp * np.sin(2 * np.pi * x) + 0.1 * rng.uniform(-1, 1, x.shape)
    v = -amp * np.sin(2 * np.pi * y)
    w = amp * np.cos(2 * np.pi * z)
    one = np.ones_like(x)
    return states.conserved(rho, u, v, w, p, G * one, P * one)


def check_uniform(R, R0, s, faces):
    """(i) |R'_q - a R_q| / ((dt/h) q max|u*|) for q = G, P"""
    umax = max(np.abs(f[7].astype(np.float64)).max() for f in faces)
    err = 0.0
    for q in (5, 6):
        qv = float(s[q].flat[0])
        if qv == 0.0:
            continue
        ref = float(A_COEF) * R0[q].astype(np.float64)
        err = max(err, np.abs(R[q] - ref).max() / (float(DTINVH) * qv * umax))
    return err


def linear_state(d, n=16):
    """G, P and the velocity along direction d linear in that coordinate;
    rho and p uniform (two-phase mixture values, P >= 0)"""
    x, y, z, h = states.centres(n + 1, n - 2, n + 3) if d == 0 else \
        states.centres(n - 2, n + 1, n + 3) if d == 1 else states.centres(n + 3, n - 2, n + 1)
    c = (x, y, z)[d] - 0.5
    G = 1.2 + 0.8 * c
    P = 40.0 + 30.0 * c
    vel = [np.zeros_like(c) for _ in range(3)]
    vel[d] = 0.3 - 1.5 * c
    rho = np.full_like(c, 1.3)
    p = np.full_like(c, 2.0)
    return states.conserved(rho, vel[0], vel[1], vel[2], p, G, P), h


def check_linear(R, R0, s, d, h):
    """(ii) interior cells (3 or more from the low and high faces of d):
    R'_q = a R_q - (dt/h) h u dq/dx_d, normalised by (dt/h) q |u|"""
    s64 = s.astype(np.float64)
    ax = 3 - d
    n = s.shape[ax]
    idx = np.arange(3, n - 3)
    take = lambda arr: np.take(arr, idx, axis=arr.ndim - 1 - d)
    u = take(s64[1 + d] / s64[0])
    err = 0.0
    for q in (5, 6):
        qv = take(s64[q])
        # exact derivative of the stored (binary32-rounded) linear profile
        grad = (np.take(s64[q], idx + 1, axis=ax - 1) - np.take(s64[q], idx - 1, axis=ax - 1)) / (2 * h)
        exact = float(A_COEF) * take(R0[q].astype(np.float64)) - float(DTINVH) * h * u * grad
        err = max(err, (np.abs(take(R[q]) - exact) / (float(DTINVH) * (np.abs(qv) * np.abs(u) + 1e-30))).max())
    return err


def noisy_two_phase(n=(12, 11, 10), seed=1):
    """mixture cells with noisy G, P and rho, smooth p, a divergent flow"""
    rng = np.random.default_rng(seed)
    x, y, z, h = states.centres(*n)
    phi = rng.uniform(0, 1, x.shape)
    G = 0.6 + 1.9 * phi
    P = 30.0 * (1 - phi) + rng.uniform(0, 5, x.shape)
    rho = 1.0 + 0.5 * rng.uniform(-1, 1, x.shape)
    p = 1.5 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    u = -np.sin(2 * np.pi * x) + 0.2 * rng.uniform(-1, 1, x.shape)
    v = np.cos(2 * np.pi * y) * 0.8
    w = -0.7 * np.sin(2 * np.pi * z)
    return states.conserved(rho, u, v, w, p, G, P)


def check_textbook(R, ref, s, faces):
    """(iii) |R'_q - R'_textbook| / ((dt/h) (|q| + 1) max|u*|), q = G, P"""
    umax = max(np.abs(f[7].astype(np.float64)).max() for f in faces)
    err = 0.0
    for q in (5, 6):
        sc = float(DTINVH) * (np.abs(s[q].astype(np.float64)).max() + 1.0) * umax
        err = max(err, np.abs(R[q] - ref[q]).max() / sc)
    return err


# ---------------------------------------------------------------------------
# tests
# ---------------------------------------------------------------------------
TOL_UNIFORM = 1e-5
TOL_LINEAR = 1e-4
TOL_TEXTBOOK = 1e-4


def test_float64_model_reproduces_the_oracle(oracle):
    """the float64 model the slips are planted in equals the oracle's own
    R' to binary32 round-off (so a slip planted there is a slip of the oracle)"""
    s = noisy_two_phase()
    R0 = states.random_smooth(12, 11, 10, seed=4, two_phase=False)
    R, faces = oracle_rhs(oracle, s, A_COEF, R0, DTINVH, (0, 1, 0, 0, 1, 0))
    dF, sG, sP, dU = assemble(faces)
    model = float(A_COEF) * R0.astype(np.float64) - float(DTINVH) * dF
    model[5] += float(DTINVH) * sG / 6 * dU
    model[6] += float(DTINVH) * sP / 6 * dU
    assert check_textbook(R, model, s, faces) < 1e-6


@pytest.mark.parametrize("phase", [states.WATER, states.AIR])
def test_uniform_material_stays_uniform(oracle, phase):
    s = compressive_uniform_state(phase=phase)
    R0 = states.random_smooth(14, 12, 10, seed=2, two_phase=False)
    bc = (0, 0, 1, 0, 0, 1)
    R, faces = oracle_rhs(oracle, s, A_COEF, R0, DTINVH, bc)
    err = check_uniform(R, R0, s, faces)
    assert err < TOL_UNIFORM, err
    for name, slip in SLIPS.items():
        if slip["pairing"] != "own":
            continue            # every face value is the same constant: a pairing slip is invisible here
        bad = check_uniform(planted(R, faces, DTINVH, **slip), R0, s, faces)
        assert bad > 100 * TOL_UNIFORM, (name, bad)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_linear_material_is_advected(oracle, d):
    s, h = linear_state(d)
    nz, ny, nx = s.shape[1:]
    R0 = states.random_smooth(nx, ny, nz, seed=5, two_phase=False)
    R, faces = oracle_rhs(oracle, s, A_COEF, R0, DTINVH, (0,) * 6)
    err = check_linear(R, R0, s, d, h)
    assert err < TOL_LINEAR, err
    for name, slip in SLIPS.items():
        if slip["pairing"] != "own":
            continue            # linear data: both sides reconstruct the same face value
        bad = check_linear(planted(R, faces, DTINVH, **slip), R0, s, d, h)
        assert bad > 100 * TOL_LINEAR, (name, bad)


@pytest.mark.parametrize("seed,bc", [(1, (0, 0, 0, 0, 0, 0)), (2, (1, 0, 0, 1, 1, 0)), (3, (0, 1, 1, 0, 0, 1))])
def test_matches_textbook_assembly(oracle, seed, bc):
    s = noisy_two_phase(seed=seed)
    R0 = states.random_smooth(12, 11, 10, seed=seed + 10, two_phase=False)
    R, faces = oracle_rhs(oracle, s, A_COEF, R0, DTINVH, bc)
    ref = rhs_textbook(s, bc, A_COEF, R0, DTINVH)
    err = check_textbook(R, ref, s, faces)
    assert err < TOL_TEXTBOOK, err
    # the conservative fields agree with the textbook too (same faces, no source)
    umax = max(np.abs(f[7].astype(np.float64)).max() for f in faces)
    for q in range(5):
        sc = float(DTINVH) * (np.abs(ref[q]).max() / float(DTINVH) + umax * np.abs(s[q]).max() + 1.0)
        assert np.abs(R[q] - ref[q]).max() / sc < TOL_TEXTBOOK, q
    for name, slip in SLIPS.items():
        bad = check_textbook(planted(R, faces, DTINVH, **slip), ref, s, faces)
        assert bad > 20 * TOL_TEXTBOOK, (name, bad)


def test_textbook_faces_match_the_oracle_faces(oracle):
    """the stencil extraction and ghost padding used above address the same
    faces as the oracle (face values agree to the face solver's tolerance)"""
    s = noisy_two_phase(seed=7)
    bc = (1, 0, 0, 1, 0, 1)
    _, faces = oracle_rhs(oracle, s, A_COEF, np.zeros_like(s), DTINVH, bc)
    for d in range(3):
        tb = textbook_faces(s, bc, d)
        got = faces[d].astype(np.float64)
        for j in (5, 6, 7, 8, 9, 10, 11):
            sc = np.abs(tb[j]).max() + 1.0
            assert np.abs(got[j] - tb[j]).max() / sc < 1e-4, (d, j)
