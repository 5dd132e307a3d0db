"""Pins the oracle's WENO5 (oracle/mpcf_oracle.hpp weno_*; reference
source/GPU/GPUonly.cuh:99-194) without retyping its formulas:

* an independent textbook WENO5-JS in float64 (Jiang & Shu 1996: beta_k in the
  13/12, 1/4 square form, d = (0.1, 0.6, 0.3), alpha_k = d_k/(eps+beta_k)^2)
  must agree to float32 round-off on generic data -- a wrong smoothness
  coefficient, weight, sign or stencil index fails it;
* constants are reproduced exactly (clipping), linear data to round-off,
  the +/- states are mirror images, and results stay inside the clip bounds.
"""
import numpy as np
import pytest

EPS = 1e-6


def weno_js_textbook(a, b, c, d, e):
    """Left-biased state at the right face of cell c (stencil a..e), float64."""
    b0 = 13 / 12 * (a - 2 * b + c) ** 2 + 1 / 4 * (a - 4 * b + 3 * c) ** 2
    b1 = 13 / 12 * (b - 2 * c + d) ** 2 + 1 / 4 * (b - d) ** 2
    b2 = 13 / 12 * (c - 2 * d + e) ** 2 + 1 / 4 * (3 * c - 4 * d + e) ** 2
    a0 = 0.1 / (EPS + b0) ** 2
    a1 = 0.6 / (EPS + b1) ** 2
    a2 = 0.3 / (EPS + b2) ** 2
    s = a0 + a1 + a2
    q0 = (2 * a - 7 * b + 11 * c) / 6
    q1 = (-b + 5 * c + 2 * d) / 6
    q2 = (2 * c + 5 * d - e) / 6
    return (a0 * q0 + a1 * q1 + a2 * q2) / s


def windows(n, seed=0, jump=5.0):
    """Noise, smooth and discontinuous windows.  The jump is kept O(1): the
    reference evaluates the smoothness indicators in expanded binary32 form,
    whose cancellation on windows offset by ~1e3 is an inherent property of
    the scheme (checked separately by the clip-bound and bit-exact tests),
    not something a textbook float64 comparison can resolve."""
    rng = np.random.default_rng(seed)
    t = np.arange(5)[None, :]
    return np.concatenate([
        rng.uniform(-1, 1, (n, 5)),
        np.sin(rng.uniform(0, 1.5, (n, 1)) * t + rng.uniform(0, 6, (n, 1))),
        jump * (t >= rng.integers(1, 5, (n, 1))) + rng.uniform(0, 1, (n, 5)),
    ]).astype(np.float32)


def test_raw_weno_matches_textbook(oracle):
    w = windows(20000)
    m, p = oracle.weno(w, clipped=False)
    w64 = w.astype(np.float64)
    ref_m = weno_js_textbook(*[w64[:, k] for k in range(5)])
    ref_p = weno_js_textbook(*[w64[:, k] for k in (4, 3, 2, 1, 0)])   # mirrored stencil
    scale = np.max(np.abs(w64), axis=1) + 1e-30
    assert np.max(np.abs(m - ref_m) / scale) < 2e-5
    assert np.max(np.abs(p - ref_p) / scale) < 2e-5


def test_textbook_check_catches_a_planted_mistake():
    """The comparison above is sharp enough: a WENO with one smoothness
    coefficient perturbed (a plausible transcription slip) differs by far more
    than the tolerance."""
    rng = np.random.default_rng(3)
    w = rng.uniform(-1, 1, (5000, 5))
    good = weno_js_textbook(*w.T)
    a, b, c, d, e = w.T
    b0 = 13 / 12 * (a - 2 * b + c) ** 2 + 1 / 4 * (a - 4 * b + 3 * c) ** 2
    b1 = 13 / 12 * (b - 2 * c + d) ** 2 + 1 / 4 * (b - d) ** 2
    b2 = 13 / 12 * (c - 2 * d + e) ** 2 + 1 / 4 * (3 * c - 4 * d + e) ** 2 * 1.2   # planted
    a0, a1, a2 = 0.1 / (EPS + b0) ** 2, 0.6 / (EPS + b1) ** 2, 0.3 / (EPS + b2) ** 2
    bad = (a0 * (2 * a - 7 * b + 11 * c) + a1 * (-b + 5 * c + 2 * d) + a2 * (2 * c + 5 * d - e)) / 6 / (a0 + a1 + a2)
    assert np.max(np.abs(bad - good)) > 1e-3


def test_constants_exact(oracle):
    vals = np.array([0.0, 1.0, 2.5, 1000.0, 4828.79, 1e-7, -3.25], np.float32)
    w = np.repeat(vals[:, None], 5, axis=1)
    m, p = oracle.weno(w)
    assert np.array_equal(m, vals) and np.array_equal(p, vals)


def test_linear_data_to_roundoff(oracle):
    rng = np.random.default_rng(1)
    s = rng.uniform(0.1, 2, 2000)
    o = rng.uniform(-10, 10, 2000)
    w = (s[:, None] * np.arange(-2, 3)[None, :] + o[:, None]).astype(np.float32)
    m, p = oracle.weno(w)
    w64 = w.astype(np.float64)
    # exact face values of the linear interpolant through the float32 data
    exact_m = (w64[:, 2] + w64[:, 3]) / 2
    exact_p = (w64[:, 1] + w64[:, 2]) / 2
    tol = 1e-5 * (np.abs(w64).max(axis=1) + 1)
    assert np.all(np.abs(m - exact_m) < tol)
    assert np.all(np.abs(p - exact_p) < tol)


def test_plus_minus_mirror(oracle):
    w = windows(5000, seed=2)
    m, p = oracle.weno(w)
    m2, p2 = oracle.weno(np.ascontiguousarray(w[:, ::-1]))
    # same mathematics in a different binary32 operation order: round-off only
    scale = np.abs(w).max(axis=1) + 1e-30
    assert np.max(np.abs(p - m2) / scale) < 1e-4
    assert np.max(np.abs(m - p2) / scale) < 1e-4


def test_clip_bounds(oracle):
    w = windows(5000, seed=4, jump=1000.0)
    m, p = oracle.weno(w)
    lo = w[:, 1:4].min(axis=1)
    hi = w[:, 1:4].max(axis=1)
    assert np.all((m >= lo) & (m <= hi) & (p >= lo) & (p <= hi))


def test_float64_instantiation_agrees(oracle):
    w = windows(3000, seed=5)
    m32, p32 = oracle.weno(w)
    m64, p64 = oracle.weno(w.astype(np.float64), dtype=np.float64)
    scale = np.abs(w).max(axis=1) + 1e-30
    assert np.max(np.abs(m32 - m64) / scale) < 2e-5
    assert np.max(np.abs(p32 - p64) / scale) < 2e-5
