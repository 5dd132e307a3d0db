"""The CPU restatement of the reference's wavelet compressor (oracle/wavelet_ref.py),
pinned against what the mathematics fixes:

* the 4th-order interpolating wavelet reproduces cubics: a polynomial of
  degree <= 3 in each coordinate, on integer samples, has every detail
  coefficient exactly zero at every level (a wrong weight, stencil or
  boundary case leaves nonzero details);
* the interpolation weights are the cubic Lagrange weights at the midpoints
  (and the extrapolation to the last point), checked against a direct
  evaluation;
* the inverse transform returns the input to rounding;
* the stream order visits every array position once;
* the bitset and the binary16 conversions on hand-computed values, including
  the reference's quirks (truncation, a binary16 zero decoding to 2^-15).
"""
import numpy as np
import pytest

import wavelet_ref as W


def lagrange(nodes, x):
    w = []
    for i, xi in enumerate(nodes):
        p = 1.0
        for j, xj in enumerate(nodes):
            if j != i:
                p *= (x - xj) / (xi - xj)
        w.append(p)
    return np.array(w)


def test_weights_are_cubic_lagrange():
    e = np.eye(4)
    got = {k: np.array([W._interp(k, *e[i]) for i in range(4)]) for k in ("first", "middle", "onetolast", "last")}
    # samples at the even points 0, 2, 4, 6 (or the last four); the detail sits at the odd point
    np.testing.assert_array_equal(got["first"], lagrange([0, 2, 4, 6], 1))
    np.testing.assert_array_equal(got["middle"], lagrange([0, 2, 4, 6], 3))
    np.testing.assert_array_equal(got["onetolast"], lagrange([0, 2, 4, 6], 5))
    np.testing.assert_array_equal(got["last"], lagrange([0, 2, 4, 6], 7))


@pytest.mark.parametrize("n", [8, 16])
def test_cubics_have_no_details(n):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    f = (2 * x ** 3 - 3 * x * y + y ** 2 * z - 5 * z ** 3 + 7 * x * y * z + 11).astype(np.float32)
    assert np.array_equal(f.astype(np.float64), 2 * x ** 3 - 3 * x * y + y ** 2 * z - 5 * z ** 3 + 7 * x * y * z + 11)
    c = W.fwt(f)
    coarse = np.zeros_like(c, bool)
    coarse[:4, :4, :4] = True
    assert np.all(c[~coarse] == 0)
    assert np.any(c[coarse] != 0)
    # and a quartic term does leave details
    c4 = W.fwt((f + x.astype(np.float32) ** 4).astype(np.float32))
    assert np.any(c4[~coarse] != 0)


@pytest.mark.parametrize("n", [8, 16])
def test_inverse_returns_input(n):
    rng = np.random.default_rng(n)
    f = rng.standard_normal((n, n, n)).astype(np.float32)
    back = W.iwt(W.fwt(f))
    np.testing.assert_allclose(back, f, rtol=0, atol=1e-4)


def test_stream_order_is_a_permutation():
    n = 16
    pos = W.stream_positions(n)
    assert len(pos) == n ** 3 and len(set(pos)) == n ** 3
    assert pos[:64] == [(z, y, x) for z in range(4) for y in range(4) for x in range(4)]
    # level 8 starts at 4^3 with octant code 1: x offset 4
    assert pos[64] == (0, 0, 4) and pos[64 + 64] == (0, 4, 0) and pos[512] == (0, 0, 8)


def test_bitset_and_binary16():
    assert list(W.serialize_bits([1, 0, 0, 0, 0, 0, 0, 0, 0, 1])) == [0x01, 0x02]
    assert list(W.serialize_bits([0, 0, 0, 0, 0, 0, 0, 1])) == [0x80]
    assert W.cvt2f16(1.0) == 0x3C00 and W.cvt2f16(-2.0) == 0xC000 and W.cvt2f16(0.5) == 0x3800
    assert W.cvt2f16(65504.0) == 0x7BFF
    assert W.cvt2f16(1.0 + 2.0 ** -11) == 0x3C00          # truncated, not rounded
    assert W.cvt2f16(1.0 + 2.0 ** -10) == 0x3C01
    assert W.cvtfromf16(0x3C00) == 1.0 and W.cvtfromf16(0xC000) == -2.0
    assert W.cvtfromf16(0x0000) == np.float32(2.0 ** -15)  # the reference's zero


@pytest.mark.parametrize("half", [False, True])
def test_encode_decode(half):
    n = 8
    rng = np.random.default_rng(5)
    f = rng.standard_normal((n, n, n)).astype(np.float32)
    buf = W.encode(f, 0.0, half)
    nbits = n ** 3 // 8
    assert len(buf) == nbits + (2 if half else 4) * (n ** 3)   # random data: every coefficient survives
    back = W.decode(buf, n, half)
    np.testing.assert_allclose(back, f, rtol=0, atol=(0.05 if half else 1e-4))
    # a large threshold keeps the coarse block only
    buf = W.encode(f, 1e9, half)
    assert len(buf) == nbits + (2 if half else 4) * 64
