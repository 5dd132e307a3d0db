"""CPU restatement of the reference's wavelet compressor (test infrastructure only:
imported by tests/, never by the product path).

Follows, operation for operation:
  * WaveletsOnInterval::WI4<false>::transform (source/WaveletCompression/
    WaveletsOnInterval.h:20-113): 4th-order interpolating wavelets on an
    interval, the interpolation weights applied in binary64 to binary32
    samples and the result rounded to binary32 (FwtAp = float);
  * WaveletSweep::sweep3D and its transposes (WaveletsOnInterval.h:116-180),
    literally: x rows, xy transpose, x rows, xz transpose, x rows, with the
    transposes never undone -- the recursion of FullTransformEngine
    (FullWaveletTransform.h:37-212) applies the next level to the low-pass
    corner of that permuted array;
  * FullTransformEngine::threshold / load (FullWaveletTransform.h:58-160):
    the 4^3 coarse block first, then per level (coarse to fine) the seven
    detail octants, survivors |c| > eps;
  * serialize_bitset and _cvt2f16 / _cvtfromf16 (WaveletCompressor.cpp:19-115),
    including their quirks: the binary16 conversion truncates the mantissa and
    clamps the exponent, and a binary16 zero decodes to 2^-15;
  * WaveletCompressorGeneric::compress / decompress (WaveletCompressor.cpp:
    117-180): bitset bytes followed by the survivors.
Pure numpy, for cubes up to 32^3 in seconds.

Pinned by tests/test_wavelet_oracle.py (not by retyping): interpolating
wavelets of order 4 annihilate cubic polynomials; the inverse transform
returns the input to rounding; the binary16 conversions on hand-checked
values; the bitset bit order.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def _interp(kind, f0, f1, f2, f3):
    # binary64 weights on binary32 samples, left to right (WaveletsOnInterval.h:23-41)
    f0, f1, f2, f3 = (np.float64(v) for v in (f0, f1, f2, f3))
    if kind == "first":
        return 5. / 16 * f0 + 15. / 16 * f1 - 5. / 16 * f2 + 1. / 16 * f3
    if kind == "middle":
        return -1. / 16 * f0 + 9. / 16 * f1 + 9. / 16 * f2 - 1. / 16 * f3
    if kind == "onetolast":
        return 1. / 16 * f0 - 5. / 16 * f1 + 15. / 16 * f2 + 5. / 16 * f3
    return -5. / 16 * f0 + 21. / 16 * f1 - 35. / 16 * f2 + 35. / 16 * f3


def transform(row, forward):
    """WI4<false>::transform<N, forward> on a binary32 row (in place on a copy)."""
    d = np.array(row, dtype=F32)
    N = d.size
    H = N // 2
    if forward:
        det = np.empty(H, F32)
        det[0] = F32(np.float64(d[1]) - _interp("first", d[0], d[2], d[4], d[6]))
        for i in range(1, H - 2):
            s = 2 * i
            det[i] = F32(np.float64(d[2 * i + 1]) - _interp("middle", d[s - 2], d[s], d[s + 2], d[s + 4]))
        det[H - 2] = F32(np.float64(d[N - 3]) - _interp("onetolast", d[N - 8], d[N - 6], d[N - 4], d[N - 2]))
        det[H - 1] = F32(np.float64(d[N - 1]) - _interp("last", d[N - 8], d[N - 6], d[N - 4], d[N - 2]))
        sca = d[0::2].copy()
        return np.concatenate([sca, det])
    sca, det = d[:H].copy(), d[H:].copy()
    out = np.empty(N, F32)
    out[0::2] = sca
    out[1] = F32(_interp("first", sca[0], sca[1], sca[2], sca[3]) + np.float64(det[0]))
    for i in range(1, H - 2):
        out[2 * i + 1] = F32(_interp("middle", sca[i - 1], sca[i], sca[i + 1], sca[i + 2]) + np.float64(det[i]))
    out[N - 3] = F32(_interp("onetolast", sca[H - 4], sca[H - 3], sca[H - 2], sca[H - 1]) + np.float64(det[H - 2]))
    out[N - 1] = F32(_interp("last", sca[H - 4], sca[H - 3], sca[H - 2], sca[H - 1]) + np.float64(det[H - 1]))
    return out


def _sweep1d(data, BS, forward):
    # data: [z][y][x] view; rows of the BS x BS x BS corner along x
    for iz in range(BS):
        for iy in range(BS):
            data[iz, iy, :BS] = transform(data[iz, iy, :BS], forward)


def _sweep1d_2d(plane, BS, forward):
    for iy in range(BS):
        plane[iy, :BS] = transform(plane[iy, :BS], forward)


def _xy_transpose(plane, BS):
    c = plane[:BS, :BS].copy()
    plane[:BS, :BS] = c.T


def _xz_transpose(data, BS):
    # data[iz][iy][ix] <-> data[ix][iy][iz] over the corner
    c = data[:BS, :BS, :BS].copy()
    data[:BS, :BS, :BS] = np.transpose(c, (2, 1, 0))


def _sweep3d(data, BS, forward):
    if forward:
        for iz in range(BS):
            _sweep1d_2d(data[iz], BS, True)
            _xy_transpose(data[iz], BS)
            _sweep1d_2d(data[iz], BS, True)
        _xz_transpose(data, BS)
        _sweep1d(data, BS, True)
    else:
        _sweep1d(data, BS, False)
        _xz_transpose(data, BS)
        for iz in range(BS):
            _sweep1d_2d(data[iz], BS, False)
            _xy_transpose(data[iz], BS)
            _sweep1d_2d(data[iz], BS, False)


def fwt(cube):
    """FullTransform<N>::fwt: the coefficient array as the reference leaves it."""
    data = np.array(cube, dtype=F32, copy=True)
    N = data.shape[0]
    BS = N
    while BS > 4:
        _sweep3d(data, BS, True)
        BS //= 2
    return data


def iwt(coef):
    data = np.array(coef, dtype=F32, copy=True)
    N = data.shape[0]
    sizes = []
    BS = N
    while BS > 4:
        sizes.append(BS)
        BS //= 2
    for BS in reversed(sizes):
        _sweep3d(data, BS, False)
    return data


def stream_positions(N):
    """(z, y, x) of every coefficient in stream order: the 4^3 coarse block,
    then per level BS = 8 .. N the octants code 1..7 (FullWaveletTransform.h:58-101)."""
    pos = [(z, y, x) for z in range(4) for y in range(4) for x in range(4)]
    BS = 8
    while BS <= N:
        H = BS // 2
        for code in range(1, 8):
            xs, ys, zs = H * (code & 1), H * (code >> 1 & 1), H * (code >> 2 & 1)
            pos += [(zs + z, ys + y, xs + x) for z in range(H) for y in range(H) for x in range(H)]
        BS *= 2
    return pos


def cvt2f16(x):
    u = int(np.array(x, dtype=F32).view(np.uint32))
    r = (u & 0x80000000) >> 16
    e = ((u & 0x7fffffff) >> 23) - 127
    r |= max(0, min(31, e + 15)) << 10
    r |= (u & 0x007fffff) >> 13
    return r & 0xffff


def cvtfromf16(h):
    r = (h & 0x8000) << 16
    e = ((0x1f << 10 & h) >> 10) - 15
    r |= ((e + 127) << 23) & 0xffffffff
    r |= (h & 0x3ff) << 13
    return np.array(r & 0xffffffff, dtype=np.uint32).view(F32)[()]


def serialize_bits(mask_flat):
    """serialize_bitset: bit i of the set is bit (i % 8) of byte i // 8."""
    m = np.asarray(mask_flat, dtype=np.uint8)
    nb = (m.size + 7) // 8
    out = np.zeros(nb, np.uint8)
    for b in range(8):
        part = m[b::8]
        out[:part.size] |= (part << b).astype(np.uint8)
    return out


def encode(cube, eps, half):
    """WaveletCompressorGeneric::compress: bitset bytes + survivors (float32,
    or the reference's binary16), the buffer handed to the encoder."""
    c = fwt(cube)
    N = c.shape[0]
    mask = np.zeros((N, N, N), np.uint8)
    vals = []
    for k, (z, y, x) in enumerate(stream_positions(N)):
        v = c[z, y, x]
        keep = k < 64 or abs(np.float64(v)) > np.float64(F32(eps))
        if keep:
            mask[z, y, x] = 1
            vals.append(v)
    bits = serialize_bits(mask.ravel())
    if half:
        payload = np.array([cvt2f16(v) for v in vals], dtype=np.uint16).tobytes()
    else:
        payload = np.array(vals, dtype=F32).tobytes()
    return bits.tobytes() + payload


def decode(buf, N, half):
    """WaveletCompressorGeneric::decompress: the reconstructed cube."""
    nbits = (N ** 3 + 7) // 8
    bits = np.frombuffer(buf[:nbits], np.uint8)
    mask = np.zeros(N ** 3, np.uint8)
    for b in range(8):
        mask[b::8] = (bits[:mask[b::8].size] >> b) & 1
    mask = mask.reshape(N, N, N)
    if half:
        h = np.frombuffer(buf[nbits:], np.uint16)
        vals = [cvtfromf16(int(v)) for v in h]
    else:
        vals = list(np.frombuffer(buf[nbits:], F32))
    c = np.zeros((N, N, N), F32)
    it = iter(vals)
    for k, (z, y, x) in enumerate(stream_positions(N)):
        if k < 64 or mask[z, y, x]:
            c[z, y, x] = next(it)
    return iwt(c)
