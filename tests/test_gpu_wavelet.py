"""Wavelet-compressed dumps on the device against the CPU restatement of the
reference's compressor (oracle/wavelet_ref.py):

* the compressor stream (bitset bytes + survivors, binary32 or the
  reference's binary16) equals the oracle's byte for byte, for cubes of edge
  8, 16, 32 and thresholds from lossless to coarse;
* decoding a stream reproduces the oracle's reconstruction bit for bit;
* a dump of a block channel follows the reference's file layout (mini
  header, zlib ocean + LUT entry, ASCII header, block metadata, LUT header),
  parsed here independently with Python's zlib, and reads back to the
  oracle's reconstruction;
* at 256^3: a lossless dump reads back to rounding (1e-4 of the field's
  magnitude) and a thresholded one to within the threshold's reach,
  compressing.
"""
import struct
import sys
import zlib

import numpy as np
import pytest

import wavelet_ref as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-wavelet]", *a, flush=True)
    sys.stdout.flush()


def field(n, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.linspace(0, 1, n),) * 3, indexing="ij")
    r = np.sqrt((x - 0.45) ** 2 + (y - 0.5) ** 2 + (z - 0.55) ** 2)
    f = 1.0 + 0.3 * np.sin(3 * x + 2 * y) * np.cos(4 * z) + 999.0 * (r < 0.25)   # smooth + a jump
    return (f + 1e-3 * rng.standard_normal(f.shape)).astype(np.float32)


@pytest.mark.parametrize("n", [8, 16, 32])
@pytest.mark.parametrize("eps", [0.0, 1e-3, 0.5])
@pytest.mark.parametrize("half", [False, True])
def test_stream_equals_oracle(mpcf, n, eps, half):
    f = field(n, n)
    say("encode", n, eps, half)
    got = mpcf.wavelet_encode(f, eps, half)
    ref = W.encode(f, eps, half)
    assert len(got) == len(ref)
    assert got == ref
    back = mpcf.wavelet_decode(got, n, half)
    assert np.array_equal(back, W.decode(ref, n, half))


def parse_dump(path):
    raw = open(path, "rb").read()
    ocean_t = b"\n==============START-BINARY-OCEAN==============\n"
    lut_t = b"\n==============START-BINARY-LUT==============\n"
    (hdr_at,) = struct.unpack_from("<Q", raw, 0)
    mini = 8 + len(ocean_t)
    assert raw[8:mini] == ocean_t
    meta_t = b"==============START-BINARY-METABLOCKS==============\n"
    m = raw.index(meta_t, hdr_at) + len(meta_t)
    header = raw[hdr_at:m].decode()
    meta = struct.unpack_from("<5i", raw, m)
    assert raw[m + 20:m + 20 + len(lut_t)] == lut_t
    agg, nchunks = struct.unpack_from("<Qi", raw, m + 20 + len(lut_t))
    assert m + 20 + len(lut_t) + 12 == len(raw)
    assert mini + agg == hdr_at and nchunks == 1
    (lut0,) = struct.unpack_from("<Q", raw, mini + agg - 8)
    chunk = zlib.decompress(raw[mini + lut0:mini + agg - 8])
    (nb,) = struct.unpack_from("<i", chunk, 0)
    assert nb + 4 == len(chunk)
    return header, meta, chunk[4:], (len(raw), agg)


@pytest.mark.parametrize("half", [False, True])
def test_dump_layout_and_read_back(mpcf, tmp_path, half):
    n = 32
    f = mpcf.ic_shock_bubble(n, n, n, 4)
    with mpcf.Block(n, n, n, 1.0 / n) as b:
        b.data = f
        b.step(0.3)
        ten = b.stream_out(mpcf.STREAM_TENSOR)       # (n, n, n, 9)
        for ch in (0, 4):
            say("dump channel", ch, half)
            eps = 1e-2 if ch == 0 else 0.0
            stats = b.wavelet_dump(ch, tmp_path / "w", eps, half)
            path = tmp_path / f"w.StreamerGridPointIterative.channel{ch}"
            header, meta, stream, (size, agg) = parse_dump(path)
            assert meta == (0, 0, 0, 0, 0)
            assert f"Blocksize: {n}\n" in header and "Blocks: 1 x 1 x 1\n" in header
            assert f"HalfFloat: {'yes' if half else 'no'}\n" in header and "Encoder: zlib\n" in header
            assert "Wavelets: InterpWavelet4thOrder\n" in header and "sizeofReal: 4\n" in header
            cube = np.ascontiguousarray(ten[..., ch])
            ref = W.encode(cube, eps, half)
            assert stream == ref
            assert stats[0] == len(ref) and stats[1] == size
            assert stats[2] == pytest.approx(n ** 3 * 4 / agg)
            back = mpcf.wavelet_read(path)
            assert np.array_equal(back, W.decode(ref, n, half))


def test_large_block_round_trip(mpcf, tmp_path):
    n = 256
    f = mpcf.ic_shock_bubble(n, n, n, 2)
    with mpcf.Block(n, n, n, 1.0 / n) as b:
        b.data = f
        rho = np.ascontiguousarray(b.stream_out(mpcf.STREAM_TENSOR)[..., 0])
        rng = float(rho.max() - rho.min())
        say("256^3 lossless")
        s0 = b.wavelet_dump(0, tmp_path / "a", 0.0)
        back = mpcf.wavelet_read(tmp_path / "a.StreamerGridPointIterative.channel0")
        # the transform in binary32 over six levels, with boundary stencils whose
        # weights sum to 6 in magnitude: a lossless dump is exact to rounding,
        # a few hundred ulps of the field's magnitude (the 32^3 cases above
        # match the oracle bit for bit)
        assert np.abs(back - rho).max() <= 1e-4 * np.abs(rho).max()
        say("256^3 thresholded")
        eps = 1e-3 * rng
        s1 = b.wavelet_dump(0, tmp_path / "b", eps)
        back = mpcf.wavelet_read(tmp_path / "b.StreamerGridPointIterative.channel0")
        # interpolating wavelets: the reconstruction error of dropped details is a
        # few times the threshold (the details' norms are bounded by eps per level)
        assert np.abs(back - rho).max() <= 20 * eps
        assert s1[2] > s0[2] and s1[2] > 4
