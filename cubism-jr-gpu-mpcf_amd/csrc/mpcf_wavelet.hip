// mpcf_wavelet.hip -- wavelet-compressed dumps of one streamer channel, the
// counterpart of the reference's SerializerIO_WaveletCompression_MPI_Simple
// (source/IO/SerializerIO_WaveletCompression_MPI_Simple.h:26-808) over its
// WaveletCompression/ library, with the transform and the thresholding on
// the device.
//
// Transform (WaveletsOnInterval.h:20-180, FullWaveletTransform.h:37-212).
// One level on the BS^3 corner of the N^3 array is three 1D passes of the
// 4th-order interpolating wavelet -- along x, then y, then z -- each output
// element a function of at most five inputs (scalings: the even samples;
// details: odd sample minus the cubic interpolant of four even neighbours,
// the dyadic weights applied in binary64 and the result rounded to
// binary32, so that every product is exact and contraction changes
// nothing).  The reference runs the passes as x rows, an xy transpose, x
// rows, an xz transpose and x rows, and never transposes back; following
// the indices through, the level leaves A[s0][s1][s2] = W[s2][s0][s1], with
// W[z][y][x] the coefficient of the separable transform, and the next level
// works on the low-pass corner of that array.  The inverse undoes z, then y,
// then x, after the inverse permutation.
//
// Thresholding (FullWaveletTransform.h:58-101, WaveletCompressor.cpp:117-146).
// The coefficient stream is the 4^3 coarse block (kept whole), then for each
// level BS = 8 .. N the detail octants code = 1..7 (x, y, z offsets BS/2 *
// bits of code), each z, y, x ordered; level BS occupies stream positions
// [(BS/2)^3, BS^3).  A coefficient survives when |c| > threshold.  The
// survivor bitset over array positions and the survivors (binary32, or the
// reference's truncating binary16) are produced by a flag kernel, an
// exclusive scan and a scatter.
//
// File (SerializerIO_WaveletCompression_MPI_Simple.h:218-397), one block:
//   size_t header address | "\n==...START-BINARY-OCEAN...==\n" |
//   zlib([int nbytes][bitset][survivors]) | size_t chunk offset (0) |
//   ASCII header | BlockMetadata {0,0,0,0,0} | "\n==...START-BINARY-LUT...==\n" |
//   HeaderLUT {ocean bytes, 1 chunk}
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "mpcf_hip.h"
#include "mpcf_internal.h"

namespace mpcf {
namespace wavelet {

using mpcf::detail::fail;

// binary64 weights of the cubic interpolants (WaveletsOnInterval.h:23-41)
__device__ __forceinline__ double interp(const int kind, const float a, const float b, const float c, const float d)
{
    const double f0 = a, f1 = b, f2 = c, f3 = d;
    if (kind == 0) return 5. / 16 * f0 + 15. / 16 * f1 - 5. / 16 * f2 + 1. / 16 * f3;
    if (kind == 1) return -1. / 16 * f0 + 9. / 16 * f1 + 9. / 16 * f2 - 1. / 16 * f3;
    if (kind == 2) return 1. / 16 * f0 - 5. / 16 * f1 + 15. / 16 * f2 + 5. / 16 * f3;
    return -5. / 16 * f0 + 21. / 16 * f1 - 35. / 16 * f2 + 35. / 16 * f3;
}

__device__ __forceinline__ int64_t at(const int n, const int z, const int y, const int x)
{
    return x + (int64_t)n * (y + (int64_t)n * z);
}

// one 1D pass along AXIS (0 x, 1 y, 2 z) over the BS^3 corner, in -> out
template <int AXIS, bool FWD>
__global__ __launch_bounds__(256) void k_pass(const float* __restrict__ in, float* __restrict__ out, const int n,
                                              const int bs)
{
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= (int64_t)bs * bs * bs) return;
    int c[3];                       // z, y, x
    c[2] = (int)(e % bs);
    c[1] = (int)((e / bs) % bs);
    c[0] = (int)(e / ((int64_t)bs * bs));
    const int ax = 2 - AXIS;        // index of the pass axis in c
    const int k = c[ax];
    const int h = bs / 2;
    auto v = [&](const int j) { int q[3] = {c[0], c[1], c[2]}; q[ax] = j; return in[at(n, q[0], q[1], q[2])]; };
    float r;
    if (FWD) {
        if (k < h) {
            r = v(2 * k);
        } else {
            const int i = k - h;
            if (i == 0)          r = (float)((double)v(1) - interp(0, v(0), v(2), v(4), v(6)));
            else if (i < h - 2)  r = (float)((double)v(2 * i + 1) - interp(1, v(2 * i - 2), v(2 * i), v(2 * i + 2), v(2 * i + 4)));
            else if (i == h - 2) r = (float)((double)v(bs - 3) - interp(2, v(bs - 8), v(bs - 6), v(bs - 4), v(bs - 2)));
            else                 r = (float)((double)v(bs - 1) - interp(3, v(bs - 8), v(bs - 6), v(bs - 4), v(bs - 2)));
        }
    } else {
        // scalings at 0 .. h-1, details at h .. bs-1
        if ((k & 1) == 0) {
            r = v(k / 2);
        } else {
            const int i = (k - 1) / 2;
            double s;
            if (i == 0)          s = interp(0, v(0), v(1), v(2), v(3));
            else if (i < h - 2)  s = interp(1, v(i - 1), v(i), v(i + 1), v(i + 2));
            else                 s = interp(i == h - 2 ? 2 : 3, v(h - 4), v(h - 3), v(h - 2), v(h - 1));
            r = (float)(s + (double)v(h + i));
        }
    }
    out[at(n, c[0], c[1], c[2])] = r;
}

// forward: A[s0][s1][s2] = W[s2][s0][s1]; inverse: W[a][b][c] = A[b][c][a]
template <bool FWD>
__global__ __launch_bounds__(256) void k_permute(const float* __restrict__ in, float* __restrict__ out, const int n,
                                                 const int bs)
{
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= (int64_t)bs * bs * bs) return;
    const int c2 = (int)(e % bs), c1 = (int)((e / bs) % bs), c0 = (int)(e / ((int64_t)bs * bs));
    if (FWD) out[at(n, c0, c1, c2)] = in[at(n, c2, c0, c1)];
    else     out[at(n, c0, c1, c2)] = in[at(n, c1, c2, c0)];
}

// array position of stream position t
__device__ __forceinline__ int64_t stream_pos(const int64_t t, const int n)
{
    if (t < 64) return at(n, (int)(t >> 4), (int)((t >> 2) & 3), (int)(t & 3));
    int bs = 8;
    while ((int64_t)bs * bs * bs <= t) bs *= 2;
    const int h = bs / 2;
    const int64_t h3 = (int64_t)h * h * h;
    const int64_t r = t - h3;
    const int code = 1 + (int)(r / h3);
    const int64_t q = r % h3;
    const int x = h * (code & 1) + (int)(q % h);
    const int y = h * ((code >> 1) & 1) + (int)((q / h) % h);
    const int z = h * ((code >> 2) & 1) + (int)(q / ((int64_t)h * h));
    return at(n, z, y, x);
}

// the reference's binary16 (WaveletCompressor.cpp:80-115): truncated mantissa,
// clamped exponent; a binary16 zero decodes to 2^-15
__device__ __forceinline__ unsigned short to_f16(const float f)
{
    const unsigned x = __float_as_uint(f);
    unsigned short r = (unsigned short)((x & 0x80000000u) >> 16);
    const int e = (int)((x & 0x7fffffffu) >> 23) - 127;
    r |= (unsigned short)(max(0, min(31, e + 15)) << 10);
    r |= (unsigned short)((x & 0x007fffffu) >> 13);
    return r;
}

__device__ __forceinline__ float from_f16(const unsigned short h)
{
    unsigned r = ((unsigned)h & 0x8000u) << 16;
    const int e = (int)(((0x1fu << 10) & h) >> 10) - 15;
    r |= (unsigned)(e + 127) << 23;
    r |= ((unsigned)h & 0x3ffu) << 13;
    return __uint_as_float(r);
}

__global__ __launch_bounds__(256) void k_mask(const float* __restrict__ a, const int n, const float eps,
                                              unsigned char* __restrict__ bits)
{
    const int64_t B = blockIdx.x * (int64_t)256 + threadIdx.x;
    const int64_t total = (int64_t)n * n * n;
    if (B * 8 >= total) return;
    unsigned char byte = 0;
    for (int b = 0; b < 8; ++b) {
        const int64_t p = B * 8 + b;
        if (p >= total) break;
        const int x = (int)(p % n), y = (int)((p / n) % n), z = (int)(p / ((int64_t)n * n));
        const bool keep = (x < 4 && y < 4 && z < 4) || fabsf(a[p]) > eps;
        byte |= (unsigned char)(keep ? 1u << b : 0u);
    }
    bits[B] = byte;
}

__global__ __launch_bounds__(256) void k_flags(const float* __restrict__ a, const int n, const float eps,
                                               int* __restrict__ fl)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= (int64_t)n * n * n) return;
    fl[t] = (t < 64 || fabsf(a[stream_pos(t, n)]) > eps) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_flags_from_bits(const unsigned char* __restrict__ bits, const int n,
                                                         int* __restrict__ fl)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= (int64_t)n * n * n) return;
    const int64_t p = stream_pos(t, n);
    fl[t] = (t < 64 || ((bits[p >> 3] >> (p & 7)) & 1)) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_scatter(const float* __restrict__ a, const int n, const int* __restrict__ fl,
                                                 const int* __restrict__ off, const int half, void* __restrict__ out)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= (int64_t)n * n * n || !fl[t]) return;
    const float v = a[stream_pos(t, n)];
    if (half) ((unsigned short*)out)[off[t]] = to_f16(v);
    else      ((float*)out)[off[t]] = v;
}

__global__ __launch_bounds__(256) void k_gather(const void* __restrict__ in, const int n, const int* __restrict__ fl,
                                                const int* __restrict__ off, const int half, float* __restrict__ a)
{
    const int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (t >= (int64_t)n * n * n) return;
    float v = 0.0f;
    if (fl[t]) v = half ? from_f16(((const unsigned short*)in)[off[t]]) : ((const float*)in)[off[t]];
    a[stream_pos(t, n)] = v;
}

__global__ __launch_bounds__(256) void k_channel(const float* __restrict__ dense, const int nc, const int ch,
                                                 const int64_t cells, float* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i < cells) out[i] = dense[i * nc + ch];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline dim3 grid_for(const int64_t items) { return dim3((unsigned)((items + 255) / 256)); }

bool cube_size_ok(const int n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }

// device scratch of one encode / decode, freed on every path
struct Scratch
{
    std::vector<void*> p;
    hipError_t e = hipSuccess;
    void* get(const size_t bytes)
    {
        void* q = nullptr;
        if (e == hipSuccess) e = hipMalloc(&q, bytes ? bytes : 1);
        if (e == hipSuccess) p.push_back(q);
        return q;
    }
    ~Scratch() { for (void* q : p) hipFree(q); }
};

#define WL_TRY(x)                                                                                 \
    do {                                                                                          \
        const hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return fail(MPCF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// in-place forward transform of the device cube a (temps t1, t2: n^3 floats)
void fwt(float* a, float* t1, float* t2, const int n, hipStream_t s)
{
    for (int bs = n; bs > 4; bs /= 2) {
        const dim3 g = grid_for((int64_t)bs * bs * bs);
        hipLaunchKernelGGL((k_pass<0, true>), g, dim3(256), 0, s, a, t1, n, bs);
        hipLaunchKernelGGL((k_pass<1, true>), g, dim3(256), 0, s, t1, t2, n, bs);
        hipLaunchKernelGGL((k_pass<2, true>), g, dim3(256), 0, s, t2, t1, n, bs);
        hipLaunchKernelGGL((k_permute<true>), g, dim3(256), 0, s, t1, a, n, bs);
    }
}

void iwt(float* a, float* t1, float* t2, const int n, hipStream_t s)
{
    for (int bs = 8; bs <= n; bs *= 2) {
        const dim3 g = grid_for((int64_t)bs * bs * bs);
        hipLaunchKernelGGL((k_permute<false>), g, dim3(256), 0, s, a, t1, n, bs);
        hipLaunchKernelGGL((k_pass<2, false>), g, dim3(256), 0, s, t1, t2, n, bs);
        hipLaunchKernelGGL((k_pass<1, false>), g, dim3(256), 0, s, t2, t1, n, bs);
        hipLaunchKernelGGL((k_pass<0, false>), g, dim3(256), 0, s, t1, a, n, bs);
    }
}

int exclusive_scan(Scratch& S, const int* fl, int* off, const int64_t items, hipStream_t s)
{
    size_t tmp = 0;
    WL_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, fl, off, (int)items, s));
    void* t = S.get(tmp);
    WL_TRY(S.e);
    WL_TRY(hipcub::DeviceScan::ExclusiveSum(t, tmp, fl, off, (int)items, s));
    return MPCF_OK;
}

// WaveletCompressorGeneric::compress of the device cube d_cube (left intact):
// bitset bytes then survivors, into raw
int encode(const float* d_cube, const int n, const float eps, const int half, std::vector<unsigned char>& raw,
           hipStream_t s)
{
    const int64_t N3 = (int64_t)n * n * n;
    const size_t nbits = (size_t)((N3 + 7) / 8);
    Scratch S;
    float* a = (float*)S.get(N3 * 4);
    float* t1 = (float*)S.get(N3 * 4);
    float* t2 = (float*)S.get(N3 * 4);
    int* fl = (int*)S.get(N3 * 4);
    int* off = (int*)S.get(N3 * 4);
    unsigned char* bits = (unsigned char*)S.get(nbits);
    WL_TRY(S.e);
    WL_TRY(hipMemcpyAsync(a, d_cube, N3 * 4, hipMemcpyDeviceToDevice, s));
    fwt(a, t1, t2, n, s);
    hipLaunchKernelGGL(k_mask, grid_for((int64_t)nbits), dim3(256), 0, s, a, n, eps, bits);
    hipLaunchKernelGGL(k_flags, grid_for(N3), dim3(256), 0, s, a, n, eps, fl);
    WL_TRY(hipGetLastError());
    int rc = exclusive_scan(S, fl, off, N3, s);
    if (rc != MPCF_OK) return rc;
    int last_off = 0, last_fl = 0;
    WL_TRY(hipMemcpyAsync(&last_off, off + N3 - 1, 4, hipMemcpyDeviceToHost, s));
    WL_TRY(hipMemcpyAsync(&last_fl, fl + N3 - 1, 4, hipMemcpyDeviceToHost, s));
    WL_TRY(hipStreamSynchronize(s));
    const int64_t survivors = (int64_t)last_off + last_fl;
    const size_t esz = half ? 2 : 4;
    void* vals = t1;   // the transform temps are free now
    hipLaunchKernelGGL(k_scatter, grid_for(N3), dim3(256), 0, s, a, n, fl, off, half, vals);
    WL_TRY(hipGetLastError());
    raw.resize(nbits + (size_t)survivors * esz);
    WL_TRY(hipMemcpyAsync(raw.data(), bits, nbits, hipMemcpyDeviceToHost, s));
    if (survivors)
        WL_TRY(hipMemcpyAsync(raw.data() + nbits, vals, (size_t)survivors * esz, hipMemcpyDeviceToHost, s));
    WL_TRY(hipStreamSynchronize(s));
    return MPCF_OK;
}

// WaveletCompressorGeneric::decompress into the device cube d_out
int decode(const unsigned char* raw, const size_t nbytes, const int n, const int half, float* d_out, hipStream_t s)
{
    const int64_t N3 = (int64_t)n * n * n;
    const size_t nbits = (size_t)((N3 + 7) / 8);
    const size_t esz = half ? 2 : 4;
    if (nbytes < nbits || (nbytes - nbits) % esz) return fail(MPCF_EINVAL, "wavelet decode: bad stream size");
    // the bitset announces the survivors (deserialize_bitset, WaveletCompressor.cpp:48-76)
    int64_t expected = 0;
    for (size_t i = 0; i < nbits; ++i) expected += __builtin_popcount(raw[i]);
    // the coarse 4^3 block is always kept (its 64 values are read whatever the
    // bits say), so its bits must be set for the count to hold
    for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                const int64_t p = x + (int64_t)n * (y + (int64_t)n * z);
                if (!((raw[p >> 3] >> (p & 7)) & 1))
                    return fail(MPCF_EINVAL, "wavelet decode: coarse block not marked in the bitset");
            }
    if ((int64_t)((nbytes - nbits) / esz) != expected)
        return fail(MPCF_EINVAL, "wavelet decode: survivor count does not match the bitset");
    Scratch S;
    float* t1 = (float*)S.get(N3 * 4);
    float* t2 = (float*)S.get(N3 * 4);
    int* fl = (int*)S.get(N3 * 4);
    int* off = (int*)S.get(N3 * 4);
    unsigned char* d_raw = (unsigned char*)S.get(nbytes);
    WL_TRY(S.e);
    WL_TRY(hipMemcpyAsync(d_raw, raw, nbytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_flags_from_bits, grid_for(N3), dim3(256), 0, s, d_raw, n, fl);
    WL_TRY(hipGetLastError());
    int rc = exclusive_scan(S, fl, off, N3, s);
    if (rc != MPCF_OK) return rc;
    hipLaunchKernelGGL(k_gather, grid_for(N3), dim3(256), 0, s, (const void*)(d_raw + nbits), n, fl, off, half, d_out);
    iwt(d_out, t1, t2, n, s);
    WL_TRY(hipGetLastError());
    WL_TRY(hipStreamSynchronize(s));
    return MPCF_OK;
}

const char* kOcean = "\n==============START-BINARY-OCEAN==============\n";
const char* kLut = "\n==============START-BINARY-LUT==============\n";
const char* kMeta = "==============START-BINARY-METABLOCKS==============\n";

#pragma pack(push, 1)
struct BlockMetadata { int idcompression, subid, ix, iy, iz; };
struct HeaderLUT { size_t aggregate_bytes; int nchunks; };
#pragma pack(pop)

std::string header_text(const int n, const double h, const float eps, const int half)
{
    // SerializerIO_WaveletCompression_MPI_Simple.h:343-394 for one block
    std::stringstream ss;
    ss << "\n==============START-ASCI-HEADER==============\n";
    const int one = 1;
    ss << "Endianess: " << (*(const char*)&one ? "little" : "big") << "\n";
    ss << "sizeofReal: " << sizeof(float) << "\n";
    ss << "sizeofsize_t: " << sizeof(size_t) << "\n";
    ss << "sizeofBlockMetadata: " << sizeof(BlockMetadata) << "\n";
    ss << "sizeofHeaderLUT: " << sizeof(HeaderLUT) << "\n";
    ss << "sizeofCompressedBlock: " << (2 * sizeof(size_t) + sizeof(int)) << "\n";
    ss << "Blocksize: " << n << "\n";
    ss << "Blocks: 1 x 1 x 1\n";
    const double ext = h * (n - 1);
    ss << "Extent: " << ext << " " << ext << " " << ext << "\n";
    ss << "SubdomainBlocks: 1 x 1 x 1\n";
    ss << "HalfFloat: " << (half ? "yes" : "no") << "\n";
    ss << "Wavelets: InterpWavelet4thOrder\n";
    ss << "WaveletThreshold: " << eps << "\n";
    ss << "Encoder: zlib\n";
    ss << kMeta;
    return ss.str();
}

int write_file(const std::string& path, const std::vector<unsigned char>& raw, const int n, const double h,
               const float eps, const int half, double stats[3])
{
    // the chunk: [int nbytes][compressor bytes], deflated (_encode_and_flush, :71-85)
    std::vector<unsigned char> chunk(sizeof(int) + raw.size());
    const int nb = (int)raw.size();
    std::memcpy(chunk.data(), &nb, sizeof(int));
    std::memcpy(chunk.data() + sizeof(int), raw.data(), raw.size());
    uLongf zlen = compressBound((uLong)chunk.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, chunk.data(), (uLong)chunk.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
        return fail(MPCF_EINVAL, "mpcf_wavelet_dump: zlib compression failed");
    const size_t lut0 = 0;                                 // the one chunk starts the ocean
    const size_t ocean = (size_t)zlen + sizeof(size_t);    // chunk + its LUT entry (:424-432)
    const size_t mini = sizeof(size_t) + std::strlen(kOcean);
    const size_t header_at = mini + ocean;
    const std::string hdr = header_text(n, h, eps, half);
    const BlockMetadata meta = {0, 0, 0, 0, 0};
    const HeaderLUT lut = {ocean, 1};
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(MPCF_EINVAL, "mpcf_wavelet_dump: cannot open " + path);
    bool ok = std::fwrite(&header_at, sizeof(size_t), 1, f) == 1 &&
              std::fwrite(kOcean, 1, std::strlen(kOcean), f) == std::strlen(kOcean) &&
              std::fwrite(z.data(), 1, zlen, f) == zlen && std::fwrite(&lut0, sizeof(size_t), 1, f) == 1 &&
              std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
              std::fwrite(&meta, sizeof(meta), 1, f) == 1 &&
              std::fwrite(kLut, 1, std::strlen(kLut), f) == std::strlen(kLut) &&
              std::fwrite(&lut, sizeof(lut), 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(MPCF_EINVAL, "mpcf_wavelet_dump: write failed for " + path);
    if (stats) {
        stats[0] = (double)raw.size();
        stats[1] = (double)(header_at + hdr.size() + sizeof(meta) + std::strlen(kLut) + sizeof(lut));
        stats[2] = (double)n * n * n * sizeof(float) / (double)ocean;   // the reference's "compr. rate"
    }
    return MPCF_OK;
}

}  // namespace wavelet
}  // namespace mpcf

using namespace mpcf::wavelet;

extern "C" int mpcf_wavelet_encode(const float* cube, int n, float threshold, int halffloat, unsigned char* out,
                                   long long cap, long long* nbytes)
{
    if (!cube || !nbytes || (cap > 0 && !out)) return fail(MPCF_EINVAL, "mpcf_wavelet_encode: NULL argument");
    if (!cube_size_ok(n)) return fail(MPCF_EINVAL, "mpcf_wavelet_encode: n must be a power of two in 8..1024");
    const int64_t N3 = (int64_t)n * n * n;
    float* d = nullptr;
    WL_TRY(hipMalloc(&d, N3 * 4));
    std::vector<unsigned char> raw;
    hipError_t e = hipMemcpy(d, cube, N3 * 4, hipMemcpyHostToDevice);
    int rc = e == hipSuccess ? encode(d, n, threshold, halffloat, raw, nullptr)
                             : fail(MPCF_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    hipFree(d);
    if (rc != MPCF_OK) return rc;
    *nbytes = (long long)raw.size();
    if (cap > 0) std::memcpy(out, raw.data(), (size_t)std::min<long long>(cap, (long long)raw.size()));
    return MPCF_OK;
}

extern "C" int mpcf_wavelet_decode(const unsigned char* in, long long nbytes, int n, int halffloat, float* cube)
{
    if (!in || !cube || nbytes < 0) return fail(MPCF_EINVAL, "mpcf_wavelet_decode: bad argument");
    if (!cube_size_ok(n)) return fail(MPCF_EINVAL, "mpcf_wavelet_decode: n must be a power of two in 8..1024");
    const int64_t N3 = (int64_t)n * n * n;
    float* d = nullptr;
    WL_TRY(hipMalloc(&d, N3 * 4));
    int rc = decode(in, (size_t)nbytes, n, halffloat, d, nullptr);
    if (rc == MPCF_OK) {
        const hipError_t e = hipMemcpy(cube, d, N3 * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(MPCF_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    hipFree(d);
    return rc;
}

extern "C" int mpcf_wavelet_dump(mpcf_ctx* c, int channel, const char* path, float threshold, int halffloat,
                                 double stats[3])
{
    if (!c || !path) return fail(MPCF_EINVAL, "mpcf_wavelet_dump: NULL argument");
    if (mpcf::detail::refuse_epair(c->flags, "mpcf_wavelet_dump") != MPCF_OK) return MPCF_EINVAL;
    if (channel < 0 || channel > 6) return fail(MPCF_EINVAL, "mpcf_wavelet_dump: channel must be 0..6");
    const int n = c->g.nx;
    if (c->g.ny != n || c->g.nz != n || !cube_size_ok(n))
        return fail(MPCF_EINVAL, "mpcf_wavelet_dump: the block must be a cube of power-of-two edge 8..1024");
    const int64_t N3 = (int64_t)n * n * n;
    // StreamerGridPointIterative channels 0..6 are the tensorial streamer's
    // first seven (Types.h:157-186): rho, u, v, w, p, G, P
    const Launchers* L = (c->flags & MPCF_FLAG_FP64) ? c->L : &kParity;
    L->stream(c->g, c->u[c->cur], (float*)c->acc, MPCF_STREAM_TENSOR, 1, c->stream);
    float* cube = nullptr;
    WL_TRY(hipMalloc(&cube, N3 * 4));
    hipLaunchKernelGGL(k_channel, grid_for(N3), dim3(256), 0, c->stream, (const float*)c->acc, 9, channel, N3, cube);
    std::vector<unsigned char> raw;
    int rc = encode(cube, n, threshold, halffloat, raw, c->stream);
    hipFree(cube);
    if (rc != MPCF_OK) return rc;
    const std::string file = std::string(path) + ".StreamerGridPointIterative.channel" + std::to_string(channel);
    return write_file(file, raw, n, c->h, threshold, halffloat, stats);
}

extern "C" int mpcf_wavelet_read(const char* file, float* cube, long long cap, int* n_out)
{
    if (!file || !n_out) return fail(MPCF_EINVAL, "mpcf_wavelet_read: NULL argument");
    FILE* f = std::fopen(file, "rb");
    if (!f) return fail(MPCF_EINVAL, std::string("mpcf_wavelet_read: cannot open ") + file);
    std::vector<unsigned char> all;
    {
        unsigned char buf[1 << 16];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) all.insert(all.end(), buf, buf + got);
        std::fclose(f);
    }
    const size_t mini = sizeof(size_t) + std::strlen(kOcean);
    if (all.size() < mini) return fail(MPCF_EINVAL, "mpcf_wavelet_read: file too short");
    size_t header_at = 0;
    std::memcpy(&header_at, all.data(), sizeof(size_t));
    if (header_at <= mini || header_at >= all.size() || std::memcmp(all.data() + sizeof(size_t), kOcean, std::strlen(kOcean)))
        return fail(MPCF_EINVAL, "mpcf_wavelet_read: not a wavelet dump");
    const std::string text(all.begin() + header_at, all.end());
    const size_t meta_at = text.find(kMeta);
    if (meta_at == std::string::npos) return fail(MPCF_EINVAL, "mpcf_wavelet_read: header without metablocks");
    auto field = [&](const char* key) -> std::string {
        const size_t p = text.find(std::string(key) + ": ");
        if (p == std::string::npos || p > meta_at) return "";
        const size_t q = text.find('\n', p);
        return text.substr(p + std::strlen(key) + 2, q - p - std::strlen(key) - 2);
    };
    const int n = std::atoi(field("Blocksize").c_str());
    const int half = field("HalfFloat") == "yes";
    if (!cube_size_ok(n) || field("Encoder") != "zlib" || field("Wavelets") != "InterpWavelet4thOrder" ||
        field("sizeofReal") != "4")
        return fail(MPCF_EINVAL, "mpcf_wavelet_read: unsupported header");
    *n_out = n;
    const int64_t N3 = (int64_t)n * n * n;
    if (!cube || cap < N3) return MPCF_OK;   // size query
    // the one chunk spans the ocean up to the header (the reader's LUT, :590-660)
    std::vector<unsigned char> chunk(sizeof(int) + (size_t)((N3 + 7) / 8) + (size_t)N3 * 4);
    uLongf clen = (uLongf)chunk.size();
    if (uncompress(chunk.data(), &clen, all.data() + mini, (uLong)(header_at - mini - sizeof(size_t))) != Z_OK ||
        clen < sizeof(int))
        return fail(MPCF_EINVAL, "mpcf_wavelet_read: zlib stream damaged");
    int nb = 0;
    std::memcpy(&nb, chunk.data(), sizeof(int));
    if (nb < 0 || (size_t)nb + sizeof(int) != clen) return fail(MPCF_EINVAL, "mpcf_wavelet_read: chunk size mismatch");
    return mpcf_wavelet_decode(chunk.data() + sizeof(int), nb, n, half, cube);
}
