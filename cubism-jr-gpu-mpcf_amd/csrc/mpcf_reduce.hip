// mpcf_reduce.hip -- binary64 reductions over node blocks, independent of the
// arithmetic build (compiled once):
//   field differences of two blocks (per-field L1 / Linf of a - b and of b),
//   the per-field sums of one block (mass, momentum, energy integrals).
// Both run in two passes -- per-block partials, then one block folding them in
// a fixed order -- so the results do not depend on scheduling and are
// bit-reproducible.  Accumulation is binary64 throughout: a thread folds at
// most a few hundred cells sequentially and the rest is a pairwise tree, so
// the relative error of a sum of N same-sign terms stays near
// (N/threads + log2(threads)) * 2^-53.
#include <hip/hip_runtime.h>
#include "mpcf_geom.h"
#include "mpcf_reduce.h"

namespace mpcf {
namespace reduce {

enum { RB = 256 };   // threads per block (64 x 4)

__device__ __forceinline__ double block_sum(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__device__ __forceinline__ double block_max(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// per field q: [0] sum |a-b|, [1] max |a-b|, [2] sum |b|, [3] max |b|.
// apair / bpair: that block stores E as a binary32 pair, its low part in
// field 7 (E = field 4 + field 7)
template <typename TA, typename TB>
__global__ __launch_bounds__(RB)
void k_diff(const Geom g, const TA* __restrict__ a, const TB* __restrict__ b, const int64_t fsa, const int64_t fsb,
            const int apair, const int bpair, double* __restrict__ part)
{
    __shared__ double sh[4];
    const int rows = g.ny * g.nz;
    for (int q = 0; q < 7; ++q) {
        double s1 = 0, m1 = 0, s2 = 0, m2 = 0;
        for (int row = blockIdx.x; row < rows; row += gridDim.x) {
            const int e0 = g.idx(0, row % g.ny, row / g.ny);
            for (int x = threadIdx.x; x < g.nx; x += RB) {
                double va = (double)a[q*fsa + e0 + x], vb = (double)b[q*fsb + e0 + x];
                if (q == 4 && apair) va += (double)a[7*fsa + e0 + x];
                if (q == 4 && bpair) vb += (double)b[7*fsb + e0 + x];
                const double d = fabs(va - vb);
                s1 += d; m1 = fmax(m1, d);
                s2 += fabs(vb); m2 = fmax(m2, fabs(vb));
            }
        }
        s1 = block_sum(s1, sh); m1 = block_max(m1, sh);
        s2 = block_sum(s2, sh); m2 = block_max(m2, sh);
        if (threadIdx.x == 0) {
            double* o = part + ((int64_t)blockIdx.x * 7 + q) * 4;
            o[0] = s1; o[1] = m1; o[2] = s2; o[3] = m2;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(RB)
void k_sums(const Geom g, const T* __restrict__ u, const int64_t fs, const int pair, double* __restrict__ part)
{
    __shared__ double sh[4];
    const int rows = g.ny * g.nz;
    for (int q = 0; q < 7; ++q) {
        double s = 0;
        for (int row = blockIdx.x; row < rows; row += gridDim.x) {
            const int e0 = g.idx(0, row % g.ny, row / g.ny);
            for (int x = threadIdx.x; x < g.nx; x += RB) {
                s += (double)u[q*fs + e0 + x];
                if (q == 4 && pair) s += (double)u[7*fs + e0 + x];
            }
        }
        s = block_sum(s, sh);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * 7 + q] = s;
    }
}

// fold nparts partial records of `width` (<= 64) values; bit k of maxmask
// set: value k is a maximum, else a sum
__global__ __launch_bounds__(RB)
void k_fold(const double* __restrict__ part, const int nparts, const int width, const unsigned long long maxmask,
            double* __restrict__ out)
{
    __shared__ double sh[4];
    for (int k = 0; k < width; ++k) {
        const bool mx = (maxmask >> k) & 1ull;
        double v = 0;
        for (int i = threadIdx.x; i < nparts; i += RB) {
            const double p = part[(int64_t)i * width + k];
            v = mx ? fmax(v, p) : v + p;
        }
        v = mx ? block_max(v, sh) : block_sum(v, sh);
        if (threadIdx.x == 0) out[k] = v;
    }
}

static int nblocks(const Geom& g)
{
    const int rows = g.ny * g.nz;
    return rows < 2048 ? rows : 2048;
}

int scratch_doubles(const Geom& g) { return nblocks(g) * 28 + 64; }

void diff(const Geom& g, const void* a, int a_f64, int64_t fsa, int a_pair, const void* b, int b_f64, int64_t fsb,
          int b_pair, double* scratch, double* out28, hipStream_t s)
{
    const int nb = nblocks(g);
    if (a_f64 && b_f64)  hipLaunchKernelGGL((k_diff<double, double>), dim3(nb), dim3(RB), 0, s, g, (const double*)a, (const double*)b, fsa, fsb, a_pair, b_pair, scratch);
    else if (a_f64)      hipLaunchKernelGGL((k_diff<double, float>), dim3(nb), dim3(RB), 0, s, g, (const double*)a, (const float*)b, fsa, fsb, a_pair, b_pair, scratch);
    else if (b_f64)      hipLaunchKernelGGL((k_diff<float, double>), dim3(nb), dim3(RB), 0, s, g, (const float*)a, (const double*)b, fsa, fsb, a_pair, b_pair, scratch);
    else                 hipLaunchKernelGGL((k_diff<float, float>), dim3(nb), dim3(RB), 0, s, g, (const float*)a, (const float*)b, fsa, fsb, a_pair, b_pair, scratch);
    unsigned long long mask = 0;
    for (int i = 0; i < 28; ++i)
        if (i % 4 == 1 || i % 4 == 3) mask |= 1ull << i;
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(RB), 0, s, scratch, nb, 28, mask, out28);
}

void sums(const Geom& g, const void* u, int f64, int64_t fs, int pair, double* scratch, double* out7, hipStream_t s)
{
    const int nb = nblocks(g);
    if (f64) hipLaunchKernelGGL((k_sums<double>), dim3(nb), dim3(RB), 0, s, g, (const double*)u, fs, pair, scratch);
    else     hipLaunchKernelGGL((k_sums<float>), dim3(nb), dim3(RB), 0, s, g, (const float*)u, fs, pair, scratch);
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(RB), 0, s, scratch, nb, 7, 0ull, out7);
}

} // namespace reduce
} // namespace mpcf
