// mpcf_noalloc.hip -- the allocation-free entry point (mpcf_max_sos_noalloc).
//
// For callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The caller
// passes its own state and output memory; the reduction's working state (the
// atomic maximum, the flag of non-physical cells, the count of finished
// blocks: one StepState) lives in `scratch`, a static device array of this
// library.  Each GPU holds its own copy of the array, at its own address,
// until the process ends; the host finds it with hipGetSymbolAddress once per
// device and passes that address to the kernels (the array's name in host
// code is only a host-side stand-in, not a device address).
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two host threads) race on it and must be ordered by the
// caller.  Each call clears the array with a small kernel on the caller's
// stream before the reduction (no hipMemsetAsync, so the pair can be captured
// in a graph and replayed).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "mpcf_internal.h"

namespace {

// 256 bytes: one StepState (the reduction targets, counters and results)
constexpr size_t kScratchBytes = 256;
static_assert(sizeof(mpcf::StepState) <= kScratchBytes, "StepState must fit the scratch array");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// zeroes the scratch array (32 lanes x 8 bytes)
__global__ __launch_bounds__(64) void k_clear_scratch(unsigned long long* p)
{
    if (threadIdx.x < kScratchBytes / sizeof(unsigned long long)) p[threadIdx.x] = 0ull;
}

std::mutex g_mu;
std::vector<void*> g_ptr;   // device address of `scratch` per device id (null: not yet looked up)

// the current device's copy of `scratch`, looked up on the first call on that
// device (the device is current: the caller's)
int scratch_ptr(void** out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    if ((size_t)dev >= g_ptr.size()) g_ptr.resize(dev + 1, nullptr);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MPCF_EHIP, "mpcf_max_sos_noalloc: hipGetSymbolAddress returned a null address");
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return MPCF_OK;
}

}  // namespace

extern "C" int mpcf_max_sos_noalloc(const void* u, long long field_stride, int nx, int ny, int nz, double h,
                                    double cfl, double dt_max, unsigned flags, double* sos_out, double* dt_out,
                                    int* bad_out, void* hip_stream)
{
    const char* who = "mpcf_max_sos_noalloc: ";
    if (!u) return fail(MPCF_EINVAL, std::string(who) + "NULL state");
    if (flags & ~(MPCF_FLAG_FAST | MPCF_FLAG_SOS_CPU | MPCF_FLAG_P64 | MPCF_FLAG_FP64))
        return fail(MPCF_EINVAL, std::string(who) + "unknown or inapplicable flags");
    if (nx < 1 || ny < 1 || nz < 1) return fail(MPCF_EINVAL, std::string(who) + "extents must be >= 1");
    const long long cells = (long long)nx * ny * nz;
    if (cells > 2147483647LL)
        return fail(MPCF_EINVAL, std::string(who) + "the state has " + std::to_string(cells) +
                                     " cells; the limit is 2^31 - 1 (32-bit cell indices)");
    if (field_stride < cells)
        return fail(MPCF_EINVAL, std::string(who) + "field_stride must be >= nx*ny*nz");
    if (!(h > 0) || !(cfl > 0) || !(dt_max > 0))
        return fail(MPCF_EINVAL, std::string(who) + "h, cfl and dt_max must be positive");
    void* p = nullptr;
    const int rc = scratch_ptr(&p);
    if (rc != MPCF_OK) return rc;
    hipStream_t s = (hipStream_t)hip_stream;
    hipLaunchKernelGGL(k_clear_scratch, dim3(1), dim3(64), 0, s, (unsigned long long*)p);
    LAUNCH_CHECK("mpcf_max_sos_noalloc: clear");
    const mpcf::Geom g{nx, ny, nz, nx, ny, nz, field_stride};
    mpcf::detail::pick_launchers(flags)->sos_dense(g, u, field_stride, (mpcf::StepState*)p,
                                                   (flags & MPCF_FLAG_SOS_CPU) ? 1 : 0, cfl, h, dt_max, sos_out,
                                                   dt_out, bad_out, s);
    LAUNCH_CHECK("mpcf_max_sos_noalloc: reduction");
    return MPCF_OK;
}
