// mpcf_capi.cpp -- the C-ABI declared in include/mpcf_hip.h.
//
// A context keeps one node block resident in HBM for its whole life:
//   u[2]  two solution buffers (stage input / stage output, swapped per stage)
//   R     the LSRK3 register (the reference's NodeBlock::tmp)
//   acc   10 accumulator fields carried between the x, y and z sweeps; also
//         the staging area of uploads and downloads
//   st    device StepState (max speed of sound, dt, dt/h)
// This replaces the reference's chunked host pipeline (GPUlab.cpp:147-403,
// 703-793) and its per-stage PCIe traffic: after mpcf_upload nothing crosses
// PCIe until mpcf_download, and dt is produced and consumed on the device.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <condition_variable>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mpcf_hip.h"
#include "mpcf_internal.h"
#include "mpcf_reduce.h"

using mpcf::Geom;
using mpcf::StepState;

namespace mpcf {
namespace detail {

namespace {
thread_local std::string g_err;
}

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

const char* last_error() { return g_err.c_str(); }

}  // namespace detail
}  // namespace mpcf

namespace {

// kernel classes reported by mpcf_timing_read
enum { K_GHOSTS = 0, K_SOS = 1, K_DT = 2, K_SWEEP_X = 3, K_SWEEP_Y = 4, K_SWEEP_Z = 5, K_UPDATE = 6, K_OTHER = 7 };

hipEvent_t take_event(mpcf_ctx* c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// brackets one launch with events on the context stream when timing is on
struct Timed
{
    mpcf_ctx* c;
    int kind;
    hipEvent_t b = nullptr;
    hipStream_t s;
    Timed(mpcf_ctx* c_, int k, hipStream_t s_ = nullptr) : c(c_), kind(k), s(s_ ? s_ : c_->stream)
    {
        if (c->timing) {
            b = take_event(c);
            if (b) hipEventRecord(b, s);
        }
    }
    ~Timed()
    {
        if (c->timing && b) {
            hipEvent_t e = take_event(c);
            if (e) {
                hipEventRecord(e, s);
                c->marks.push_back({kind, b, e});
            } else {
                c->ev_pool.push_back(b);
            }
        }
    }
};

// the fused x+y sweep wins when the arithmetic is cheap (FAST); the split
// pair wins when it is VALU-bound (bit-exact build), unless forced
bool split_xy(const mpcf_ctx* c)
{
    // default: the fused x+y sweep, except in the binary64 build (fused and
    // split give the same results; the fused one reads u once and writes the
    // accumulators once, the split x sweep's LDS-staged runs over-fetch)
    return (c->flags & MPCF_FLAG_SPLIT_XY) ? true : (c->flags & MPCF_FLAG_FUSED_XY) ? false
                                                   : (c->flags & MPCF_FLAG_FP64) != 0;
}

// R <- a R - dtinvh L(u) (stage == 0), or that and u_out <- u + b R (stage == 1;
// stage == 2: u_out only, R not stored)
void run_rhs(mpcf_ctx* c, const void* u, void* uo, double a, double b, double dtinvh, const StepState* st, int stage,
             StepState* sos_next = nullptr)
{
    mpcf::SweepParams p;
    p.u = u; p.R = c->R; p.acc = c->acc; p.uo = uo;
    p.a = a; p.b = b; p.dtinvh = dtinvh; p.st = st;
    p.sos_next = sos_next; p.sos_cpu = (c->flags & MPCF_FLAG_SOS_CPU) ? 1 : 0;
    p.bcmask = mpcf::detail::bc_mask(c->bc);
    if (split_xy(c)) {
        { Timed t(c, K_SWEEP_X); c->L->sweep(c->g, 0, p, stage, c->stream); }
        { Timed t(c, K_SWEEP_Y); c->L->sweep(c->g, 1, p, stage, c->stream); }
    } else {
        Timed t(c, K_SWEEP_X);
        c->L->sweep(c->g, 3, p, stage, c->stream);
    }
    { Timed t(c, K_SWEEP_Z); c->L->sweep(c->g, 2, p, stage, c->stream); }
}

// z-slabs of a pipelined step (0: none): the count mpcf_set_pipeline asked
// for, else one from the block's size
int pipeline_slabs(const mpcf_ctx* c)
{
    if (c->pipe_req >= 0) return c->pipe_req;
    // the split sweeps are all VALU-bound: slabs give them nothing to overlap
    // (measured on MI355X at 512^3, bit-exact build: whole-block 59.4 ms a
    // step, 4 slabs 60.1, 8 slabs 61.4)
    if (split_xy(c)) return 0;
    // so is the bit-exact binary32 build's fused sweep, whose division and
    // binary64 weight sequences make both sweeps VALU-bound (512^3: whole-
    // block 57.05 ms a step, 4 slabs 57.6, 8 slabs 57.9; 384^3: whole 25.16,
    // 3 slabs 25.17, 6 slabs 25.3)
    if (!(c->flags & MPCF_FLAG_FAST)) return 0;
    // blocks of 2^24 cells or more, slabs by depth as measured on MI355X (fast
    // build, cubes, ms a step): 256^3 4 slabs (2.66; 2.74 with 2, 2.81
    // whole); 320^3 3 (4.83; 4.97 with 4, 4.98 with 5); 384^3 3 (8.29; 8.34
    // with 4, 8.49 with 6); 448^3 4 (13.0; 13.2 with 7); 512^3 8 (19.8;
    // 20.0-20.1 with 6 or 10, 20.3 with 12); 640^3 8 (10 slabs 3 % faster on
    // one box, equal on another); 768^3 and 1024^3 8 (12 slabs were 6 %
    // slower at 1024^3).  128^3 runs best whole.
    const int64_t cells = (int64_t)c->g.nx * c->g.ny * c->g.nz;
    const int nz = c->g.nz;
    const int S = nz < 320 ? nz / 64 : nz < 512 ? nz / 100 : std::min(8, nz / 64);
    return (cells >= (int64_t(1) << 24) && S >= 2) ? S : 0;
}

// the x+y sweeps of a step's first stage need no dt (the FAST builds store
// their part unscaled when a = 0, and the z sweeps scale it)
bool dt_deferred(const mpcf_ctx* c)
{
    return (c->flags & MPCF_FLAG_FAST) && !(c->flags & MPCF_FLAG_FP64);
}

// the three stages of a step over S z-slabs on two streams.  The x+y sweep
// (the fused one, or the x and the y sweep in turn) of slab j needs the
// previous stage's z sweep of slab j only (its stencil is in x and y; it also
// rewrites that slab's R and accumulators), so it waits for that event; the
// z sweep of slab j waits for the x+y sweep of slab j and,
// by the order of the second stream, for every z sweep of the previous
// stage (its stencil reaches into the neighbouring slabs, and its output
// buffer is the one they read).  The x+y sweeps of the next slab run while a
// z sweep runs: VALU-bound and bandwidth-bound work overlap.
// prev_open: the previous step was pipelined with S slabs and left open, so
// the first stage's x+y sweep of slab j waits for that step's last z sweep of
// slab j (which wrote the slab's u) instead of the whole step.  open: leave
// this step open the same way (the caller's stream has not joined it).
// skipR: the last stage's z sweeps do not store R.
int run_stages_pipelined(mpcf_ctx* c, const int S, const bool prev_open, const bool open, const bool skipR)
{
    if (!c->stream2) HIP_TRY(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    while (c->pipe_ev.size() < (size_t)(2 * 3 * S)) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->pipe_ev.push_back(e);
    }
    const bool split = split_xy(c);   // x and y as two sweeps per slab
    auto exy = [&](int s, int j) { return c->pipe_ev[(s * S + j) * 2]; };
    auto ez = [&](int s, int j) { return c->pipe_ev[(s * S + j) * 2 + 1]; };
    for (int s = 0; s < 3; ++s) {
        mpcf::SweepParams p;
        p.u = c->u[c->cur]; p.R = c->R; p.acc = c->acc; p.uo = c->u[1 - c->cur];
        p.a = kA[s]; p.b = kB[s]; p.dtinvh = 0.0; p.st = c->st;
        p.sos_next = s == 2 ? c->st : nullptr; p.sos_cpu = (c->flags & MPCF_FLAG_SOS_CPU) ? 1 : 0;
        p.bcmask = mpcf::detail::bc_mask(c->bc);
        for (int j = 0; j < S; ++j) {
            p.z0 = (int)((int64_t)c->g.nz * j / S);
            p.z1 = (int)((int64_t)c->g.nz * (j + 1) / S);
            if (s > 0) HIP_TRY(hipStreamWaitEvent(c->stream, ez(s - 1, j), 0));
            else if (prev_open) HIP_TRY(hipStreamWaitEvent(c->stream, ez(2, j), 0));
            if (split) {
                { Timed t(c, K_SWEEP_X); c->L->sweep(c->g, 0, p, 1, c->stream); }
                { Timed t(c, K_SWEEP_Y); c->L->sweep(c->g, 1, p, 1, c->stream); }
            } else {
                Timed t(c, K_SWEEP_X);
                c->L->sweep(c->g, 3, p, 1, c->stream);
            }
            HIP_TRY(hipEventRecord(exy(s, j), c->stream));
            HIP_TRY(hipStreamWaitEvent(c->stream2, exy(s, j), 0));
            { Timed t(c, K_SWEEP_Z, c->stream2); c->L->sweep(c->g, 2, p, (s == 2 && skipR) ? 2 : 1, c->stream2); }
            HIP_TRY(hipEventRecord(ez(s, j), c->stream2));
        }
        LAUNCH_CHECK("stage");
        c->cur = 1 - c->cur;
    }
    // everything after the step is ordered after its last z sweep (now, or
    // when the next entry point joins)
    if (open) c->pipe_open = S;
    else HIP_TRY(hipStreamWaitEvent(c->stream, ez(2, S - 1), 0));
    return MPCF_OK;
}

// the context's stream joins the z sweeps a pipelined step left open
int join_pipeline(mpcf_ctx* c)
{
    if (c->pipe_open) {
        const int S = c->pipe_open;
        c->pipe_open = 0;
        HIP_TRY(hipStreamWaitEvent(c->stream, c->pipe_ev[(2 * S + S - 1) * 2 + 1], 0));   // last z sweep
    }
    return MPCF_OK;
}

} // namespace

// message of MPCF_ENUMERIC for a flagged cell (the reference's assertions,
// GPUkernels.cu:1171-1174,1218, MaxSpeedOfSound.cpp:31-48)
const char* const kBadCell = "has a non-finite or non-physical cell (NaN / inf, rho <= 0, E <= 0, G <= 0 or c <= 0)";

extern "C" const char* mpcf_last_error(void) { return mpcf::detail::last_error(); }

extern "C" const char* mpcf_version(void) { return "mpcf-mi355x 0.1 (gfx950)"; }

extern "C" int mpcf_create(mpcf_ctx** out, int nx, int ny, int nz, double h, unsigned flags)
{
    if (!out) return fail(MPCF_EINVAL, "mpcf_create: out is NULL");
    *out = nullptr;
    if (nx < 5 || ny < 5 || nz < 5) return fail(MPCF_EINVAL, "mpcf_create: extents must be >= 5");
    if (!(h > 0) || !std::isfinite(h)) return fail(MPCF_EINVAL, "mpcf_create: h must be positive and finite");
    if (flags & ~kAllFlags) return fail(MPCF_EINVAL, "mpcf_create: unknown flag");
    if ((flags & MPCF_FLAG_SPLIT_XY) && (flags & MPCF_FLAG_FUSED_XY))
        return fail(MPCF_EINVAL, "mpcf_create: SPLIT_XY and FUSED_XY are exclusive");
    if ((flags & MPCF_FLAG_FP64) && (flags & (MPCF_FLAG_FAST | MPCF_FLAG_P64 | MPCF_FLAG_E_PAIR)))
        return fail(MPCF_EINVAL, "mpcf_create: FP64 excludes FAST, P64 and E_PAIR");
    if ((flags & MPCF_FLAG_E_PAIR) && (flags & MPCF_FLAG_SPLIT_XY))
        return fail(MPCF_EINVAL, "mpcf_create: E_PAIR runs the fused x+y sweep only");
    if (flags & MPCF_FLAG_E_PAIR) flags |= MPCF_FLAG_P64;  // the pair is the P64 arithmetic plus E's low part
    if (flags & MPCF_FLAG_P64) flags |= MPCF_FLAG_FAST;   // P64 is the FAST arithmetic plus binary64 steps
    const Geom g = mpcf::make_geom(nx, ny, nz);
    if ((int64_t)g.px * g.py * g.pz >= (int64_t(1) << 31))
        return fail(MPCF_EINVAL, "mpcf_create: block too large for 32-bit element indices");

    mpcf_ctx* c = new mpcf_ctx();
    c->g = g; c->h = h; c->flags = flags; c->cur = 0; c->sos_folded = false;
    c->pipe_req = (flags & MPCF_FLAG_NO_PIPELINE) ? 0 : -1;
    for (int i = 0; i < 6; ++i) c->bc[i] = MPCF_BC_ABSORBING;
    c->L = pick_launchers(flags);
    c->esize = elem_bytes(flags);
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) { delete c; return fail(MPCF_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e)); }

    // a solution buffer holds 7 fields, or 8 with E's low part (E_PAIR)
    const int nu = (flags & MPCF_FLAG_E_PAIR) ? 8 : 7;
    const size_t nfields = nu + nu + 7 + mpcf::NACC;
    c->bytes = nfields * (size_t)g.fstride * c->esize + 256;
    e = hipMalloc((void**)&c->base, c->bytes);
    if (e != hipSuccess) {
        delete c;
        return fail(MPCF_ENOMEM, std::string("hipMalloc of block state: ") + hipGetErrorString(e));
    }
    e = hipHostMalloc((void**)&c->st_host, sizeof(StepState), hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // zero the state on the context's own stream and wait: a non-blocking
    // stream is not ordered after work on the null stream, so a plain
    // hipMemset could land after the first upload or stage (freed memory of
    // an earlier block may hold NaN, and R must start at zero)
    if (e == hipSuccess) e = hipMemsetAsync(c->base, 0, c->bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        if (c->stream) hipStreamDestroy(c->stream);
        if (c->st_host) hipHostFree(c->st_host);
        hipFree(c->base);
        delete c;
        return fail(MPCF_EHIP, std::string("mpcf_create: ") + hipGetErrorString(e));
    }
    c->own_stream = true;
    const size_t fb = (size_t)g.fstride * c->esize;   // bytes per field (a multiple of 256)
    c->u[0] = c->base;
    c->u[1] = c->base + nu * fb;
    c->R = c->base + 2 * nu * fb;
    c->acc = c->base + (2 * nu + 7) * fb;
    c->st = (StepState*)(c->base + nfields * fb);
    static_assert(sizeof(StepState) <= 256, "step state must fit the reserved tail");
    std::memset(c->st_host, 0, sizeof(StepState));
    *out = c;
    return MPCF_OK;
}

extern "C" int mpcf_destroy(mpcf_ctx* c)
{
    if (!c) return MPCF_OK;
    hipStreamSynchronize(c->stream);
    if (c->stream2) {
        hipStreamSynchronize(c->stream2);
        for (hipEvent_t e : c->pipe_ev) if (e) hipEventDestroy(e);
        if (c->pipe_fork) hipEventDestroy(c->pipe_fork);
        hipStreamDestroy(c->stream2);
    }
    if (c->own_stream) hipStreamDestroy(c->stream);
    for (auto& m : c->marks) { hipEventDestroy(m.b); hipEventDestroy(m.e); }
    for (auto e : c->ev_pool) hipEventDestroy(e);
    hipFree(c->base);
    hipHostFree(c->st_host);
    delete c;
    return MPCF_OK;
}

extern "C" int mpcf_set_pipeline(mpcf_ctx* c, int slabs)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_set_pipeline: NULL context");
    if (slabs < -1 || slabs == 1 || slabs > std::min(64, c->g.nz))
        return fail(MPCF_EINVAL, "mpcf_set_pipeline: slabs must be -1 (automatic), 0 (off) or 2..min(64, NZ)");
    c->pipe_req = slabs;
    return MPCF_OK;
}

extern "C" int mpcf_set_stream(mpcf_ctx* c, void* stream)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_set_stream: NULL context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->own_stream) {
        // a new stream first: on failure the context keeps its old one
        if (!stream) {
            hipStream_t fresh = nullptr;
            HIP_TRY(hipStreamCreateWithFlags(&fresh, hipStreamNonBlocking));
            hipStreamDestroy(c->stream);
            c->stream = fresh;
            return MPCF_OK;
        }
        HIP_TRY(hipStreamDestroy(c->stream));
    }
    if (stream) {
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    return MPCF_OK;
}

extern "C" void* mpcf_get_stream(mpcf_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int mpcf_set_bc(mpcf_ctx* c, const int bc[6])
{
    if (!c || !bc) return fail(MPCF_EINVAL, "mpcf_set_bc: NULL argument");
    for (int i = 0; i < 6; ++i)
        if (bc[i] != MPCF_BC_ABSORBING && bc[i] != MPCF_BC_REFLECTING && bc[i] != MPCF_BC_HALO)
            return fail(MPCF_EINVAL, "mpcf_set_bc: unknown boundary kind");
    if (mpcf::detail::has_halo(bc) && mpcf::detail::refuse_epair(c->flags, "mpcf_set_bc (halo faces)") != MPCF_OK)
        return MPCF_EINVAL;
    for (int i = 0; i < 6; ++i) c->bc[i] = bc[i];
    return MPCF_OK;
}

namespace {

int halo_args(mpcf_ctx* c, const int face, const void* buf, const char* who)
{
    if (!c || !buf) return fail(MPCF_EINVAL, std::string(who) + ": NULL argument");
    if (face < 0 || face > 5) return fail(MPCF_EINVAL, std::string(who) + ": face must be 0..5");
    return mpcf::detail::refuse_epair(c->flags, who);
}

long long halo_count(const mpcf_ctx* c, const int face)
{
    const int dir = face >> 1;
    const long long area = dir == 0 ? (long long)c->g.ny * c->g.nz : dir == 1 ? (long long)c->g.nx * c->g.nz
                                                                              : (long long)c->g.nx * c->g.ny;
    return 7 * 3 * area;
}

}  // namespace

extern "C" int mpcf_halo_count(mpcf_ctx* c, int face, long long* count)
{
    const int rc = halo_args(c, face, count, "mpcf_halo_count");
    if (rc != MPCF_OK) return rc;
    *count = halo_count(c, face);
    return MPCF_OK;
}

extern "C" int mpcf_halo_export(mpcf_ctx* c, int face, void* dst)
{
    const int rc = halo_args(c, face, dst, "mpcf_halo_export");
    if (rc != MPCF_OK) return rc;
    c->L->halo(c->g, c->u[c->cur], dst, face, 0, c->stream);
    LAUNCH_CHECK("halo export");
    return MPCF_OK;
}

extern "C" int mpcf_ghosts_read(mpcf_ctx* c, int face, void* dst)
{
    const int rc = halo_args(c, face, dst, "mpcf_ghosts_read");
    if (rc != MPCF_OK) return rc;
    c->L->halo(c->g, c->u[c->cur], dst, face, 2, c->stream);
    LAUNCH_CHECK("ghosts read");
    return MPCF_OK;
}

extern "C" int mpcf_halo_import(mpcf_ctx* c, int face, const void* src)
{
    const int rc = halo_args(c, face, src, "mpcf_halo_import");
    if (rc != MPCF_OK) return rc;
    c->L->halo(c->g, c->u[c->cur], const_cast<void*>(src), face, 1, c->stream);
    LAUNCH_CHECK("halo import");
    return MPCF_OK;
}

namespace {

// host fields <-> the padded device layout, one field at a time through the
// accumulator area (free between stages), converting between the host type
// (binary32 or binary64) and the storage type on the device
// Downloads into pageable host memory: the runtime's own path moves about
// 9 GB/s, a copy into pinned memory about 55 GB/s.  Large downloads go through
// one process-wide pair of pinned chunks (allocated on first use, kept for the
// life of the process): the copy engine fills one chunk while host threads
// move the other into the caller's array.
struct PinnedChunks
{
    std::mutex mu;
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t cap = 0;
};

PinnedChunks& pinned_chunks()
{
    static PinnedChunks p;
    return p;
}

// A host copy split over a pool of worker threads (the copy and the first
// touch of the destination's pages are the cost).  The pool is created on
// first use and lives as long as the process; one copy runs at a time
// (`busy` serialises the callers).
struct CopyPool
{
    std::mutex busy;
    std::mutex mu;
    std::condition_variable go, done;
    std::vector<std::thread> workers;
    char* dst = nullptr;
    const char* src = nullptr;
    size_t n = 0, per = 0;
    unsigned parts = 0, left = 0;
    unsigned long long round = 0;

    void work(const unsigned i)
    {
        unsigned long long seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu);
            go.wait(lk, [&] { return round != seen; });
            seen = round;
            const size_t off = (size_t)(i + 1) * per;
            const bool mine = i + 1 < parts && off < n;
            char* d = dst;
            const char* sp = src;
            const size_t len = mine ? std::min(per, n - off) : 0;
            lk.unlock();
            if (mine) std::memcpy(d + off, sp + off, len);
            lk.lock();
            if (--left == 0) done.notify_one();
        }
    }

    void copy(char* d, const char* sp, const size_t bytes)
    {
        const size_t kPiece = size_t(1) << 20;
        const unsigned nw = (unsigned)workers.size();
        const unsigned p = (unsigned)std::min<size_t>(nw + 1, std::max<size_t>(1, bytes / kPiece));
        if (p < 2) {
            std::memcpy(d, sp, bytes);
            return;
        }
        std::lock_guard<std::mutex> one(busy);
        {
            std::lock_guard<std::mutex> lk(mu);
            dst = d; src = sp; n = bytes; parts = p; per = (bytes + p - 1) / p;
            left = nw;
            ++round;
        }
        go.notify_all();
        std::memcpy(d, sp, std::min(per, bytes));
        std::unique_lock<std::mutex> lk(mu);
        done.wait(lk, [&] { return left == 0; });
    }
};

CopyPool& copy_pool()
{
    static CopyPool* pool = [] {
        CopyPool* P = new CopyPool();   // never destroyed: its workers wait until the process ends
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned nw = std::min(7u, hw - 1);
        for (unsigned i = 0; i < nw; ++i) {
            P->workers.emplace_back([P, i] { P->work(i); });
            P->workers.back().detach();
        }
        return P;
    }();
    return *pool;
}

}  // namespace

void mpcf::detail::parallel_copy(void* dst, const void* src, const size_t n)
{
    copy_pool().copy((char*)dst, (const char*)src, n);
}

namespace {

// n bytes at device address src into host memory dst, ordered on stream s and
// complete on return
int download_host(hipStream_t s, void* dst, const void* src, const size_t n)
{
    if (n < (size_t(1) << 20)) {
        HIP_TRY(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        return MPCF_OK;
    }
    PinnedChunks& P = pinned_chunks();
    std::lock_guard<std::mutex> lk(P.mu);
    if (!P.buf[0]) {
        const size_t cap = size_t(64) << 20;
        char* b[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        hipError_t e = hipSuccess;
        for (int i = 0; i < 2 && e == hipSuccess; ++i) {
            e = hipHostMalloc((void**)&b[i], cap, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            for (int i = 0; i < 2; ++i) {
                if (b[i]) hipHostFree(b[i]);
                if (ev[i]) hipEventDestroy(ev[i]);
            }
            return fail(MPCF_EHIP, std::string("download staging: ") + hipGetErrorString(e));
        }
        for (int i = 0; i < 2; ++i) { P.buf[i] = b[i]; P.ev[i] = ev[i]; }
        P.cap = cap;
    }
    const size_t nchunk = (n + P.cap - 1) / P.cap;
    // chunk k goes to buffer k % 2; chunk k+1 is queued before chunk k is
    // copied out (the buffer it reuses was emptied in the previous round)
    auto issue = [&](const size_t k) -> hipError_t {
        const size_t off = k * P.cap, len = std::min(P.cap, n - off);
        hipError_t e = hipMemcpyAsync(P.buf[k & 1], (const char*)src + off, len, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(P.ev[k & 1], s);
        return e;
    };
    HIP_TRY(issue(0));
    for (size_t k = 0; k < nchunk; ++k) {
        if (k + 1 < nchunk) HIP_TRY(issue(k + 1));
        HIP_TRY(hipEventSynchronize(P.ev[k & 1]));
        const size_t off = k * P.cap;
        copy_pool().copy((char*)dst + off, P.buf[k & 1], std::min(P.cap, n - off));
    }
    return MPCF_OK;
}

int transfer(mpcf_ctx* c, const int which, void* const* fields, const int host_f64, const bool to_device,
             const char* what)
{
    if (!c || !fields) return fail(MPCF_EINVAL, std::string(what) + ": NULL argument");
    if (which != MPCF_DATA && which != MPCF_TMP) return fail(MPCF_EINVAL, std::string(what) + ": which must be DATA or TMP");
    for (int q = 0; q < 7; ++q)
        if (!fields[q]) return fail(MPCF_EINVAL, std::string(what) + ": NULL field pointer");
    const size_t n = (size_t)c->g.nx * c->g.ny * c->g.nz;
    const size_t hb = host_f64 ? sizeof(double) : sizeof(float);
    const size_t fb = (size_t)c->g.fstride * c->esize;
    char* blk = (char*)(which == MPCF_DATA ? c->u[c->cur] : c->R);
    // E of an E_PAIR solution: the pair (hi in field 4, lo in field 7)
    const bool pair = which == MPCF_DATA && (c->flags & MPCF_FLAG_E_PAIR);
    // the accumulator area stages dense fields.  Uploads stage all seven at
    // once (binary32 host data), or in groups of at least five (binary64 host
    // data into a binary32 block), with one wait per group; downloads go one
    // field at a time through download_host
    const size_t per = n * hb;
    const int group = to_device ? (int)std::min<size_t>(7, mpcf::NACC * fb / per) : 1;
    for (int q0 = 0; q0 < 7; q0 += group) {
        const int q1 = std::min(7, q0 + group);
        for (int q = q0; q < q1; ++q) {
            void* dq = blk + q * fb;
            char* st = (char*)c->acc + (size_t)(q - q0) * per;
            if (to_device) {
                HIP_TRY(hipMemcpyAsync(st, fields[q], per, hipMemcpyHostToDevice, c->stream));
                if (pair && q == 4) c->L->pack_epair(c->g, dq, blk + 7 * fb, st, 1, host_f64, c->stream);
                else                c->L->pack(c->g, dq, st, 1, 1, host_f64, c->stream);
                LAUNCH_CHECK("pack");
            } else {
                if (pair && q == 4) c->L->pack_epair(c->g, dq, blk + 7 * fb, st, 0, host_f64, c->stream);
                else                c->L->pack(c->g, dq, st, 0, 1, host_f64, c->stream);
                LAUNCH_CHECK("unpack");
                const int rc = download_host(c->stream, fields[q], st, per);
                if (rc != MPCF_OK) return rc;
            }
        }
        HIP_TRY(hipStreamSynchronize(c->stream));   // the staging area is reused by the next group
    }
    return MPCF_OK;
}

} // namespace

extern "C" int mpcf_upload(mpcf_ctx* c, int which, const float* const fields[7])
{
    if (c && which == MPCF_DATA) c->sos_folded = false;
    return transfer(c, which, (void* const*)fields, 0, true, "mpcf_upload");
}

extern "C" int mpcf_download(mpcf_ctx* c, int which, float* const fields[7])
{
    return transfer(c, which, (void* const*)fields, 0, false, "mpcf_download");
}

extern "C" int mpcf_upload_f64(mpcf_ctx* c, int which, const double* const fields[7])
{
    if (c && which == MPCF_DATA) c->sos_folded = false;
    return transfer(c, which, (void* const*)fields, 1, true, "mpcf_upload_f64");
}

extern "C" int mpcf_download_f64(mpcf_ctx* c, int which, double* const fields[7])
{
    return transfer(c, which, (void* const*)fields, 1, false, "mpcf_download_f64");
}

extern "C" int mpcf_stream_channels(int kind)
{
    return kind == MPCF_STREAM_TENSOR ? 9 : kind == MPCF_STREAM_SAVE ? 7 : kind == MPCF_STREAM_SCALAR ? 1 : 0;
}

extern "C" int mpcf_stream_out(mpcf_ctx* c, int kind, float* out)
{
    if (!c || !out) return fail(MPCF_EINVAL, "mpcf_stream_out: NULL argument");
    if (mpcf::detail::refuse_epair(c->flags, "mpcf_stream_out") != MPCF_OK) return MPCF_EINVAL;
    const int nc = mpcf_stream_channels(kind);
    if (!nc) return fail(MPCF_EINVAL, "mpcf_stream_out: unknown streamer kind");
    const size_t n = (size_t)c->g.nx * c->g.ny * c->g.nz;
    // binary32 storage streams through the bit-exact build's arithmetic
    const Launchers* L = (c->flags & MPCF_FLAG_FP64) ? c->L : &kParity;
    L->stream(c->g, c->u[c->cur], (float*)c->acc, kind, 1, c->stream);
    LAUNCH_CHECK("stream out");
    HIP_TRY(hipMemcpyAsync(out, c->acc, n * nc * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MPCF_OK;
}

extern "C" int mpcf_stream_in(mpcf_ctx* c, const float* in)
{
    if (!c || !in) return fail(MPCF_EINVAL, "mpcf_stream_in: NULL argument");
    if (mpcf::detail::refuse_epair(c->flags, "mpcf_stream_in") != MPCF_OK) return MPCF_EINVAL;
    c->sos_folded = false;
    const size_t n = (size_t)c->g.nx * c->g.ny * c->g.nz;
    const Launchers* L = (c->flags & MPCF_FLAG_FP64) ? c->L : &kParity;
    HIP_TRY(hipMemcpyAsync(c->acc, in, n * 7 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    L->stream(c->g, c->u[c->cur], (float*)c->acc, MPCF_STREAM_SAVE, 0, c->stream);
    LAUNCH_CHECK("stream in");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MPCF_OK;
}

extern "C" int mpcf_load_ghosts(mpcf_ctx* c)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_load_ghosts: NULL context");
    if (mpcf::detail::refuse_epair(c->flags, "mpcf_load_ghosts") != MPCF_OK) return MPCF_EINVAL;
    { Timed t(c, K_GHOSTS); c->L->ghosts(c->g, c->u[c->cur], c->bc, c->stream); }
    LAUNCH_CHECK("ghosts");
    return MPCF_OK;
}

extern "C" int mpcf_max_sos(mpcf_ctx* c, float* sos)
{
    if (!c || !sos) return fail(MPCF_EINVAL, "mpcf_max_sos: NULL argument");
    c->sos_folded = false;   // the full reduction clears the folded maximum
    {
        Timed t(c, K_SOS);
        c->L->sos(c->g, c->u[c->cur], c->st, (c->flags & MPCF_FLAG_SOS_CPU) ? 1 : 0, 0, 0.0, c->h, 0.0, c->stream);
    }
    LAUNCH_CHECK("sos");
    HIP_TRY(hipMemcpyAsync(c->st_host, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *sos = c->st_host->sos;
    if (c->st_host->bad) return fail(MPCF_ENUMERIC, std::string("mpcf_max_sos: ") + kBadCell);
    return MPCF_OK;
}

extern "C" int mpcf_convection(mpcf_ctx* c, float a, float dtinvh)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_convection: NULL context");
    c->sos_folded = false;
    run_rhs(c, c->u[c->cur], nullptr, a, 0.0, dtinvh, nullptr, 0);
    LAUNCH_CHECK("convection");
    return MPCF_OK;
}

extern "C" int mpcf_update(mpcf_ctx* c, float b)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_update: NULL context");
    c->sos_folded = false;
    { Timed t(c, K_UPDATE); c->L->update(c->g, c->u[c->cur], c->R, b, c->stream); }
    LAUNCH_CHECK("update");
    return MPCF_OK;
}

extern "C" int mpcf_process_all(mpcf_ctx* c, float a, float b, float dtinvh)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_process_all: NULL context");
    c->sos_folded = false;
    run_rhs(c, c->u[c->cur], c->u[1 - c->cur], a, b, dtinvh, nullptr, 1);
    LAUNCH_CHECK("process_all");
    c->cur = 1 - c->cur;
    return MPCF_OK;
}

namespace {

int step_args(mpcf_ctx* c, const double cfl, const double dt_max, const char* who)
{
    if (!c) return fail(MPCF_EINVAL, std::string(who) + ": NULL context");
    if (!(cfl > 0) || !std::isfinite(cfl)) return fail(MPCF_EINVAL, std::string(who) + ": cfl must be positive and finite");
    if (!(dt_max > 0)) return fail(MPCF_EINVAL, std::string(who) + ": dt_max must be positive");
    if (mpcf::detail::has_halo(c->bc))
        return fail(MPCF_EINVAL, std::string(who) + ": halo faces need an exchange before every stage; "
                                 "drive the stages with mpcf_max_sos and mpcf_process_all");
    return MPCF_OK;
}

// one LSRK3 step, queued.  keep_open: a pipelined FAST step may leave its
// last z sweeps unjoined (c->pipe_open) so that the next step's first x+y
// sweeps overlap them; the caller joins before it returns.
int step_queue(mpcf_ctx* c, const double cfl, const double dt_max, const bool keep_open)
{
    // the sweeps apply the boundary conditions as they load: no ghost fill
    const int slabs = pipeline_slabs(c);
    // pipelined FAST steps run dt on the second stream, ahead of the first
    // stage's z sweeps, while the first stage's x+y sweeps (which need no dt)
    // start on the context's stream as soon as their slab of the previous
    // step is done
    const bool defer = slabs && dt_deferred(c);
    const bool prev_open = defer && c->pipe_open == slabs;
    if (!prev_open) {
        const int rc = join_pipeline(c);
        if (rc != MPCF_OK) return rc;
    }
    hipStream_t sdt = c->stream;
    if (defer) {
        if (!c->stream2) HIP_TRY(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        if (!c->pipe_fork) HIP_TRY(hipEventCreateWithFlags(&c->pipe_fork, hipEventDisableTiming));
        // an open previous step orders the second stream already; otherwise
        // it follows what the context's stream queued (an upload, say)
        if (!prev_open) {
            HIP_TRY(hipEventRecord(c->pipe_fork, c->stream));
            HIP_TRY(hipStreamWaitEvent(c->stream2, c->pipe_fork, 0));
        }
        sdt = c->stream2;
    }
    {
        Timed t(c, K_SOS, sdt);
        // the maximum of this state was folded into the previous step's last
        // z sweep unless something else wrote u since: then only dt remains
        if (c->sos_folded) c->L->sos_folded(c->st, cfl, c->h, dt_max, sdt);
        else c->L->sos(c->g, c->u[c->cur], c->st, (c->flags & MPCF_FLAG_SOS_CPU) ? 1 : 0, 1, cfl, c->h, dt_max, sdt);
    }
    LAUNCH_CHECK("sos/dt");
    c->sos_folded = false;
    c->pipe_open = 0;
    // a FAST step that another step of the same call follows does not store
    // R in its last stage: the next step's first stage multiplies it by a = 0
    const bool skipR = keep_open && dt_deferred(c);
    if (slabs) {
        const int rc = run_stages_pipelined(c, slabs, prev_open, defer && keep_open, skipR);
        if (rc != MPCF_OK) return rc;
    } else {
        for (int s = 0; s < 3; ++s) {
            run_rhs(c, c->u[c->cur], c->u[1 - c->cur], kA[s], kB[s], 0.0, c->st, (s == 2 && skipR) ? 2 : 1,
                    s == 2 ? c->st : nullptr);
            LAUNCH_CHECK("stage");
            c->cur = 1 - c->cur;
        }
    }
    c->sos_folded = true;
    return MPCF_OK;
}

}  // namespace

extern "C" int mpcf_step(mpcf_ctx* c, double cfl, double dt_max, double* dt)
{
    int rc = step_args(c, cfl, dt_max, "mpcf_step");
    if (rc != MPCF_OK) return rc;
    rc = step_queue(c, cfl, dt_max, false);
    if (rc != MPCF_OK) return rc;
    if (dt) {
        double d = 0;
        float sos = 0;
        rc = mpcf_last_dt(c, &d, &sos);
        *dt = d;
        if (rc != MPCF_OK) return rc;
        if (!(d > 0) || !std::isfinite(d) || !(sos > 0) || !std::isfinite(sos))
            return fail(MPCF_ENUMERIC, "mpcf_step: non-finite or non-positive dt / speed of sound");
    }
    return MPCF_OK;
}

extern "C" int mpcf_steps(mpcf_ctx* c, int nsteps, double cfl, double dt_max)
{
    int rc = step_args(c, cfl, dt_max, "mpcf_steps");
    if (rc != MPCF_OK) return rc;
    if (nsteps < 0) return fail(MPCF_EINVAL, "mpcf_steps: nsteps must be >= 0");
    for (int k = 0; k < nsteps; ++k) {
        rc = step_queue(c, cfl, dt_max, k + 1 < nsteps);
        if (rc != MPCF_OK) {
            join_pipeline(c);
            return rc;
        }
    }
    return join_pipeline(c);
}

extern "C" int mpcf_last_dt(mpcf_ctx* c, double* dt, float* sos)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_last_dt: NULL context");
    HIP_TRY(hipMemcpyAsync(c->st_host, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const StepState& h = *c->st_host;
    if (dt) *dt = h.dt;
    if (sos) *sos = h.sos;
    // the state dt was taken from, and (after a step, which folds it) the state
    // the step wrote
    if (h.bad) return fail(MPCF_ENUMERIC, std::string("mpcf_last_dt: the state the time step was taken from ") + kBadCell);
    if ((c->sos_folded || h.folded) && h.next_bad)
        return fail(MPCF_ENUMERIC, std::string("mpcf_last_dt: the state the last step wrote ") + kBadCell);
    return MPCF_OK;
}

extern "C" int mpcf_last_sos64(mpcf_ctx* c, double* sos)
{
    if (!c || !sos) return fail(MPCF_EINVAL, "mpcf_last_sos64: NULL argument");
    HIP_TRY(hipMemcpyAsync(c->st_host, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *sos = c->st_host->sos_d;
    return MPCF_OK;
}

extern "C" int mpcf_compare(mpcf_ctx* a, mpcf_ctx* b, int which, double out[28])
{
    if (!a || !b || !out) return fail(MPCF_EINVAL, "mpcf_compare: NULL argument");
    if (which != MPCF_DATA && which != MPCF_TMP) return fail(MPCF_EINVAL, "mpcf_compare: which must be DATA or TMP");
    if (a->g.nx != b->g.nx || a->g.ny != b->g.ny || a->g.nz != b->g.nz)
        return fail(MPCF_EINVAL, "mpcf_compare: blocks differ in extent");
    if (a->device != b->device) return fail(MPCF_EINVAL, "mpcf_compare: blocks live on different devices");
    HIP_TRY(hipStreamSynchronize(b->stream));
    const void* ua = which == MPCF_DATA ? a->u[a->cur] : a->R;
    const void* ub = which == MPCF_DATA ? b->u[b->cur] : b->R;
    // the accumulator area of `a` holds the partials; the result follows them
    double* scratch = (double*)a->acc;
    double* res = scratch + mpcf::reduce::scratch_doubles(a->g);
    const int ap = which == MPCF_DATA && (a->flags & MPCF_FLAG_E_PAIR) ? 1 : 0;
    const int bp = which == MPCF_DATA && (b->flags & MPCF_FLAG_E_PAIR) ? 1 : 0;
    mpcf::reduce::diff(a->g, ua, a->esize == 8, a->g.fstride, ap, ub, b->esize == 8, b->g.fstride, bp, scratch, res,
                       a->stream);
    LAUNCH_CHECK("compare");
    HIP_TRY(hipMemcpyAsync(out, res, 28 * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    HIP_TRY(hipStreamSynchronize(a->stream));
    return MPCF_OK;
}

extern "C" int mpcf_integrals(mpcf_ctx* c, int which, double out[7])
{
    if (!c || !out) return fail(MPCF_EINVAL, "mpcf_integrals: NULL argument");
    if (which != MPCF_DATA && which != MPCF_TMP) return fail(MPCF_EINVAL, "mpcf_integrals: which must be DATA or TMP");
    const void* u = which == MPCF_DATA ? c->u[c->cur] : c->R;
    double* scratch = (double*)c->acc;
    double* res = scratch + mpcf::reduce::scratch_doubles(c->g);
    const int pair = which == MPCF_DATA && (c->flags & MPCF_FLAG_E_PAIR) ? 1 : 0;
    mpcf::reduce::sums(c->g, u, c->esize == 8, c->g.fstride, pair, scratch, res, c->stream);
    LAUNCH_CHECK("integrals");
    double sums[7];
    HIP_TRY(hipMemcpyAsync(sums, res, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double vol = c->h * c->h * c->h;
    for (int q = 0; q < 7; ++q) out[q] = sums[q] * vol;
    return MPCF_OK;
}

extern "C" int mpcf_synchronize(mpcf_ctx* c)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_synchronize: NULL context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MPCF_OK;
}

extern "C" int mpcf_query(mpcf_ctx* c, long long out[8])
{
    if (!c || !out) return fail(MPCF_EINVAL, "mpcf_query: NULL argument");
    out[0] = c->g.nx; out[1] = c->g.ny; out[2] = c->g.nz;
    out[3] = c->g.px; out[4] = c->g.py; out[5] = c->g.pz;
    out[6] = (long long)c->g.fstride;
    out[7] = (long long)c->bytes;
    return MPCF_OK;
}

extern "C" int mpcf_timing_enable(mpcf_ctx* c, int enable)
{
    if (!c) return fail(MPCF_EINVAL, "mpcf_timing_enable: NULL context");
    c->timing = enable != 0;
    return MPCF_OK;
}

extern "C" int mpcf_timing_read(mpcf_ctx* c, double ms[8], long long count[8])
{
    if (!c || !ms || !count) return fail(MPCF_EINVAL, "mpcf_timing_read: NULL argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 8; ++k) { ms[k] = 0; count[k] = 0; }
    for (auto& m : c->marks) {
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, m.b, m.e));
        const int k = (m.kind >= 0 && m.kind < 8) ? m.kind : K_OTHER;
        ms[k] += t;
        count[k] += 1;
        c->ev_pool.push_back(m.b);
        c->ev_pool.push_back(m.e);
    }
    c->marks.clear();
    return MPCF_OK;
}

namespace {

struct DevBuf
{
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};

} // namespace

extern "C" int mpcf_test_weno(int n, const float* in5, float* minus, float* plus, unsigned flags)
{
    if (n < 0 || (n > 0 && (!in5 || !minus || !plus))) return fail(MPCF_EINVAL, "mpcf_test_weno: bad argument");
    if (n == 0) return MPCF_OK;
    if (flags & MPCF_FLAG_FP64) return fail(MPCF_EINVAL, "mpcf_test_weno: binary32 arrays only");
    const Launchers* L = pick_launchers(flags);
    DevBuf a, m, p;
    HIP_TRY(hipMalloc(&a.p, sizeof(float) * 5 * (size_t)n));
    HIP_TRY(hipMalloc(&m.p, sizeof(float) * (size_t)n));
    HIP_TRY(hipMalloc(&p.p, sizeof(float) * (size_t)n));
    HIP_TRY(hipMemcpy(a.p, in5, sizeof(float) * 5 * (size_t)n, hipMemcpyHostToDevice));
    L->test_weno(n, a.p, m.p, p.p, nullptr);
    LAUNCH_CHECK("test_weno");
    HIP_TRY(hipMemcpy(minus, m.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(plus, p.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return MPCF_OK;
}

extern "C" int mpcf_test_face(int n, int dir, const float* st, float* out, unsigned flags)
{
    if (n < 0 || dir < 0 || dir > 2 || (n > 0 && (!st || !out))) return fail(MPCF_EINVAL, "mpcf_test_face: bad argument");
    if (n == 0) return MPCF_OK;
    if (flags & MPCF_FLAG_FP64) return fail(MPCF_EINVAL, "mpcf_test_face: binary32 arrays only");
    const Launchers* L = pick_launchers(flags);
    DevBuf a, o;
    HIP_TRY(hipMalloc(&a.p, sizeof(float) * 42 * (size_t)n));
    HIP_TRY(hipMalloc(&o.p, sizeof(float) * 12 * (size_t)n));
    HIP_TRY(hipMemcpy(a.p, st, sizeof(float) * 42 * (size_t)n, hipMemcpyHostToDevice));
    L->test_face(n, dir, a.p, o.p, nullptr);
    LAUNCH_CHECK("test_face");
    HIP_TRY(hipMemcpy(out, o.p, sizeof(float) * 12 * (size_t)n, hipMemcpyDeviceToHost));
    return MPCF_OK;
}

// ===========================================================================
// Batched ensembles (mpcf_batch_*): many independent cases of uneven sizes
// advanced together, one launch per kernel kind per stage, per-case dt on the
// device, a whole step optionally captured once as a hipGraph and replayed.
// Each case is a full context (its own resident block); the batch owns the
// device tables that the batch kernels read.
// ===========================================================================
#include <algorithm>
#include <functional>
#include <map>

struct mpcf_batch
{
    unsigned flags;
    const Launchers* L;
    hipStream_t stream;
    std::vector<mpcf_ctx*> cases;
    std::vector<int> ids;
    std::vector<int> chunk_xy, chunk_z;      // per case; 0 = default
    int next_id = 1;
    // device tables
    mpcf::CaseDesc* d_cases = nullptr;
    int* d_cur = nullptr;
    int* d_off = nullptr;                    // kBatchGroups x 3 x (cap + 1)
    int cap = 0;
    bool dirty = true;
    bool cur_on_device = false;              // d_cur holds the truth (after steps)
    // the cases in kBatchGroups groups of about equal cells, each stepped on
    // its own stream (groups 1.. on side streams forked from and joined back
    // to `stream`), so that one group's x+y sweep runs beside another's z sweep
    int tot_xy[kBatchGroups] = {}, tot_z[kBatchGroups] = {}, tot_sos[kBatchGroups] = {};
    mpcf::BatchDev dev[kBatchGroups]{};
    hipStream_t side[kBatchGroups] = {};
    hipEvent_t fork = nullptr, join[kBatchGroups] = {};
    // graph of one step
    bool use_graph = false;
    bool graph_valid = false;
    double g_cfl = 0, g_dtmax = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long captures = 0;
    long long steps = 0;
    // autotuned march lengths of the x+y sweep by NY and of the z sweep by NZ
    std::map<int, int> tuned_xy, tuned_z;
    // how a step is launched: MPCF_BATCH_MODE_PERSISTENT -- per stage, two
    // persistent launches (x+y items, then z items) whose workgroups pull
    // (case, tile) items from a device queue, all on the batch's stream;
    // MPCF_BATCH_MODE_GROUPS -- one flat launch per kernel kind and case group,
    // the groups on their own streams
    int mode = MPCF_BATCH_MODE_GROUPS;
    int p_threads = 256;                     // workgroup size of the persistent kernel
    int p_blocks[2] = {0, 0};                // grids of the x+y and the z launch
    mpcf::PSeg* d_seg = nullptr;             // queue segments (2 per case + 1)
    unsigned* d_done = nullptr;              // per case: finished z items of the running stage
    mpcf::PCtl* d_ctl = nullptr;             // queue head, workgroups done
    int p_cap = 0;
    mpcf::PersistDev pdev{};
};

namespace {

void batch_drop_graph(mpcf_batch* b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

// bring the device-side buffer indices back into the contexts
int batch_pull_cur(mpcf_batch* b)
{
    if (!b->cur_on_device || b->cases.empty()) return MPCF_OK;
    std::vector<int> cur(b->cases.size());
    HIP_TRY(hipMemcpyAsync(cur.data(), b->d_cur, cur.size() * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < cur.size(); ++i) b->cases[i]->cur = cur[i];
    b->cur_on_device = false;
    return MPCF_OK;
}

// march length of case i in the x+y (which 0) or z sweep; `columns` is the
// marching columns of all cases in that sweep
int march_len(const mpcf_batch* b, int i, int nd, int which, long long columns)
{
    const int forced = which == 0 ? b->chunk_xy[i] : b->chunk_z[i];
    const std::map<int, int>& tuned_map = which == 0 ? b->tuned_xy : b->tuned_z;
    auto it = tuned_map.find(nd);
    const int tuned = it != tuned_map.end() ? it->second : 0;
    if (forced > 0 || tuned > 0) return mpcf::detail::batch_march(nd, forced, tuned);
    return mpcf::detail::adaptive_march(nd, columns);
}

// the queues of the persistent stage launches: the x+y items of every case,
// largest case first, then their z items in the same order
int persist_tables(mpcf_batch* b, std::vector<mpcf::CaseDesc>& desc)
{
    const int n = (int)desc.size();
    if (n > b->p_cap) {
        if (b->d_seg) hipFree(b->d_seg);
        if (b->d_done) hipFree(b->d_done);
        if (b->d_ctl) hipFree(b->d_ctl);
        b->d_seg = nullptr; b->d_done = nullptr; b->d_ctl = nullptr;
        const int cap = std::max(16, 2 * n);
        hipError_t e = hipMalloc((void**)&b->d_seg, (2 * cap + 1) * sizeof(mpcf::PSeg));
        if (e == hipSuccess) e = hipMalloc((void**)&b->d_done, cap * sizeof(unsigned));
        if (e == hipSuccess) e = hipMalloc((void**)&b->d_ctl, sizeof(mpcf::PCtl));
        if (e != hipSuccess) return fail(MPCF_ENOMEM, std::string("batch queue tables: ") + hipGetErrorString(e));
        b->p_cap = cap;
    }
    const int W = b->p_threads / 64;
    for (mpcf::CaseDesc& d : desc) mpcf::detail::case_items(d, W);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto cells = [&](int i) { return (long long)desc[i].g.nx * desc[i].g.ny * desc[i].g.nz; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cells(x) > cells(y); });
    std::vector<mpcf::PSeg> seg;
    unsigned total = 0, items[2] = {0, 0};
    for (int kind = 0; kind < 2; ++kind) {
        const unsigned start = total;
        for (int k = 0; k < n; ++k) {
            const int c = order[k];
            seg.push_back({c, kind, total});
            total += kind == 0 ? desc[c].items_xy : desc[c].items_z;
        }
        items[kind] = total - start;
    }
    const int nseg = (int)seg.size();
    seg.push_back({-1, 0, total});
    HIP_TRY(hipMemcpyAsync(b->d_seg, seg.data(), seg.size() * sizeof(mpcf::PSeg), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemsetAsync(b->d_done, 0, n * sizeof(unsigned), b->stream));
    HIP_TRY(hipMemsetAsync(b->d_ctl, 0, sizeof(mpcf::PCtl), b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->cases[0]->device));
    for (int ph = 0; ph < 2; ++ph) {
        const int occ = std::max(1, b->L->persist_occ(b->p_threads, ph));
        b->p_blocks[ph] = (int)std::max(1LL, std::min((long long)occ * cus, (long long)items[ph]));
    }
    mpcf::PersistDev& P = b->pdev;
    P.cases = b->d_cases; P.cur = b->d_cur; P.seg = b->d_seg; P.nseg = nseg; P.nxy = n;
    P.done = b->d_done; P.ctl = b->d_ctl;
    P.nblocks = 0;
    return MPCF_OK;
}

int batch_rebuild(mpcf_batch* b)
{
    const int rc = batch_pull_cur(b);
    if (rc != MPCF_OK) return rc;
    const int n = (int)b->cases.size();
    if (n > b->cap) {
        if (b->d_cases) hipFree(b->d_cases);
        if (b->d_cur) hipFree(b->d_cur);
        if (b->d_off) hipFree(b->d_off);
        b->d_cases = nullptr; b->d_cur = nullptr; b->d_off = nullptr;
        const int cap = std::max(16, 2 * n);
        hipError_t e = hipMalloc((void**)&b->d_cases, cap * sizeof(mpcf::CaseDesc));
        if (e == hipSuccess) e = hipMalloc((void**)&b->d_cur, cap * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void**)&b->d_off, kBatchGroups * 3 * (cap + 1) * sizeof(int));
        if (e != hipSuccess) return fail(MPCF_ENOMEM, std::string("batch tables: ") + hipGetErrorString(e));
        b->cap = cap;
    }
    std::vector<mpcf::CaseDesc> desc(n);
    const int tab = 3 * (b->cap + 1);
    std::vector<int> cur(n), off(kBatchGroups * tab, 0);
    long long cols_xy = 0, cols_z = 0;
    for (mpcf_ctx* c : b->cases) {
        cols_xy += (long long)(c->g.nx + 6) * c->g.nz;
        cols_z += (long long)c->g.nx * c->g.ny;
    }
    int first[kBatchGroups + 1];
    {
        std::vector<long long> cc(n);
        for (int i = 0; i < n; ++i) cc[i] = (long long)b->cases[i]->g.nx * b->cases[i]->g.ny * b->cases[i]->g.nz;
        split_groups(cc.data(), n, kBatchGroups, first);
    }
    for (int i = 0; i < n; ++i) {
        mpcf_ctx* c = b->cases[i];
        int grp = 0;
        while (i >= first[grp + 1]) ++grp;
        const int j = i - first[grp];                                      // index within the group
        int* oxy = &off[grp * tab];
        int* oz = oxy + (b->cap + 1);
        int* os = oxy + 2 * (b->cap + 1);
        int blocks[3];
        mpcf::detail::case_desc(c, c->cur, march_len(b, i, c->g.ny, 0, cols_xy), march_len(b, i, c->g.nz, 1, cols_z),
                                desc[i], blocks);
        oxy[j + 1] = oxy[j] + blocks[0];
        oz[j + 1] = oz[j] + blocks[1];
        os[j + 1] = os[j] + blocks[2];
        cur[i] = c->cur;
    }
    if (b->mode == MPCF_BATCH_MODE_PERSISTENT && n > 0) {
        const int prc = persist_tables(b, desc);   // fills the descriptors' item counts
        if (prc != MPCF_OK) return prc;
    }
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(b->d_cases, desc.data(), n * sizeof(mpcf::CaseDesc), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_cur, cur.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
    }
    for (int grp = 0; grp < kBatchGroups; ++grp) {
        const int count = first[grp + 1] - first[grp];
        const int* ho = &off[grp * tab];
        b->tot_xy[grp] = ho[count]; b->tot_z[grp] = ho[(b->cap + 1) + count]; b->tot_sos[grp] = ho[2 * (b->cap + 1) + count];
        mpcf::BatchDev& d = b->dev[grp];
        d.cases = b->d_cases + first[grp]; d.cur = b->d_cur + first[grp]; d.n = count;
        d.off_xy = b->d_off + grp * tab; d.off_z = d.off_xy + (b->cap + 1); d.off_sos = d.off_xy + 2 * (b->cap + 1);
    }
    b->dirty = false;
    batch_drop_graph(b);
    return MPCF_OK;
}

int batch_enqueue_step(mpcf_batch* b, double cfl, double dt_max)
{
    const int cpu_form = (b->flags & MPCF_FLAG_SOS_CPU) ? 1 : 0;
    if (b->mode == MPCF_BATCH_MODE_PERSISTENT) {
        // dt of every case (from the maximum the previous step folded, or a
        // full reduction), then one persistent launch per stage
        for (int grp = 0; grp < kBatchGroups; ++grp)
            if (b->dev[grp].n > 0) b->L->b_sos(b->dev[grp], b->tot_sos[grp], cpu_form, cfl, dt_max, b->stream);
        for (int s = 0; s < 3; ++s)
            b->L->persist(b->pdev, b->p_blocks, b->p_threads, kA[s], kB[s], s == 2, cpu_form, b->stream);
        return MPCF_OK;
    }
    // the sweeps apply each case's boundary conditions as they load: a step is
    // seven launches per group; groups 1.. run on side streams
    bool forked = false;
    for (int grp = kBatchGroups - 1; grp >= 0; --grp) {
        const mpcf::BatchDev& d = b->dev[grp];
        if (d.n == 0) continue;
        hipStream_t st = b->stream;
        if (grp > 0) {
            if (!b->side[grp]) HIP_TRY(hipStreamCreateWithFlags(&b->side[grp], hipStreamNonBlocking));
            if (!b->join[grp]) HIP_TRY(hipEventCreateWithFlags(&b->join[grp], hipEventDisableTiming));
            if (!b->fork) HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
            if (!forked) HIP_TRY(hipEventRecord(b->fork, b->stream));
            forked = true;
            st = b->side[grp];
            HIP_TRY(hipStreamWaitEvent(st, b->fork, 0));
        }
        b->L->b_sos(d, b->tot_sos[grp], cpu_form, cfl, dt_max, st);
        for (int s = 0; s < 3; ++s) {
            b->L->b_xy(d, b->tot_xy[grp], kA[s], kB[s], st);
            b->L->b_z(d, b->tot_z[grp], kA[s], kB[s], s == 2, cpu_form, 1, st);
        }
        if (grp > 0) HIP_TRY(hipEventRecord(b->join[grp], st));
    }
    for (int grp = 1; grp < kBatchGroups; ++grp)
        if (b->dev[grp].n > 0) HIP_TRY(hipStreamWaitEvent(b->stream, b->join[grp], 0));
    return MPCF_OK;
}

int find_case_index(const mpcf_batch* b, int id)
{
    for (size_t i = 0; i < b->ids.size(); ++i)
        if (b->ids[i] == id) return (int)i;
    return -1;
}


} // namespace

extern "C" int mpcf_case_bytes(int nx, int ny, int nz, long long* bytes)
{
    if (!bytes || nx < 5 || ny < 5 || nz < 5) return fail(MPCF_EINVAL, "mpcf_case_bytes: bad argument");
    const Geom g = mpcf::make_geom(nx, ny, nz);
    *bytes = (long long)((7 + 7 + 7 + mpcf::NACC) * (size_t)g.fstride * sizeof(float) + 256);
    return MPCF_OK;
}

extern "C" int mpcf_batch_create(mpcf_batch** out, unsigned flags)
{
    if (!out) return fail(MPCF_EINVAL, "mpcf_batch_create: out is NULL");
    *out = nullptr;
    if (flags & ~(unsigned)(MPCF_FLAG_FAST | MPCF_FLAG_SOS_CPU | MPCF_FLAG_P64 | MPCF_FLAG_FP64))
        return fail(MPCF_EINVAL, "mpcf_batch_create: unknown flag");
    if ((flags & MPCF_FLAG_FP64) && (flags & (MPCF_FLAG_FAST | MPCF_FLAG_P64)))
        return fail(MPCF_EINVAL, "mpcf_batch_create: FP64 excludes FAST and P64");
    if (flags & MPCF_FLAG_P64) flags |= MPCF_FLAG_FAST;
    mpcf_batch* b = new mpcf_batch();
    b->flags = flags;
    b->L = pick_launchers(flags);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return fail(MPCF_EHIP, std::string("mpcf_batch_create: ") + hipGetErrorString(e)); }
    *out = b;
    return MPCF_OK;
}

extern "C" int mpcf_batch_destroy(mpcf_batch* b)
{
    if (!b) return MPCF_OK;
    hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (mpcf_ctx* c : b->cases) mpcf_destroy(c);
    if (b->d_cases) hipFree(b->d_cases);
    if (b->d_cur) hipFree(b->d_cur);
    if (b->d_off) hipFree(b->d_off);
    if (b->d_seg) hipFree(b->d_seg);
    if (b->d_done) hipFree(b->d_done);
    if (b->d_ctl) hipFree(b->d_ctl);
    for (int grp = 1; grp < kBatchGroups; ++grp)
    {
        if (b->side[grp]) {
            hipStreamSynchronize(b->side[grp]);
            hipStreamDestroy(b->side[grp]);
        }
        if (b->join[grp]) hipEventDestroy(b->join[grp]);
    }
    if (b->fork) hipEventDestroy(b->fork);
    hipStreamDestroy(b->stream);
    delete b;
    return MPCF_OK;
}

extern "C" int mpcf_batch_add(mpcf_batch* b, int nx, int ny, int nz, double h, const int bc[6],
                              const float* const fields[7], int* case_id)
{
    if (!b || !fields || !case_id) return fail(MPCF_EINVAL, "mpcf_batch_add: NULL argument");
    if (bc && mpcf::detail::has_halo(bc))
        return fail(MPCF_EINVAL, "mpcf_batch_add: a batched case is a whole domain (no halo faces)");
    int rc = batch_pull_cur(b);
    if (rc != MPCF_OK) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    mpcf_ctx* c = nullptr;
    rc = mpcf_create(&c, nx, ny, nz, h, b->flags);
    if (rc != MPCF_OK) return rc;
    if (bc) {
        rc = mpcf_set_bc(c, bc);
        if (rc != MPCF_OK) { mpcf_destroy(c); return rc; }
    }
    rc = mpcf_upload(c, MPCF_DATA, fields);
    if (rc != MPCF_OK) { mpcf_destroy(c); return rc; }
    b->cases.push_back(c);
    b->ids.push_back(b->next_id);
    b->chunk_xy.push_back(0);
    b->chunk_z.push_back(0);
    *case_id = b->next_id++;
    b->dirty = true;
    return MPCF_OK;
}

extern "C" int mpcf_batch_remove(mpcf_batch* b, int case_id)
{
    if (!b) return fail(MPCF_EINVAL, "mpcf_batch_remove: NULL batch");
    const int i = find_case_index(b, case_id);
    if (i < 0) return fail(MPCF_EINVAL, "mpcf_batch_remove: unknown case id");
    int rc = batch_pull_cur(b);
    if (rc != MPCF_OK) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    mpcf_destroy(b->cases[i]);
    b->cases.erase(b->cases.begin() + i);
    b->ids.erase(b->ids.begin() + i);
    b->chunk_xy.erase(b->chunk_xy.begin() + i);
    b->chunk_z.erase(b->chunk_z.begin() + i);
    b->dirty = true;
    return MPCF_OK;
}

extern "C" int mpcf_batch_download(mpcf_batch* b, int case_id, int which, float* const fields[7])
{
    if (!b) return fail(MPCF_EINVAL, "mpcf_batch_download: NULL batch");
    const int i = find_case_index(b, case_id);
    if (i < 0) return fail(MPCF_EINVAL, "mpcf_batch_download: unknown case id");
    HIP_TRY(hipStreamSynchronize(b->stream));
    const int rc = batch_pull_cur(b);
    if (rc != MPCF_OK) return rc;
    return mpcf_download(b->cases[i], which, fields);
}

extern "C" int mpcf_batch_set_march(mpcf_batch* b, int case_id, int len_xy, int len_z)
{
    if (!b || len_xy < 0 || len_z < 0) return fail(MPCF_EINVAL, "mpcf_batch_set_march: bad argument");
    const int i = find_case_index(b, case_id);
    if (i < 0) return fail(MPCF_EINVAL, "mpcf_batch_set_march: unknown case id");
    b->chunk_xy[i] = len_xy;
    b->chunk_z[i] = len_z;
    b->dirty = true;
    return MPCF_OK;
}

extern "C" int mpcf_batch_use_graph(mpcf_batch* b, int enable)
{
    if (!b) return fail(MPCF_EINVAL, "mpcf_batch_use_graph: NULL batch");
    b->use_graph = enable != 0;
    if (!b->use_graph) batch_drop_graph(b);
    return MPCF_OK;
}

extern "C" int mpcf_batch_step(mpcf_batch* b, double cfl, double dt_max)
{
    if (!b) return fail(MPCF_EINVAL, "mpcf_batch_step: NULL batch");
    if (!(cfl > 0) || !std::isfinite(cfl) || !(dt_max > 0)) return fail(MPCF_EINVAL, "mpcf_batch_step: bad cfl/dt_max");
    if (b->dirty) {
        const int rc = batch_rebuild(b);
        if (rc != MPCF_OK) return rc;
    }
    if (b->cases.empty()) return MPCF_OK;
    if (b->use_graph) {
        if (!b->graph_valid || b->g_cfl != cfl || b->g_dtmax != dt_max) {
            batch_drop_graph(b);
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const int rc = batch_enqueue_step(b, cfl, dt_max);
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamEndCapture(b->stream, &g));
            if (rc != MPCF_OK) { if (g) hipGraphDestroy(g); return rc; }
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_valid = true;
            b->g_cfl = cfl;
            b->g_dtmax = dt_max;
            ++b->captures;
        }
        HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
        const int rc = batch_enqueue_step(b, cfl, dt_max);
        if (rc != MPCF_OK) return rc;
        LAUNCH_CHECK("batch step");
    }
    b->cur_on_device = true;
    ++b->steps;
    return MPCF_OK;
}

extern "C" int mpcf_batch_last_dt(mpcf_batch* b, int case_id, double* dt, float* sos)
{
    if (!b) return fail(MPCF_EINVAL, "mpcf_batch_last_dt: NULL batch");
    const int i = find_case_index(b, case_id);
    if (i < 0) return fail(MPCF_EINVAL, "mpcf_batch_last_dt: unknown case id");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return mpcf_last_dt(b->cases[i], dt, sos);
}

extern "C" int mpcf_batch_synchronize(mpcf_batch* b)
{
    if (!b) return fail(MPCF_EINVAL, "mpcf_batch_synchronize: NULL batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return MPCF_OK;
}

extern "C" void* mpcf_batch_get_stream(mpcf_batch* b) { return b ? (void*)b->stream : nullptr; }

extern "C" int mpcf_batch_query(mpcf_batch* b, long long out[6])
{
    if (!b || !out) return fail(MPCF_EINVAL, "mpcf_batch_query: NULL argument");
    long long cells = 0, bytes = 0;
    for (mpcf_ctx* c : b->cases) {
        cells += (long long)c->g.nx * c->g.ny * c->g.nz;
        bytes += (long long)c->bytes;
    }
    out[0] = (long long)b->cases.size();
    out[1] = cells;
    out[2] = bytes;
    out[3] = b->captures;
    out[4] = b->steps;
    std::map<int, int> ext(b->tuned_xy);
    ext.insert(b->tuned_z.begin(), b->tuned_z.end());
    out[5] = (long long)ext.size();
    return MPCF_OK;
}

extern "C" int mpcf_batch_set_mode(mpcf_batch* b, int mode, int threads)
{
    if (!b) return fail(MPCF_EINVAL, "mpcf_batch_set_mode: NULL batch");
    if (mode != MPCF_BATCH_MODE_GROUPS && mode != MPCF_BATCH_MODE_PERSISTENT)
        return fail(MPCF_EINVAL, "mpcf_batch_set_mode: mode must be MPCF_BATCH_MODE_GROUPS or MPCF_BATCH_MODE_PERSISTENT");
    if (threads == 0) threads = 256;
    if (threads != 64 && threads != 128 && threads != 256)
        return fail(MPCF_EINVAL, "mpcf_batch_set_mode: threads must be 64, 128 or 256");
    if (b->mode != mode || b->p_threads != threads) {
        const int rc = batch_pull_cur(b);
        if (rc != MPCF_OK) return rc;
        b->mode = mode;
        b->p_threads = threads;
        b->dirty = true;
    }
    return MPCF_OK;
}

extern "C" int mpcf_batch_tuning(mpcf_batch* b, int extent, int out[4])
{
    if (!b || !out) return fail(MPCF_EINVAL, "mpcf_batch_tuning: NULL argument");
    out[0] = b->mode;
    out[1] = b->p_threads;
    auto ixy = b->tuned_xy.find(extent);
    auto iz = b->tuned_z.find(extent);
    out[2] = ixy != b->tuned_xy.end() ? ixy->second : 0;
    out[3] = iz != b->tuned_z.end() ? iz->second : 0;
    return MPCF_OK;
}

// Autotuner: the launch mode and the persistent workgroup size, then per case
// extent the two march lengths (how many cells one wave marches in the x+y
// and z sweeps before the next chunk takes over), by coordinate descent.
// Every candidate is timed over `steps` batch steps after a warm-up step,
// from a saved state that is restored afterwards, so tuning does not advance
// the cases.
namespace {

struct TuneState
{
    mpcf_batch* b;
    double cfl;
    int steps;
    std::vector<void*> save;
    std::vector<int> cur0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

int tune_save(TuneState& T)
{
    mpcf_batch* b = T.b;
    for (mpcf_ctx* c : b->cases) {
        void* p = nullptr;
        const size_t reg = 7 * (size_t)c->g.fstride * c->esize;   // one 7-field register
        const hipError_t e = hipMalloc(&p, 2 * reg);
        if (e != hipSuccess) return fail(MPCF_ENOMEM, "mpcf_batch_autotune: no room to save the state");
        T.save.push_back(p);
        HIP_TRY(hipMemcpyAsync(p, c->u[c->cur], reg, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync((char*)p + reg, c->R, reg, hipMemcpyDeviceToDevice, b->stream));
        T.cur0.push_back(c->cur);
    }
    HIP_TRY(hipEventCreate(&T.e0));
    HIP_TRY(hipEventCreate(&T.e1));
    return MPCF_OK;
}

int tune_restore(TuneState& T)
{
    mpcf_batch* b = T.b;
    int rc = batch_pull_cur(b);
    for (size_t i = 0; i < b->cases.size(); ++i) {
        mpcf_ctx* c = b->cases[i];
        c->cur = T.cur0[i];
        const size_t reg = 7 * (size_t)c->g.fstride * c->esize;
        HIP_TRY(hipMemcpyAsync(c->u[c->cur], T.save[i], reg, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(c->R, (char*)T.save[i] + reg, reg, hipMemcpyDeviceToDevice, b->stream));
        // the folded speed of sound belongs to the trial state: reduce again
        HIP_TRY(hipMemsetAsync(&c->st->folded, 0, sizeof(int), b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return rc;
}

void tune_free(TuneState& T)
{
    for (void* q : T.save) hipFree(q);
    if (T.e0) hipEventDestroy(T.e0);
    if (T.e1) hipEventDestroy(T.e1);
}

// milliseconds of `steps` steps in the batch's current configuration
int tune_trial(TuneState& T, float* ms)
{
    mpcf_batch* b = T.b;
    b->dirty = true;
    int rc = mpcf_batch_step(b, T.cfl, 1e30);        // warm-up in this configuration
    if (rc == MPCF_OK) {
        HIP_TRY(hipEventRecord(T.e0, b->stream));
        for (int k = 0; k < T.steps && rc == MPCF_OK; ++k) rc = mpcf_batch_step(b, T.cfl, 1e30);
        HIP_TRY(hipEventRecord(T.e1, b->stream));
        HIP_TRY(hipEventSynchronize(T.e1));
        HIP_TRY(hipEventElapsedTime(ms, T.e0, T.e1));
    }
    const int rc2 = tune_restore(T);
    return rc != MPCF_OK ? rc : rc2;
}

}  // namespace

extern "C" int mpcf_batch_autotune(mpcf_batch* b, int steps, double cfl, int* best_len)
{
    if (!b || steps < 1 || !(cfl > 0)) return fail(MPCF_EINVAL, "mpcf_batch_autotune: bad argument");
    if (b->cases.empty()) return MPCF_OK;
    int rc = batch_pull_cur(b);
    if (rc != MPCF_OK) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    TuneState T{b, cfl, steps};
    rc = tune_save(T);
    const bool graph = b->use_graph;
    b->use_graph = false;
    batch_drop_graph(b);
    // 1. launch mode and workgroup size, with the lengths tuned so far
    const int modes[][2] = {{MPCF_BATCH_MODE_PERSISTENT, 256}, {MPCF_BATCH_MODE_PERSISTENT, 128},
                            {MPCF_BATCH_MODE_PERSISTENT, 64}, {MPCF_BATCH_MODE_GROUPS, 256}};
    float best = 1e30f;
    int bm = b->mode, bt = b->p_threads;
    for (const auto& m : modes) {
        if (rc != MPCF_OK) break;
        b->mode = m[0];
        b->p_threads = m[1];
        float ms = 0;
        rc = tune_trial(T, &ms);
        if (rc == MPCF_OK && ms < best) { best = ms; bm = m[0]; bt = m[1]; }
    }
    b->mode = bm;
    b->p_threads = bt;
    // 2. per extent, largest first: the x+y march length by NY, then the z
    // march length by NZ, the other extents' lengths held at their best
    std::vector<int> ext;
    for (mpcf_ctx* c : b->cases) { ext.push_back(c->g.ny); ext.push_back(c->g.nz); }
    std::sort(ext.begin(), ext.end(), std::greater<int>());
    ext.erase(std::unique(ext.begin(), ext.end()), ext.end());
    for (int which = 0; which < 2 && rc == MPCF_OK; ++which) {
        std::map<int, int>& tuned = which == 0 ? b->tuned_xy : b->tuned_z;
        for (int E : ext) {
            if (rc != MPCF_OK) break;
            bool used = false;
            for (mpcf_ctx* c : b->cases) used |= (which == 0 ? c->g.ny : c->g.nz) == E;
            if (!used) continue;
            const int keep = tuned.count(E) ? tuned[E] : 0;
            int bl = keep;
            std::vector<int> cand = {8, 16, 32, 64, E};
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            for (int l : cand) {
                if (l > E || rc != MPCF_OK) continue;
                tuned[E] = l;
                float ms = 0;
                rc = tune_trial(T, &ms);
                if (rc == MPCF_OK && ms < best) { best = ms; bl = l; }
            }
            if (bl > 0) tuned[E] = bl; else tuned.erase(E);
        }
    }
    tune_free(T);
    b->dirty = true;
    b->use_graph = graph;
    b->steps = 0;
    if (best_len && !ext.empty()) {
        int E = 0;
        for (mpcf_ctx* c : b->cases) E = std::max(E, c->g.ny);
        auto it = b->tuned_xy.find(E);
        *best_len = it != b->tuned_xy.end() ? it->second : 0;
    }
    return rc;
}
