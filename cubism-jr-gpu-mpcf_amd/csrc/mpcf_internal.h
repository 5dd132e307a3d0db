// mpcf_internal.h -- library-internal definitions shared by the C-ABI
// translation units (mpcf_capi.cpp, mpcf_serve.cpp): error reporting, the
// launcher tables of the arithmetic builds, the context struct.
#pragma once

#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "mpcf_hip.h"
#include "mpcf_launch.h"

using mpcf::Geom;
using mpcf::StepState;

namespace mpcf {
namespace detail {

// records msg as this thread's last error (mpcf_last_error) and returns code
int fail(int code, const std::string& msg);

// host memcpy of n bytes split over a pool of worker threads (created on first
// use, kept until the process ends); concurrent callers take turns
void parallel_copy(void* dst, const void* src, size_t n);
const char* last_error();

}  // namespace detail
}  // namespace mpcf

using mpcf::detail::fail;

namespace mpcf {
namespace detail {

#define HIP_TRY(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(MPCF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK(what)                                                                \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess)                                                             \
            return fail(MPCF_EHIP, std::string(what) + " launch: " + hipGetErrorString(e_)); \
    } while (0)

struct Launchers
{
    decltype(&mpcf::parity::launch_ghosts) ghosts;
    decltype(&mpcf::parity::launch_halo) halo;
    decltype(&mpcf::parity::launch_sos) sos;
    decltype(&mpcf::parity::launch_sos_folded) sos_folded;
    decltype(&mpcf::parity::launch_sos_dense) sos_dense;
    decltype(&mpcf::parity::launch_sweep) sweep;
    decltype(&mpcf::parity::launch_update) update;
    decltype(&mpcf::parity::launch_pack) pack;
    decltype(&mpcf::parity::launch_pack_epair) pack_epair;
    decltype(&mpcf::parity::launch_stream) stream;
    decltype(&mpcf::parity::launch_test_weno) test_weno;
    decltype(&mpcf::parity::launch_test_face) test_face;
    decltype(&mpcf::parity::batch_sos) b_sos;
    decltype(&mpcf::parity::batch_xy) b_xy;
    decltype(&mpcf::parity::batch_z) b_z;
    decltype(&mpcf::parity::persist_stage) persist;
    decltype(&mpcf::parity::persist_occupancy) persist_occ;
};

#define MPCF_LAUNCHERS(NS)                                                                           \
    {mpcf::NS::launch_ghosts, mpcf::NS::launch_halo, mpcf::NS::launch_sos, mpcf::NS::launch_sos_folded,         \
     mpcf::NS::launch_sos_dense,                                                                                \
     mpcf::NS::launch_sweep, mpcf::NS::launch_update,                                                            \
     mpcf::NS::launch_pack, mpcf::NS::launch_pack_epair, mpcf::NS::launch_stream,                      \
     mpcf::NS::launch_test_weno,                                                                      \
     mpcf::NS::launch_test_face,                                                                      \
     mpcf::NS::batch_sos, mpcf::NS::batch_xy, mpcf::NS::batch_z, mpcf::NS::persist_stage,        \
     mpcf::NS::persist_occupancy}

inline const Launchers kParity = MPCF_LAUNCHERS(parity);
inline const Launchers kFast = MPCF_LAUNCHERS(fast);
inline const Launchers kFastP = MPCF_LAUNCHERS(fastp);
inline const Launchers kFastPE = MPCF_LAUNCHERS(fastpe);
inline const Launchers kF64 = MPCF_LAUNCHERS(f64);
#undef MPCF_LAUNCHERS

// batches and the server step their cases as this many groups of about equal
// cells, each on its own stream (measured on a 64-case ensemble: four groups
// serve a static batch best; the server, whose p99 under a request stream
// suffers from more groups, uses two)
#ifndef MPCF_BATCH_GROUPS
#define MPCF_BATCH_GROUPS 4
#endif
#ifndef MPCF_SERVER_GROUPS
#define MPCF_SERVER_GROUPS 2
#endif
constexpr int kBatchGroups = MPCF_BATCH_GROUPS;
constexpr int kServerGroups = MPCF_SERVER_GROUPS;

// group g of the n cases (cells[i] each) is [first[g], first[g + 1]): every
// group gets one case while cases remain, then the cases whose midpoint lies
// below g + 1 parts of G of the total cells, keeping one case for each later
// group
inline void split_groups(const long long* cells, const int n, const int G, int* first)
{
    long long total = 0, acc = 0;
    for (int i = 0; i < n; ++i) total += cells[i];
    int i = 0;
    first[0] = 0;
    for (int g = 1; g < G; ++g) {
        const long long target2 = 2 * (total / G * g + total % G * g / G);
        if (i < n) acc += cells[i++];
        while (i < n - (G - g) && 2 * acc + cells[i] < target2) acc += cells[i++];
        first[g] = i;
    }
    first[G] = n;
}

inline const unsigned kAllFlags = MPCF_FLAG_FAST | MPCF_FLAG_SOS_CPU | MPCF_FLAG_SPLIT_XY | MPCF_FLAG_FUSED_XY |
                           MPCF_FLAG_P64 | MPCF_FLAG_FP64 | MPCF_FLAG_NO_PIPELINE | MPCF_FLAG_E_PAIR;

// arithmetic build selected by the flags
inline const Launchers* pick_launchers(const unsigned flags)
{
    if (flags & MPCF_FLAG_FP64) return &kF64;
    if (flags & MPCF_FLAG_E_PAIR) return &kFastPE;
    if (flags & MPCF_FLAG_P64) return &kFastP;
    return (flags & MPCF_FLAG_FAST) ? &kFast : &kParity;
}

// storage bytes per element of a context with these flags
inline size_t elem_bytes(const unsigned flags) { return (flags & MPCF_FLAG_FP64) ? sizeof(double) : sizeof(float); }

// the entry points that read or write the solution field by field without
// E's low part (ghost band, halos, streamers, dumps, wavelets) refuse an
// E_PAIR context
inline int refuse_epair(const unsigned flags, const char* who)
{
    if (flags & MPCF_FLAG_E_PAIR)
        return fail(MPCF_EINVAL, std::string(who) + ": not available for an MPCF_FLAG_E_PAIR context");
    return MPCF_OK;
}

// query helper used by mpcf_query: which x/y variant runs
// LSRK3 coefficients (LSRK3_IntegratorMPI.cpp:46,51,56), rounded to the
// storage type by the launchers like GPUlab::process_all's Real arguments
// (GPUlab.h:258): binary32, or binary64 in the binary64 build
inline const double kA[3] = {0.0, -17. / 32, -32. / 27};
inline const double kB[3] = {1. / 4, 8. / 9, 3. / 4};

}  // namespace detail
}  // namespace mpcf

using mpcf::detail::Launchers;
using mpcf::detail::pick_launchers;
using mpcf::detail::kBatchGroups;
using mpcf::detail::kServerGroups;
using mpcf::detail::split_groups;
using mpcf::detail::elem_bytes;
using mpcf::detail::kA;
using mpcf::detail::kB;
using mpcf::detail::kAllFlags;
using mpcf::detail::kParity;
using mpcf::detail::kFast;

// one resident node block
struct mpcf_ctx
{
    Geom g;
    double h;
    unsigned flags;
    int bc[6];
    int device;
    size_t esize;       // bytes per stored element (4, or 8 with MPCF_FLAG_FP64)
    char* base;         // single allocation
    void* u[2];
    int cur;
    void* R;
    void* acc;
    StepState* st;      // device
    StepState* st_host; // pinned host mirror
    hipStream_t stream;
    bool own_stream;
    size_t bytes;
    const Launchers* L;
    // the device holds the max speed of sound of the current u, folded into
    // the last stage of the previous mpcf_step (cleared by anything else that
    // writes u or runs the full reduction)
    bool sos_folded;
    // pipelined steps: a second stream for the z sweeps of the slabs and the
    // events that order them against the x+y sweeps (created on first use)
    hipStream_t stream2;
    std::vector<hipEvent_t> pipe_ev;
    int pipe_req;   // -1 automatic, 0 whole-block stages, k >= 2: k slabs (mpcf_set_pipeline)
    // slabs of the last pipelined step whose z sweeps the context's stream has
    // not joined yet (0: none): within one mpcf_steps call, a FAST step leaves
    // them open so that the next step's first x+y sweeps overlap them; the
    // call joins them before it returns
    int pipe_open;
    hipEvent_t pipe_fork;   // orders the second stream after the context's stream
    // optional per-kernel event timing (mpcf_timing_enable)
    bool timing;
    std::vector<hipEvent_t> ev_pool;
    struct Mark { int kind; hipEvent_t b, e; };
    std::vector<Mark> marks;
};

namespace mpcf {
namespace detail {

// march length of the batch kernels for an axis of extent nd: `forced` if
// positive, else the autotuned length for that extent, else min(nd, 64)
// (enough parallelism comes from the other cases)
inline int batch_march(const int nd, const int forced, const int tuned)
{
    if (forced > 0) return forced < nd ? forced : nd;
    if (tuned > 0) return tuned < nd ? tuned : nd;
    return nd < 64 ? nd : 64;
}

// descriptor of one case for the batch kernels, and its block counts in the
// flat grids: blocks[0] x+y sweep, [1] z sweep, [2] speed of sound
inline void case_desc(const mpcf_ctx* c, const int cur, const int len_xy, const int len_z, CaseDesc& d, int blocks[3])
{
    d.g = c->g;
    d.u[0] = c->u[0]; d.u[1] = c->u[1]; d.R = c->R; d.acc = c->acc; d.st = c->st; d.h = c->h;
    for (int f = 0; f < 6; ++f) d.bc[f] = c->bc[f];
    d.chunk_xy = len_xy;
    d.chunk_z = len_z;
    int dims[3];
    parity::xy_grid(c->g, d.chunk_xy, dims);
    d.nbx_xy = dims[0]; d.nby_xy = dims[1];
    blocks[0] = dims[0] * dims[1] * dims[2];
    parity::z_grid(c->g, d.chunk_z, dims);
    d.nbx_z = dims[0]; d.nby_z = dims[1];
    blocks[1] = dims[0] * dims[1] * dims[2];
    const int rows4 = (c->g.ny * c->g.nz + 3) / 4;
    d.nb_sos = rows4 < 2048 ? rows4 : 2048;
    blocks[2] = d.nb_sos;
    (void)cur;
}

// the persistent stage kernel's items of a case with workgroups of W waves
inline void case_items(CaseDesc& d, const int W)
{
    const int tiles = ((d.g.nx + 6) * d.g.nz + kXYTile - 1) / kXYTile;
    d.wxy = (tiles + W - 1) / W;
    d.items_xy = (unsigned)d.wxy * (unsigned)((d.g.ny + d.chunk_xy - 1) / d.chunk_xy);
    d.wz = (d.g.nx * d.g.ny + 64 * W - 1) / (64 * W);
    d.items_z = (unsigned)d.wz * (unsigned)((d.g.nz + d.chunk_z - 1) / d.chunk_z);
}

}  // namespace detail
}  // namespace mpcf

namespace mpcf {
namespace detail {

// the sweeps' boundary mask: bit f set -- face f reflecting; bit 8+f set --
// face f a halo face (its ghost band is read as imported)
inline int bc_mask(const int bc[6])
{
    int m = 0;
    for (int f = 0; f < 6; ++f) m |= (bc[f] == MPCF_BC_REFLECTING) << f | (bc[f] == MPCF_BC_HALO) << (8 + f);
    return m;
}

inline bool has_halo(const int bc[6])
{
    for (int f = 0; f < 6; ++f)
        if (bc[f] == MPCF_BC_HALO) return true;
    return false;
}

// march length of the batch kernels when neither forced nor tuned: split the
// marched axis until the active cases together have about 2^17 marching
// threads (two waves per SIMD), then round the length down to a power of two
// between 8 and 64.  A step's latency is the march length times the per-cell
// latency, so few small cases run short marches; long marches recompute fewer
// warm-up cells.  Lengths that leave the grid just over a whole number of
// resident rounds of blocks are slow, and powers of two avoid them for the
// cubic sizes served here (measured on MI355X, one 192^3 case: 2.06 ms per
// step with 64-cell marches, 2.16 with 32, 2.35 with 48, 2.83 with 96).
// `columns` is the marching columns of all active cases in this sweep.
inline int adaptive_march(const int nd, const long long columns)
{
    const long long target = 1LL << 17;
    long long nchunks = columns > 0 ? (target + columns - 1) / columns : 1;
    if (nchunks < 1) nchunks = 1;
    const int want = (int)((nd + nchunks - 1) / nchunks);
    int len = 8;
    while (len * 2 <= want && len < 64) len *= 2;
    return len < nd ? len : nd;
}

}  // namespace detail
}  // namespace mpcf
