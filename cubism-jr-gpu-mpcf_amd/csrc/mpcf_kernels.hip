// mpcf_kernels.hip -- CDNA4 kernels of one LSRK3 stage and the dt reduction.
//
// Built four times (see build.py):
//   MPCF_FAST=0, -ffp-contract=off  -> namespace mpcf::parity  (bit-exact scheme)
//   MPCF_FAST=1, -ffp-contract=on   -> namespace mpcf::fast
//   MPCF_FAST=1, MPCF_P64=1         -> namespace mpcf::fastp   (fast + binary64
//                                      pressure recovery and speed-of-sound reduction)
//   MPCF_F64=1,  -ffp-contract=off  -> namespace mpcf::f64     (the scheme in binary64:
//                                      storage, arithmetic and reductions)
//
// One RK stage (GPUlab::process_all, source/GPUlab.cpp:703-793) is
//   ghost fill -> x sweep -> y sweep -> z sweep (+ register update)
// Each sweep marches its direction with a rolling five-cell primitive window
// in registers: every cell is converted to primitives once, reconstructed
// once (both face states), and every face is solved once.  The x sweep
// stages x-runs of 64 pencils through LDS so its loads stay contiguous; the
// y and z sweeps put the 64 lanes of a wave along x and march a strided axis.
// Accumulators carry the partial flux sums between sweeps in the reference's
// summation order (((((fxp-fxm)+fyp)-fym)+fzp)-fzm) (GPUkernels.cu:1111).
#include <hip/hip_runtime.h>
#include <type_traits>
#include "mpcf_scheme.h"
#include "mpcf_launch.h"

#ifndef MPCF_FAST
#define MPCF_FAST 0
#endif
#ifndef MPCF_P64
#define MPCF_P64 0
#endif
#ifndef MPCF_F64
#define MPCF_F64 0
#endif
#ifndef MPCF_EPAIR
#define MPCF_EPAIR 0
#endif

#if MPCF_F64
#define MPCF_NS f64
#elif MPCF_EPAIR
#define MPCF_NS fastpe
#elif MPCF_P64
#define MPCF_NS fastp
#elif MPCF_FAST
#define MPCF_NS fast
#else
#define MPCF_NS parity
#endif

namespace mpcf {
namespace MPCF_NS {

#if MPCF_F64
typedef double Real;
#else
typedef float Real;
#endif
typedef PrimT<Real> Prim;
constexpr bool kP64 = MPCF_P64 != 0;
// E stored as a binary32 pair: the solution buffers hold an eighth field,
// the low part of E (E = E[hi] + E[lo], exact in binary64), read wherever E
// is read and written wherever E is written
constexpr bool kEPair = MPCF_EPAIR != 0;
constexpr int NU = 7 + (kEPair ? 1 : 0);   // fields of a solution buffer
constexpr int FELO = 7;                   // the low part of E
// the speed-of-sound reduction and dt/h run in binary64
constexpr bool kSosD = MPCF_P64 != 0 || MPCF_F64 != 0 || MPCF_EPAIR != 0;
// scalar type of the speed-of-sound reduction
typedef std::conditional<kSosD, double, float>::type SosT;

// c + max(|rho u|, |rho v|, |rho w|) / rho of one cell (GPUkernels.cu:1206-1212,
// or MaxSpeedOfSound.cpp:43-47 with 0.5/rho in binary64 for CPU_FORM).  bad:
// the cell fails one of the reference's assertions (rho > 0, E > 0, G > 0,
// c > 0, no NaN: GPUkernels.cu:1171-1173, MaxSpeedOfSound.cpp:31-48) or its
// speed is infinite.  The update's P >= 0 (GPUkernels.cu:1174) is left out:
// rounding leaves P a few ulp of the neighbouring water's P below zero in air
// cells next to a sharp interface in one arithmetic and at +0 in another, and
// a P negative enough to matter drives c^2 <= 0, which is flagged.  Every comparison is false for a NaN, and a NaN
// in any field reaches p and c, so a NaN cell is always flagged -- the maximum
// itself drops NaN operands (fmax) and would not show it.
__device__ __forceinline__ SosT sos_cell(const bool cpu_form, const SosT r, const SosT mu, const SosT mv,
                                         const SosT mw, const SosT E, const SosT G, const SosT P, bool& bad)
{
    SosT p;
    if (cpu_form) p = (SosT)((E - (mu*mu + mv*mv + mw*mw)*(0.5/r) - P)/G);
    else          p = (E - (mu*mu + mv*mv + mw*mw)*(SosT(0.5f)/r) - P) / G;
    const SosT c = msqrt(((p + P) / G + p) / r);
    const SosT s = c + mmax(mmax(fabs(mu), fabs(mv)), fabs(mw)) / r;
    const SosT huge = kSosD ? SosT(__DBL_MAX__) : SosT(__FLT_MAX__);
    bad = !(r > SosT(0) && E > SosT(0) && G > SosT(0) && c > SosT(0) && s <= huge);
    return s;
}

// moves the maximum (and flag) folded into the previous step's last stage to
// the reduction targets, where sos_finalize reads them
__device__ __forceinline__ void take_folded(StepState* st)
{
    if (kSosD) atomicExch(&st->sos_bits_d, atomicOr(&st->next_sos_bits_d, 0ull));
    else       atomicExch(&st->sos_bits, atomicOr(&st->next_sos_bits, 0u));
    atomicExch(&st->bad_bits, atomicOr(&st->next_bad, 0u));
}

typedef FaceFluxT<Real> FaceFlux;
typedef CellRecT<Real> CellRec;
typedef RecT<Real> Rec;
constexpr bool kFast = MPCF_FAST != 0;
// FAST builds fold the x and y parts of a stage into R as soon as they are
// known, R <- a R - dtinvh (x+y flux differences), and keep only the three
// G/P accumulators (sum G, sum P, div u) for the z sweep, which then finishes
// R <- R - dtinvh (z flux differences - div u sum G / 6 ...).  The same
// algebra as the reference's single update, rounded in two parts; it moves
// the R read from the bandwidth-bound z sweep to the x+y sweep, which has
// bandwidth to spare, and drops 7 accumulator fields from the z sweep's
// reads.  The bit-exact builds keep the reference's single update.
constexpr bool kFold = kFast;
// In a stage with a = 0 (the first stage of a step) the folded x+y part is
// stored unscaled, R' = a R - (x+y flux differences), and the z sweep scales
// it: R = dtinvh * R' - dtinvh (z part).  dtinvh * (-t) rounds to exactly
// -(dtinvh * t), so the results are bit for bit those of scaling in the x+y
// sweep -- but the x+y sweep of a step's first stage no longer needs dt, and
// can run while the previous step's last z sweeps (which fold the maximum
// speed of sound dt is taken from) are still in flight.
__device__ __forceinline__ bool unscaled_stage(const Real a) { return kFold && a == Real(0.0f); }
// accelerator complex dies of an MI355X; workgroups are dispatched to them in
// turn by linear workgroup id
constexpr int kXCDs = 8;
// the fused x+y march unrolled by five in the binary32 FAST builds (in the
// bit-exact and binary64 builds the unrolled march spills)
constexpr bool kUnrollXY = kFast && !MPCF_F64;
// likewise the y / z marches (a separate switch: their register budget differs)
constexpr bool kUnrollMarch = kFast && !MPCF_F64;

// MODE_STAGE_NOR: a stage whose R is not stored -- the last stage of a step
// that another step of the same mpcf_steps call follows (that step's first
// stage multiplies R by a = 0), 28 B/cell less for the bandwidth-bound z sweep
enum { MODE_FIRST = 0, MODE_ACC = 1, MODE_CONV = 2, MODE_STAGE = 3, MODE_STAGE_NOR = 4 };
constexpr bool is_stage(const int mode) { return mode == MODE_STAGE || mode == MODE_STAGE_NOR; }

struct StageArgs
{
    const Real* __restrict__ u;    // input solution (7 fields, ghosts filled)
    Real* __restrict__ R;          // RK register (7 fields), updated in place
    Real* __restrict__ acc;        // sweep accumulators (10 fields)
    Real* __restrict__ uo;         // output solution (MODE_STAGE)
    int64_t fs;                    // field stride (elements)
    Real a, b, dtinvh;
    const StepState* st;           // if non-null, dtinvh is read from st->dtinvh
    StepState* sos_next;           // non-null: fold the max speed of sound of the written u into
                                   // sos_next->next_sos_bits[_d] (last stage of a step)
    int sos_cpu;                   // its form: MaxSpeedOfSound_CPP (1) or the GPU kernel's (0)
    int z0, z1;                    // the z range of interior cells this launch covers: the planes
                                   // of the x, y and x+y sweeps, the marched rows of the z sweep
                                   // (a slab of a pipelined step; 0, nz otherwise)
    int bcmask;                    // bit f: face f reflecting; bit 8+f: face f a halo
                                   // face (ghost band imported); else absorbing
};

// Boundary conditions at load time (BoundaryConditions.h:60-109).  A stencil
// read at index i outside [0, n) along an axis is served by the cell its
// ghost would copy: absorbing -- the boundary cell; reflecting -- the mirror
// cell (-1-i on the low face, 2n-1-i on the high face) with the axis' normal
// momentum negated (sg = -1).  The values are the ghost values bit for bit,
// so no kernel fills or reads a ghost band -- except on a halo face, whose
// ghost band holds a neighbouring block's cells (imported by k_halo) and is
// read as it is.
__device__ __forceinline__ int bc_index(const int i, const int n, const int bcmask, const int face_lo, Real& sg)
{
    sg = Real(1.0f);
    if (i < 0) {
        if ((bcmask >> (8 + face_lo)) & 1) return i;
        if ((bcmask >> face_lo) & 1) { sg = Real(-1.0f); return -1 - i; }
        return 0;
    }
    if (i >= n) {
        if ((bcmask >> (9 + face_lo)) & 1) return i;
        if ((bcmask >> (face_lo + 1)) & 1) { sg = Real(-1.0f); return 2 * n - 1 - i; }
        return n - 1;
    }
    return i;
}

// A lane's element index split into a part that is the same for every lane of
// the wave (base, a scalar) and the lane's byte offset from it (lo, small and
// non-negative).  The fused x+y sweep accesses its fields through buffer
// resources built on wave-uniform element pointers (the field, row and wave
// parts of the index, added by the scalar unit), with the lane's 32-bit
// offset: as flat global accesses the same loads and stores took one 64-bit
// vector add each, 3 % of its vector instructions.  The y and z sweeps use the
// same split with flat global accesses (see ld_lane).
struct LaneIdx { int64_t base; uint32_t lo; };

// e: the lane's element; floor_e: an element no greater than the element of
// any active lane of the wave (the first active lane's value is taken)
__device__ __forceinline__ LaneIdx lane_idx(const int e, const int floor_e)
{
    const int b = __builtin_amdgcn_readfirstlane(floor_e);
    return LaneIdx{b, (uint32_t)(e - b) * (uint32_t)sizeof(Real)};
}

// raw buffer resource at a wave-uniform address: no stride, no record limit,
// 32-bit data format (a zero format would read as zero)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(const void* p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, -1, 0x00020000);
}

// element p[u] + lane offset lo (u wave-uniform).  BUF: through a buffer
// resource; else a flat global access (the y/z sweeps: their marches keep
// more scalar state live, and a resource per access runs them out of scalar
// registers)
template <bool BUF, typename V>
__device__ __forceinline__ V ld_lane(const V* p, const int64_t u, const uint32_t lo)
{
    if constexpr (!BUF) return p[u + lo / sizeof(V)];
    else if constexpr (sizeof(V) == 4) return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(rsrc_at(p + u), lo, 0, 0));
    else return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(rsrc_at(p + u), lo, 0, 0));
}

// store; STREAM: non-temporal, for data that no kernel reads again soon --
// the solution and register written by the z sweep (read by the next stage, a
// whole block later) and the G/P accumulators of the x+y sweep (read once by
// the z sweep of the slab).  The hint keeps these streams from displacing
// lines in the caches; at 512^3 it takes the pipelined step from 20.66 to
// 20.0 ms.
template <bool BUF, bool STREAM = false, typename V>
__device__ __forceinline__ void st_lane(V* p, const int64_t u, const uint32_t lo, const V v)
{
    constexpr int aux = STREAM ? 2 : 0;   // bit 1: nt
    if constexpr (!BUF) {
        if (STREAM) __builtin_nontemporal_store(v, &p[u + lo / sizeof(V)]);
        else        p[u + lo / sizeof(V)] = v;
    } else if constexpr (sizeof(V) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rsrc_at(p + u), lo, 0, aux);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v),
                                              rsrc_at(p + u), lo, 0, aux);
}

// raw conserved values of the cell at index `row` of a marched column (L =
// element of row 0, s = element stride of the axis), boundary-mapped.  Returns
// the sign the axis' normal momentum takes; the caller applies it where the
// values are consumed, so that a prefetch is not waited on when it is issued.
template <bool BUF>
__device__ __forceinline__ Real load_row(const StageArgs& A, const LaneIdx& L, const int row, const int n, const int s,
                                         const int face_lo, Real* out)
{
    Real sg;
    const int r = bc_index(row, n, A.bcmask, face_lo, sg);
#pragma unroll
    for (int q = 0; q < NU; ++q) out[q] = ld_lane<BUF>(A.u, L.base + q*A.fs + (int64_t)r * s, L.lo);
    return sg;
}

// primitives of a loaded row of NU conserved values
__device__ __forceinline__ Prim prim_of(const Real* w)
{
    return to_prim<Real, kFast, kP64, kEPair>(w[0], w[1], w[2], w[3], w[4], w[5], w[6], kEPair ? w[FELO] : Real(0));
}

// operands of one cell's accumulator update, loaded one step ahead
template <int MODE>
struct EmitIn
{
    Real acc[NACC];
    Real R[NVAR];
    Real u[NU];
};

template <int MODE>
__device__ __forceinline__ void emit_load(const StageArgs& A, const LaneIdx& L, const int64_t row, EmitIn<MODE>& in)
{
    const int64_t fs = A.fs;
    const int64_t e = L.base + row;   // the wave-uniform part of the cell's element
    if (MODE == MODE_FIRST) return;
    if (!kFold || MODE == MODE_ACC) {
#pragma unroll
        for (int q = 0; q < NACC; ++q) in.acc[q] = ld_lane<false>(A.acc, q*fs + e, L.lo);
    } else {
        // folded: only the G/P accumulators remain
        in.acc[A_SUMG] = ld_lane<false>(A.acc, A_SUMG*fs + e, L.lo);
        in.acc[A_SUMP] = ld_lane<false>(A.acc, A_SUMP*fs + e, L.lo);
        in.acc[A_DIVU] = ld_lane<false>(A.acc, A_DIVU*fs + e, L.lo);
    }
    if (MODE == MODE_CONV || is_stage(MODE) || (kFold && MODE == MODE_ACC)) {
#pragma unroll
        for (int q = 0; q < NVAR; ++q) in.R[q] = ld_lane<false>(A.R, q*fs + e, L.lo);
    }
}

// write the contribution of one cell.  F = flux at the cell's right face,
// Fl = flux at its left face; cG/cP = plus-state + minus-state of G/P.
template <int MODE>
__device__ __forceinline__ void emit(const StageArgs& A, const Real dtinvh, const LaneIdx& L, const int64_t row, const EmitIn<MODE>& in,
                                     const FaceFlux& F, const FaceFlux& Fl, const Real cG, const Real cP,
                                     SosT& smax, bool& bad)
{
    Real* __restrict__ acc = A.acc;
    const int64_t fs = A.fs;
    const int64_t e = L.base + row;   // the wave-uniform part of the cell's element
    const uint32_t lo = L.lo;
    if (MODE == MODE_FIRST) {
        // _xflux + _xextraterm_hllc ("=" assignment, GPUkernels.cu:57-59)
#pragma unroll
        for (int q = 0; q < NVAR; ++q) st_lane<false>(acc, q*fs + e, lo, F.f[q] - Fl.f[q]);
        st_lane<false>(acc, A_SUMG*fs + e, lo, cG);
        st_lane<false>(acc, A_SUMP*fs + e, lo, cP);
        st_lane<false>(acc, A_DIVU*fs + e, lo, F.vel - Fl.vel);
    } else if (MODE == MODE_ACC) {
        // _yextraterm_hllc ("+=", GPUkernels.cu:439-441)
#pragma unroll
        for (int q = 0; q < NVAR; ++q) {
            const Real t = (in.acc[q] + F.f[q]) - Fl.f[q];
            if (kFold) st_lane<false>(A.R, q*fs + e, lo, unscaled_stage(A.a) ? A.a*in.R[q] - t : A.a*in.R[q] - dtinvh*t);
            else       st_lane<false>(acc, q*fs + e, lo, t);
        }
        st_lane<false>(acc, A_SUMG*fs + e, lo, in.acc[A_SUMG] + cG);
        st_lane<false>(acc, A_SUMP*fs + e, lo, in.acc[A_SUMP] + cP);
        st_lane<false>(acc, A_DIVU*fs + e, lo, in.acc[A_DIVU] + (F.vel - Fl.vel));
    } else {
        // _zextraterm_hllc + _divergence (+ _update), GPUkernels.cu:812-814,1104-1136,1164-1170
        const Real sG = in.acc[A_SUMG] + cG;
        const Real sP = in.acc[A_SUMP] + cP;
        const Real dU = in.acc[A_DIVU] + (F.vel - Fl.vel);
        const Real factor6 = Real(1.0f) / Real(6.0f);
        Real un[NVAR];
        double Eun = 0.0;   // E + b R in binary64 (E pair builds)
#pragma unroll
        for (int q = 0; q < NVAR; ++q) {
            // folded: R already holds a R - dtinvh (x+y part)
            const Real t = kFold ? F.f[q] - Fl.f[q] : (in.acc[q] + F.f[q]) - Fl.f[q];
            Real rhs;
            if (q == FG)      rhs = dtinvh*(t - dU * sG * factor6);
            else if (q == FP) rhs = dtinvh*(t - dU * sP * factor6);
            else              rhs = dtinvh*t;
            // folded: R holds a R - dtinvh (x+y part), or a R - (x+y part)
            // when unscaled (an explicit product: no contraction into the
            // subtraction, so it rounds as the x+y sweep's product would)
            const Real Rx = !kFold ? A.a*in.R[q] : unscaled_stage(A.a) ? __fmul_rn(dtinvh, in.R[q]) : in.R[q];
            const Real Rn = Rx - rhs;
            if (MODE != MODE_STAGE_NOR) st_lane<false, true>(A.R, q*fs + e, lo, Rn);
            if (is_stage(MODE)) {
                if (kEPair && q == FE) {
                    // E + b R in binary64, stored as the rounded sum and its
                    // remainder (exact: |remainder| <= half an ulp of the sum)
                    const double Ed = (double(in.u[FE]) + double(in.u[FELO])) + double(A.b) * double(Rn);
                    const Real hi = Real(Ed);
                    un[q] = hi;
                    st_lane<false, true>(A.uo, q*fs + e, lo, hi);
                    st_lane<false, true>(A.uo, FELO*fs + e, lo, Real(Ed - double(hi)));
                    Eun = Ed;
                } else {
                    un[q] = A.b*Rn + in.u[q];
                    st_lane<false, true>(A.uo, q*fs + e, lo, un[q]);
                }
            }
        }
        // the next step's speed of sound, from the values just stored
        if (is_stage(MODE) && A.sos_next) {
            bool b;
            smax = mmax(smax, sos_cell(A.sos_cpu != 0, un[0], un[1], un[2], un[3], kEPair ? SosT(Eun) : SosT(un[4]),
                                       un[5], un[6], b));
            bad = bad || b;
        }
    }
}

// marching state along direction D
template <int D, int MODE>
struct March
{
    Prim w0, w1, w2, w3, w4;   // primitives of cells k-2 .. k+2
    Prim pm;                   // minus-state (face k-1/2) of cell k-1
    Real pGp, pPp;             // plus-state G,P (face k-3/2) of cell k-1
    FaceFlux Fl;               // flux at face k-3/2
    SosT smax = 0;             // running max speed of sound of the emitted cells (folded step)
    bool bad = false;          // an emitted cell is non-finite or non-physical (folded step)

    __device__ __forceinline__ void push(const Prim& q) { w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = q; }

    // reconstruct cell k, solve face k-1/2 (FLUX: k >= c0), emit cell k-1
    // (EMIT: k >= c0 + 1; operands in `in`).  FLUX and EMIT are bools or
    // std::bool_constant (the peeled steps of an unrolled march)
    template <class FLUX, class EMIT>
    __device__ __forceinline__ void step(const StageArgs& A, const Real dtinvh, const FLUX flux, const EMIT em,
                                         const LaneIdx& L, const int64_t row_km1, const EmitIn<MODE>& in)
    {
        const CellRec r = reconstruct<Real, kFast>(w0, w1, w2, w3, w4);
        if (flux) {
            const FaceFlux F = riemann<D, Real, kFast>(pm, r.p);
            if (em) emit<MODE>(A, dtinvh, L, row_km1, in, F, Fl, pGp + pm.G, pPp + pm.P, smax, bad);
            Fl = F;
        }
        pm = r.m; pGp = r.p.G; pPp = r.p.P;
    }
};

__device__ __forceinline__ Prim load_prim(const Real* __restrict__ u, const int64_t fs, const int e)
{
    return to_prim<Real, kFast, kP64, kEPair>(u[e], u[fs + e], u[2*fs + e], u[3*fs + e], u[4*fs + e], u[5*fs + e],
                                              u[6*fs + e], kEPair ? u[FELO*fs + e] : Real(0));
}

__device__ __forceinline__ Real stage_dtinvh(const StageArgs& A)
{
    if (!A.st) return A.dtinvh;
    return MPCF_F64 ? Real(A.st->dtinvh_d) : Real(A.st->dtinvh);
}

// ---------------------------------------------------------------------------
// y / z sweeps: block of 256 lanes over the flattened (x, transverse) columns,
// blockIdx.y = chunk of the marched axis.
// ---------------------------------------------------------------------------
// one column of a y / z sweep; returns the running max speed of sound of the
// cells it wrote and sets bad if one of them is non-finite or non-physical
// (zero / false unless A.sos_next)
template <int D, int MODE>
__device__ __forceinline__ SosT sweep_yz_column(const Geom& g, const StageArgs& A, const int chunk, const int col,
                                                const int bz, bool& bad)
{
    const int nd = D == 1 ? g.ny : g.nz;
    const int x = col % g.nx;
    // the y sweep's columns lie in the launch's planes [z0, z1); the z sweep
    // marches its launch's rows [z0, z1) (boundary conditions still use the
    // whole extent nd)
    const int t = (D == 1 ? A.z0 : 0) + col / g.nx;
    const int c0 = (D == 2 ? A.z0 : 0) + bz * chunk;
    const int c1 = min(c0 + chunk, D == 2 ? A.z1 : nd);
    const int s = g.stride(D);
    const int e0 = D == 1 ? g.idx(x, 0, t) : g.idx(x, t, 0);
    // the columns of a wave are consecutive: the first active lane's element is the least
    const LaneIdx L = lane_idx(e0, e0);
    // the y sweep of a first stage stores its part unscaled and does not read dt
    const Real dtinvh = (MODE == MODE_ACC && unscaled_stage(A.a)) ? Real(0.0f) : stage_dtinvh(A);
    const int64_t fs = A.fs;
    constexpr bool kRaw = is_stage(MODE);   // the update needs the raw u of the emitted cell

    constexpr int NEGQ = D == 1 ? FV : FW;   // normal momentum of the marched axis
    constexpr int FLO = 2 * D;               // its low face
    March<D, MODE> m;
    // raw conserved values of cells k-1 .. k+2 (kept for the register update;
    // the emitted cell k-1 is always interior, so its raw values are exact)
    Real r0[NU], r1[NU], r2[NU], r3[NU], w[NU];
    r1[NEGQ] *= load_row<false>(A, L, c0 - 2, nd, s, FLO, r1);
    r2[NEGQ] *= load_row<false>(A, L, c0 - 1, nd, s, FLO, r2);
    r3[NEGQ] *= load_row<false>(A, L, c0, nd, s, FLO, r3);
    w[NEGQ] *= load_row<false>(A, L, c0 - 3, nd, s, FLO, w);
    m.w1 = prim_of(w);
    m.w2 = prim_of(r1);
    m.w3 = prim_of(r2);
    m.w4 = prim_of(r3);
    // software pipeline: raw values of the next incoming cell and the emit
    // operands of the next step are loaded one step ahead
    Real nxt[NU];
    Real nsg = load_row<false>(A, L, c0 + 1, nd, s, FLO, nxt);
    EmitIn<MODE> ein, ein_next;
    // one march step at row k.  FLUX: face k-1/2 is solved (k >= c0); EMIT:
    // cell k-1 is written (k >= c0 + 1); PF: row k+3 is prefetched and the
    // emit operands of cell k are loaded (k < c1; the latter from k >= c0)
    auto step = [&](auto FLUX, auto EMIT, auto PF, const int k) {
        nxt[NEGQ] *= nsg;   // boundary sign of the prefetched row
        // rotate the raw ring: r0 = cell k-1, r3 = cell k+2
#pragma unroll
        for (int q = 0; q < NU; ++q) { r0[q] = r1[q]; r1[q] = r2[q]; r2[q] = r3[q]; r3[q] = nxt[q]; }
        const Prim q = prim_of(nxt);  // cell k+2
        if (PF) nsg = load_row<false>(A, L, k + 3, nd, s, FLO, nxt);
        if (PF && FLUX) emit_load<MODE>(A, L, (int64_t)k * s, ein_next);   // emitted at step k+1
        if (kRaw) {
#pragma unroll
            for (int f = 0; f < NU; ++f) ein.u[f] = r0[f];                   // cell k-1
        }
        m.push(q);
        m.step(A, dtinvh, FLUX, EMIT, L, (int64_t)(k - 1) * s, ein);
        ein = ein_next;
    };
    if constexpr (kUnrollMarch) {
        using T_ = std::true_type;
        using F_ = std::false_type;
        step(F_{}, F_{}, T_{}, c0 - 1);
        step(T_{}, F_{}, T_{}, c0);
        int k = c0 + 1;
        // five steps per trip, the period of the window of primitives
        for (; k + 5 <= c1; k += 5) {
            step(T_{}, T_{}, T_{}, k);
            step(T_{}, T_{}, T_{}, k + 1);
            step(T_{}, T_{}, T_{}, k + 2);
            step(T_{}, T_{}, T_{}, k + 3);
            step(T_{}, T_{}, T_{}, k + 4);
        }
        for (; k < c1; ++k) step(T_{}, T_{}, T_{}, k);
        step(T_{}, T_{}, F_{}, c1);      // c1 >= c0 + 1: every chunk has a row
    } else {
        for (int k = c0 - 1; k <= c1; ++k) step(k >= c0, k >= c0 + 1, k < c1, k);
    }
    bad = m.bad;
    return m.smax;
}

// column `col` of the columns (x, t) flattened x-fastest: consecutive lanes
// are consecutive x (coalesced) and no lane idles when nx is not a multiple
// of 64.  Every lane of the wave takes part (the fold reduction is per wave)
template <int D, int MODE>
__device__ __forceinline__ void sweep_yz_body(const Geom& g, const StageArgs& A, const int chunk,
                                              const int col, const int bz)
{
    const int nt = D == 1 ? A.z1 - A.z0 : g.ny;
    SosT smax = 0;
    bool bad = false;
    if (col < g.nx * nt) smax = sweep_yz_column<D, MODE>(g, A, chunk, col, bz, bad);   // no barriers in the body
    if (is_stage(MODE) && A.sos_next) {
        // one atomic max (and, for a bad cell, one atomic or) per wave; every
        // lane of the wave takes part
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) smax = mmax(smax, __shfl_xor(smax, off, 64));
        const bool wbad = __ballot(bad) != 0;
        if ((threadIdx.x & 63) == 0) {
            if (smax > SosT(0)) {
                if (kSosD) atomicMax(&A.sos_next->next_sos_bits_d, (unsigned long long)__double_as_longlong((double)smax));
                else       atomicMax(&A.sos_next->next_sos_bits, __float_as_uint((float)smax));
            }
            if (wbad) atomicOr(&A.sos_next->next_bad, 1u);
        }
    }
}

template <int D, int MODE>
__global__ __launch_bounds__(256)
void k_sweep_yz(const Geom g, const StageArgs A, const int chunk)
{
    sweep_yz_body<D, MODE>(g, A, chunk, blockIdx.x * 256 + threadIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// x sweep: one wave per block, lane = pencil (y,z) flattened.  x-runs of SX
// cells of the block's 64 pencils are loaded with contiguous segments,
// converted to primitives and staged in LDS; each lane then marches its own
// pencil through the run.  The run's outputs are staged in LDS as well and
// stored back as contiguous segments (lane-per-pencil stores would touch 64
// cache lines per instruction).  The raw values of the next run are loaded
// into registers while the current run is computed.
// ---------------------------------------------------------------------------
enum { SX = 8 };

struct XStage
{
    Real in[NVAR][64][SX + 1];     // primitives of the current run
    Real out[NACC][64][SX + 1];    // outputs of cells xs-3 .. xs+SX-4
};

// x-sweep output for one cell, written to LDS instead of global memory
__device__ __forceinline__ void emit_first_lds(XStage& S, const int lane, const int slot, const FaceFlux& F,
                                               const FaceFlux& Fl, const Real cG, const Real cP)
{
#pragma unroll
    for (int q = 0; q < NVAR; ++q) S.out[q][lane][slot] = F.f[q] - Fl.f[q];
    S.out[A_SUMG][lane][slot] = cG;
    S.out[A_SUMP][lane][slot] = cP;
    S.out[A_DIVU][lane][slot] = F.vel - Fl.vel;
}

__global__ __launch_bounds__(64)
void k_sweep_x(const Geom g, const StageArgs A, const int chunk)
{
    __shared__ XStage S;
    const int lane = threadIdx.x;
    const int npen = g.ny * (A.z1 - A.z0);     // pencils (y, z) of the launch's planes [z0, z1)
    const int pb = blockIdx.x * 64;
    const int c0 = blockIdx.y * chunk;
    const int c1 = min(c0 + chunk, g.nx);
    const int64_t fs = A.fs;
    const int xl = lane % SX;

    // element index of x = 0 for the 8 pencils this lane stages/stores
    int sbase[8];
    bool sok[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int p = pb + i * 8 + lane / SX;
        sok[i] = p < npen;
        const int pc = sok[i] ? p : npen - 1;
        sbase[i] = g.idx(0, pc % g.ny, A.z0 + pc / g.ny);
    }
    const bool active = pb + lane < npen;

    // raw values of the next run: cell xs + xl of the lane's 8 staging pencils
    Real pre[8][NVAR];
    Real pre_sg;   // boundary sign of the prefetched run's x-momentum, applied when it is staged
    auto load_run = [&](const int xs) {
        const int xc = min(xs + xl, c1 + 2);   // reads past the run end are clamped to the stencil reach
        const int xm = bc_index(xc, g.nx, A.bcmask, 0, pre_sg);
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int q = 0; q < NVAR; ++q) pre[i][q] = A.u[q * fs + sbase[i] + xm];
    };

    March<0, MODE_FIRST> m;
    int xs = c0 - 3;
    load_run(xs);
    for (; xs <= c1 + 2; xs += SX) {
        // stage the run (previous run's consumers and storers are past the barrier below)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const Prim q = to_prim<Real, kFast, kP64>(pre[i][0], pre[i][1] * pre_sg, pre[i][2], pre[i][3], pre[i][4],
                                                      pre[i][5], pre[i][6]);
            const int pl = i * 8 + lane / SX;
            S.in[FR][pl][xl] = q.r; S.in[FU][pl][xl] = q.u; S.in[FV][pl][xl] = q.v; S.in[FW][pl][xl] = q.w;
            S.in[FE][pl][xl] = q.p; S.in[FG][pl][xl] = q.G; S.in[FP][pl][xl] = q.P;
        }
        if (xs + SX <= c1 + 2) load_run(xs + SX);
        __syncthreads();
        const int nj = min(SX, c1 + 3 - xs);
        for (int j = 0; j < nj; ++j) {
            Prim q;
            q.r = S.in[FR][lane][j]; q.u = S.in[FU][lane][j]; q.v = S.in[FV][lane][j]; q.w = S.in[FW][lane][j];
            q.p = S.in[FE][lane][j]; q.G = S.in[FG][lane][j]; q.P = S.in[FP][lane][j];
            m.push(q);
            const int k = xs + j - 2;              // cell reconstructed at this step
            if (k >= c0 - 1) {
                const CellRec r = reconstruct<Real, kFast>(m.w0, m.w1, m.w2, m.w3, m.w4);
                if (k >= c0) {
                    const FaceFlux F = riemann<0, Real, kFast>(m.pm, r.p);
                    if (k >= c0 + 1) emit_first_lds(S, lane, j, F, m.Fl, m.pGp + m.pm.G, m.pPp + m.pm.P);
                    m.Fl = F;
                }
                m.pm = r.m; m.pGp = r.p.G; m.pPp = r.p.P;
            }
        }
        __syncthreads();
        // outputs of this run are cells xs-3+j for the slots j written above
        const int xo = xs - 3 + xl;
        if (xo >= c0 && xo < c1 && xl < nj) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (!sok[i]) continue;
                const int pl = i * 8 + lane / SX;
#pragma unroll
                for (int q = 0; q < NACC; ++q) A.acc[q * fs + sbase[i] + xo] = S.out[q][pl][xl];
            }
        }
    }
    (void)active;
}

// ---------------------------------------------------------------------------
// fused x + y sweep: lanes along x, marching y.  A wave covers 64 consecutive
// x-cells starting 3 cells left of its tile and writes the 58 cells whose
// x-stencils lie inside the wave (lanes 3..60); neighbouring tiles overlap by
// six cells.  The x-direction of row k-1 is computed from the rolling window's
// row k-1 with wave-shifted copies (DPP wave_shr/wave_shl) for the five-point
// windows, the left face state and the right face flux; the y-direction is
// the same register march as the y sweep.  The accumulators leave this kernel
// already summed in the x-then-y order of the reference (GPUkernels.cu:1111),
// so the x sweep's accumulator round trip disappears.
// ---------------------------------------------------------------------------
enum { XY_TILE = kXYTile };

// value of lane l-1 / lane l+1.  The wave's edge lanes (no source lane) read
// zero: they are halo positions whose results are never stored, and with no
// old value to keep the shift is one v_mov_b32_dpp (no copy of the operand),
// which the compiler can also fold into the consuming VALU instruction
template <int CTRL>
__device__ __forceinline__ float dpp_move(const float v)
{
    const int b = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(b, CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(const double v)
{
    // the two 32-bit halves move with the same lane shift
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    const int lo2 = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
    const int hi2 = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((long long)hi2 << 32) | (long long)(unsigned)lo2);
}
__device__ __forceinline__ Real from_left(const Real v) { return dpp_move<0x138>(v); }    // wave_shr:1
__device__ __forceinline__ Real from_right(const Real v) { return dpp_move<0x130>(v); }   // wave_shl:1
__device__ __forceinline__ Prim prim_from_left(const Prim& q)
{
    Prim o;
    o.r = from_left(q.r); o.u = from_left(q.u); o.v = from_left(q.v); o.w = from_left(q.w);
    o.p = from_left(q.p); o.G = from_left(q.G); o.P = from_left(q.P);
    return o;
}
__device__ __forceinline__ Prim prim_from_right(const Prim& q)
{
    Prim o;
    o.r = from_right(q.r); o.u = from_right(q.u); o.v = from_right(q.v); o.w = from_right(q.w);
    o.p = from_right(q.p); o.G = from_right(q.G); o.P = from_right(q.P);
    return o;
}

// workgroups go to the XCDs round-robin by linear id; this maps linear id
// `lin` of a grid of n to a tile index such that each XCD takes a contiguous
// run of tiles, so that neighbouring tiles of the fused x+y sweep (which
// overlap by six positions and share partial cache lines) meet in one L2.
// For grids of equal-cost tiles (one block)
__device__ __forceinline__ int xcd_order(const int lin, const int n)
{
    const int xcd = lin % kXCDs, k = lin / kXCDs;
    const int per = n / kXCDs, extra = n % kXCDs;
    return xcd * per + min(xcd, extra) + k;
}

// x-direction reconstruction of each lane's cell from its lane neighbours.
// FAST: the first differences are formed once per lane and shifted (a lane's
// b - a is its left neighbour's c - b, the same subtraction of the same
// values), six instructions per field instead of eight
__device__ __forceinline__ RecT<Real> weno_lanes(const Real c)
{
    const Real b = from_left(c), d = from_right(c);
    if constexpr (kFast) {
        const Real dbc = c - b, dcd = d - c;
        return weno_fast_diff<Real>(b, c, d, from_left(dbc), dbc, dcd, from_right(dcd));
    } else {
        return weno_cell<Real, kFast>(from_left(b), b, c, d, from_right(d));
    }
}
__device__ __forceinline__ CellRec reconstruct_lanes(const Prim& c)
{
    CellRec o;
    RecT<Real> t;
    t = weno_lanes(c.r); o.m.r = t.m; o.p.r = t.p;
    t = weno_lanes(c.u); o.m.u = t.m; o.p.u = t.p;
    t = weno_lanes(c.v); o.m.v = t.m; o.p.v = t.p;
    t = weno_lanes(c.w); o.m.w = t.m; o.p.w = t.p;
    t = weno_lanes(c.p); o.m.p = t.m; o.p.p = t.p;
    t = weno_lanes(c.G); o.m.G = t.m; o.p.G = t.p;
    t = weno_lanes(c.P); o.m.P = t.m; o.p.P = t.p;
    return o;
}

// wave `wv` of the x+y sweep over rows [bz * chunk, bz * chunk + chunk).
// FIRST (folded builds, a = 0 known at launch): R is not read -- the stage
// multiplies it by zero -- which takes 28 B/cell from that stage's traffic
template <bool FIRST = false>
__device__ __forceinline__ void sweep_xy_body(const Geom& g, const StageArgs& A, const int chunk,
                                              const int wv, const int bz)
{
    // positions q along the padded rows (x = -3 .. nx+2) of all z-planes,
    // flattened; wave w covers q = 58 w .. 58 w + 63 and stores lanes 3..60,
    // so consecutive waves overlap by six positions and ghost positions are
    // computed but never stored
    const int lane = threadIdx.x & 63;
    const int rowlen = g.nx + 6;
    const int nzr = A.z1 - A.z0;                  // planes of this launch
    const int q = wv * XY_TILE + lane;
    int z = q / rowlen;
    const int x = q - z * rowlen - 3;
    const bool store = lane >= 3 && lane <= 60 && x >= 0 && x < g.nx && z < nzr;
    if (wv * XY_TILE >= rowlen * nzr) return;     // whole wave past the end
    z = A.z0 + min(z, nzr - 1);                   // lanes past the last plane read its last row
    // lanes outside [0, nx) read the cell their x-ghost copies
    Real sgx;
    const int xm = bc_index(x, g.nx, A.bcmask, 0, sgx);
    const bool xmirror = (A.bcmask & 3) != 0;
    const int c0 = bz * chunk;
    const int c1 = min(c0 + chunk, g.ny);
    const int s = g.px;
    const int e0 = g.idx(xm, 0, z);
    // every lane's element is at least that of the first ghost column of lane
    // 0's plane (lanes left of x = 0 read it on a halo face)
    const LaneIdx L = lane_idx(e0, g.idx(-GHOST, 0, z));
    const int64_t fs = A.fs;
    Real* __restrict__ acc = A.acc;
    // boundary signs of a loaded row: y-momentum by the row's mapping,
    // x-momentum by the lane's (applied where the values are consumed)
    auto fix = [&](Real* o, const Real sgy) {
        o[FV] *= sgy;
        if (xmirror) o[FU] *= sgx;
    };

    March<1, MODE_ACC> m;
    Real w[NU];
    fix(w, load_row<true>(A, L, c0 - 3, g.ny, s, 2, w)); m.w1 = prim_of(w);
    fix(w, load_row<true>(A, L, c0 - 2, g.ny, s, 2, w)); m.w2 = prim_of(w);
    fix(w, load_row<true>(A, L, c0 - 1, g.ny, s, 2, w)); m.w3 = prim_of(w);
    fix(w, load_row<true>(A, L, c0, g.ny, s, 2, w));     m.w4 = prim_of(w);
    Real nxt[NU];
    Real nsg = load_row<true>(A, L, c0 + 1, g.ny, s, 2, nxt);
    // a first stage stores its x+y part unscaled and does not read dt
    const bool unscaled = FIRST || unscaled_stage(A.a);
    const Real dtinvh = unscaled ? Real(0.0f) : stage_dtinvh(A);
    // one march step at row k.  FLUX: the face below row k is solved
    // (k >= c0); EMIT: row k-1 is finished and written (k >= c0 + 1); PF: row
    // k+3 is prefetched (k < c1)
    auto step = [&](auto FLUX, auto EMIT, auto PF, const int k) {
        fix(nxt, nsg);
        const Prim q = prim_of(nxt);  // row k+2
        if (PF) nsg = load_row<true>(A, L, k + 3, g.ny, s, 2, nxt);
        // folded builds: R of row k-1, emitted below (every lane's address is
        // an interior cell, stored or not)
        Real Ro[NVAR];
        if (kFold && EMIT && !FIRST) {
#pragma unroll
            for (int f = 0; f < NVAR; ++f) Ro[f] = ld_lane<true>(A.R, L.base + f*fs + (int64_t)(k - 1) * s, L.lo);
        }
        m.push(q);
        const CellRec ry = reconstruct<Real, kFast>(m.w0, m.w1, m.w2, m.w3, m.w4);   // row k, y-direction
        if (FLUX) {
            const FaceFlux Fy = riemann<1, Real, kFast>(m.pm, ry.p);                // face k-1/2
            if (EMIT) {
                // x-direction of row k-1 (window slot w1)
                const CellRec rx = reconstruct_lanes(m.w1);
                const Prim Lm = prim_from_left(rx.m);
                const FaceFlux Fxl = riemann<0, Real, kFast>(Lm, rx.p);            // face x-1/2
                FaceFlux Fxr;                                                        // face x+1/2
#pragma unroll
                for (int f = 0; f < NVAR; ++f) Fxr.f[f] = from_right(Fxl.f[f]);
                Fxr.vel = from_right(Fxl.vel);
                if (store) {
                    const int64_t e = L.base + (int64_t)(k - 1) * s;   // wave-uniform part
                    const Real cGy = m.pGp + m.pm.G, cPy = m.pPp + m.pm.P;
#pragma unroll
                    for (int f = 0; f < NVAR; ++f) {
                        const Real t = ((Fxr.f[f] - Fxl.f[f]) + Fy.f[f]) - m.Fl.f[f];
                        // first stage: -t, which is 0 R - t up to the sign of a zero
                        if (kFold) st_lane<true>(A.R, f*fs + e, L.lo, unscaled ? -t : A.a*Ro[f] - dtinvh*t);
                        else       st_lane<true>(acc, f*fs + e, L.lo, t);
                    }
                    st_lane<true, true>(acc, A_SUMG*fs + e, L.lo, (rx.p.G + rx.m.G) + cGy);
                    st_lane<true, true>(acc, A_SUMP*fs + e, L.lo, (rx.p.P + rx.m.P) + cPy);
                    st_lane<true, true>(acc, A_DIVU*fs + e, L.lo, (Fxr.vel - Fxl.vel) + (Fy.vel - m.Fl.vel));
                }
            }
            m.Fl = Fy;
        }
        m.pm = ry.m; m.pGp = ry.p.G; m.pPp = ry.p.P;
    };
    if constexpr (kUnrollXY) {
        using T_ = std::true_type;
        using F_ = std::false_type;
        step(F_{}, F_{}, T_{}, c0 - 1);
        step(T_{}, F_{}, T_{}, c0);
        int k = c0 + 1;
        // five steps per trip, the period of the window of primitives: the
        // window then rotates by register renaming instead of ~50 moves a step
        for (; k + 5 <= c1; k += 5) {
            step(T_{}, T_{}, T_{}, k);
            step(T_{}, T_{}, T_{}, k + 1);
            step(T_{}, T_{}, T_{}, k + 2);
            step(T_{}, T_{}, T_{}, k + 3);
            step(T_{}, T_{}, T_{}, k + 4);
        }
        for (; k < c1; ++k) step(T_{}, T_{}, T_{}, k);
        step(T_{}, T_{}, F_{}, c1);      // c1 >= c0 + 1: every chunk has a row
    } else {
        for (int k = c0 - 1; k <= c1; ++k) step(k >= c0, k >= c0 + 1, k < c1, k);
    }
}

// the unrolled FAST march could take more than 256 VGPRs (one wave per SIMD);
// two waves per SIMD with a few spills outside the march run faster
#if MPCF_FAST && !MPCF_F64
#define MPCF_XY_WAVES __attribute__((amdgpu_waves_per_eu(2, 2)))
#else
#define MPCF_XY_WAVES
#endif
template <bool FIRST>
__global__ __launch_bounds__(256) MPCF_XY_WAVES
void k_sweep_xy(const Geom g, const StageArgs A, const int chunk)
{
    // where a pipelined step's z sweep shares a SIMD with these waves, they
    // issue first (wave priority, not a queue priority): 512^3 19.92 ->
    // 19.85 ms a step; the z sweep's waves first instead was 7 % slower
    __builtin_amdgcn_s_setprio(2);
    const int gx = gridDim.x;
    const int t = xcd_order(blockIdx.x + blockIdx.y * gx, gx * gridDim.y);
    sweep_xy_body<FIRST>(g, A, chunk, (t % gx) * 4 + threadIdx.y, t / gx);
}

// ---------------------------------------------------------------------------
// ghost fill (BoundaryConditions.h:60-109 applied to the padded band).
// blockIdx.y = face = 2*dir + side.  kind: 0 absorbing, 1 reflecting, 2 halo
// (left as imported).
// ---------------------------------------------------------------------------
struct BCSet { int kind[6]; };

__device__ __forceinline__ void ghosts_body(const Geom& g, Real* __restrict__ u, const int64_t fs, const int* bc,
                                            const int bx, const int face)
{
    const int dir = face >> 1, side = face & 1;
    const int na = dir == 0 ? g.ny : g.nx;
    const int nb = dir == 2 ? g.ny : g.nz;
    const int t = bx * 256 + threadIdx.x;
    if (t >= na * nb) return;
    const int ta = t % na, tb = t / na;
    const int n = g.n(dir);
    const int kind = bc[face];
    if (kind == 2) return;
#pragma unroll
    for (int gl = 0; gl < 3; ++gl) {
        const int dst = side == 0 ? gl - 3 : n + gl;
        const int src = kind == 0 ? (side == 0 ? 0 : n - 1) : (side == 0 ? 2 - gl : n - 1 - gl);
        int ed, es;
        if (dir == 0)      { ed = g.idx(dst, ta, tb); es = g.idx(src, ta, tb); }
        else if (dir == 1) { ed = g.idx(ta, dst, tb); es = g.idx(ta, src, tb); }
        else               { ed = g.idx(ta, tb, dst); es = g.idx(ta, tb, src); }
#pragma unroll
        for (int q = 0; q < NVAR; ++q) {
            const Real v = u[q*fs + es];
            u[q*fs + ed] = (kind == 1 && q == 1 + dir) ? -v : v;
        }
    }
}

__global__ __launch_bounds__(256)
void k_ghosts(const Geom g, Real* __restrict__ u, const int64_t fs, const BCSet bc)
{
    ghosts_body(g, u, fs, bc.kind, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// face halos of a decomposed domain (GPUlab::_copysend_halos,
// GPUlab.cpp:76-112, and the receive side that becomes the ghost band,
// GPUlab.cpp:636-641).  The buffer holds the 7 fields one after the other,
// each in the reference's ghost map order (Types.h:31-69): x faces
// [NZ][NY][3], y faces [NZ][3][NX], z faces [3][NY][NX], layer l = the l-th
// of the three cells in increasing index.  MODE 0 (export) reads the three
// boundary cells (0..2 on the low face, n-3..n-1 on the high face), MODE 1
// (import) writes the ghost band (-3..-1 low, n..n+2 high): a high face's
// export is the high neighbour's low-face import and vice versa.  MODE 2
// reads the ghost band back in the same order (the halo buffers GPUlab's
// load_ghosts leaves, GPUlab.cpp:622-644).
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256)
void k_halo(const Geom g, Real* __restrict__ u, const int64_t fs, const int face, Real* __restrict__ buf)
{
    constexpr bool IMPORT = MODE == 1;
    constexpr bool BAND = MODE != 0;   // the ghost band, not the boundary cells
    const int dir = face >> 1, side = face & 1;
    const int area = dir == 0 ? g.ny * g.nz : dir == 1 ? g.nx * g.nz : g.nx * g.ny;
    const int count = 3 * area;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    int x, y, z, layer;
    if (dir == 0)      { layer = t % 3;                 y = (t / 3) % g.ny;    z = t / (3 * g.ny); x = 0; }
    else if (dir == 1) { x = t % g.nx;                  layer = (t / g.nx) % 3; z = t / (3 * g.nx); y = 0; }
    else               { x = t % g.nx;                  y = (t / g.nx) % g.ny;  layer = t / area;   z = 0; }
    const int n = g.n(dir);
    const int cell = BAND ? (side == 0 ? layer - 3 : n + layer) : (side == 0 ? layer : n - 3 + layer);
    if (dir == 0) x = cell; else if (dir == 1) y = cell; else z = cell;
    const int e = g.idx(x, y, z);
#pragma unroll
    for (int q = 0; q < NVAR; ++q) {
        if (IMPORT) u[q * fs + e] = buf[(int64_t)q * count + t];
        else        buf[(int64_t)q * count + t] = u[q * fs + e];
    }
}

// ---------------------------------------------------------------------------
// max speed of sound and dt in one launch (GPUkernels.cu:1182-1222,
// MaxSpeedOfSound.cpp:18-55, LSRK3_IntegratorMPI.cpp:37-38): every block
// folds its maximum into the step state with an atomic max, then counts
// itself done; the last block reads the maximum, derives dt and dt/h, and puts
// the reduction target and the counter back to zero for the next step.
// ---------------------------------------------------------------------------

struct DtArgs
{
    int with_dt;                  // 0: only the maximum (mpcf_max_sos)
    double cfl, h, dt_max;
};

__device__ __forceinline__ void sos_finalize(StepState* st, const DtArgs& d)
{
    double sos;
    if (kSosD) {
        const unsigned long long bits = atomicOr(&st->sos_bits_d, 0ull);   // coherent read
        sos = __longlong_as_double((long long)bits);
        st->sos_d = sos;
        st->sos = (float)sos;
        atomicExch(&st->sos_bits_d, 0ull);
    } else {
        const float f = __uint_as_float(atomicOr(&st->sos_bits, 0u));
        sos = f;
        st->sos = f;
        st->sos_d = f;
        atomicExch(&st->sos_bits, 0u);
    }
    st->sos_is_d = kSosD ? 1 : 0;
    st->bad = atomicExch(&st->bad_bits, 0u);
    // the folded maximum (and flag) of the next step starts from zero
    atomicExch(&st->next_sos_bits, 0u);
    atomicExch(&st->next_sos_bits_d, 0ull);
    atomicExch(&st->next_bad, 0u);
    if (d.with_dt) {
        // binary64 dt from the binary32 (or binary64) maximum
        double dt = d.cfl*d.h/sos;
        dt = d.dt_max < dt ? d.dt_max : dt;
        st->dt = dt;
        st->h = d.h; st->cfl = d.cfl; st->dt_max = d.dt_max;
        st->dtinvh = (float)(dt/d.h);
        st->dtinvh_d = dt/d.h;
    }
    atomicExch(&st->sos_count, 0u);
}

// the reduction over the rows of a block, blocks b = 0..nb-1 of it in this
// launch; returns true in the one thread that finalized (thread (0, 0) of the
// last block to finish).  DENSE: the fields are in the host layout, cell
// (x, y, z) at x + nx*(y + ny*z), instead of the padded one
template <int CPU_FORM, bool DENSE = false>
__device__ __forceinline__ bool sos_body(const Geom& g, const Real* __restrict__ u, const int64_t fs, StepState* st,
                                         const int b, const int nb, SosT* wmax, int* wbad, const DtArgs& d)
{
    const int rows = g.ny * g.nz;
    SosT sos = 0.0f;
    bool bad = false;
    for (int row = b * 4 + threadIdx.y; row < rows; row += nb * 4) {
        const int e0 = DENSE ? row * g.nx : g.idx(0, row % g.ny, row / g.ny);
        for (int x = threadIdx.x; x < g.nx; x += 64) {
            const int e = e0 + x;
            bool cb;
            const SosT Ec = (kEPair && !DENSE) ? SosT(double(u[4*fs + e]) + double(u[FELO*fs + e]))
                                               : SosT(u[4*fs + e]);
            sos = mmax(sos, sos_cell(CPU_FORM != 0, u[e], u[fs + e], u[2*fs + e], u[3*fs + e], Ec,
                                     u[5*fs + e], u[6*fs + e], cb));
            bad = bad || cb;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sos = mmax(sos, __shfl_xor(sos, off, 64));
    const int anybad = __ballot(bad) != 0;
    if (threadIdx.x == 0) { wmax[threadIdx.y] = sos; wbad[threadIdx.y] = anybad; }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const SosT m = mmax(mmax(wmax[0], wmax[1]), mmax(wmax[2], wmax[3]));
        if (kSosD) atomicMax(&st->sos_bits_d, (unsigned long long)__double_as_longlong((double)m));
        else       atomicMax(&st->sos_bits, __float_as_uint((float)m));
        if (wbad[0] | wbad[1] | wbad[2] | wbad[3]) atomicOr(&st->bad_bits, 1u);
        __threadfence();
        if (atomicAdd(&st->sos_count, 1u) == (unsigned)nb - 1u) {
            __threadfence();
            sos_finalize(st, d);
            return true;
        }
    }
    return false;
}

// dt from the maximum folded into the previous step's last z sweep
__global__ __launch_bounds__(64)
void k_sos_folded(StepState* st, const DtArgs d)
{
    if (threadIdx.x != 0) return;
    take_folded(st);
    __threadfence();
    sos_finalize(st, d);
}

template <int CPU_FORM>
__global__ __launch_bounds__(256)
void k_sos(const Geom g, const Real* __restrict__ u, const int64_t fs, StepState* st, const DtArgs d)
{
    __shared__ SosT wmax[4];
    __shared__ int wbad[4];
    sos_body<CPU_FORM>(g, u, fs, st, blockIdx.x, gridDim.x, wmax, wbad, d);
}

// the same reduction over fields in the host layout (cell x + nx*(y + ny*z),
// field q at u + q*fs), with st a caller-provided reduction state that must
// hold zeros in sos_bits[_d], bad_bits and sos_count; the finalizing thread
// copies the maximum (binary64), dt and the flag to the outputs (any may be
// null)
template <int CPU_FORM>
__global__ __launch_bounds__(256)
void k_sos_dense(const Geom g, const Real* __restrict__ u, const int64_t fs, StepState* st, const DtArgs d,
                 double* sos_out, double* dt_out, int* bad_out)
{
    __shared__ SosT wmax[4];
    __shared__ int wbad[4];
    if (sos_body<CPU_FORM, true>(g, u, fs, st, blockIdx.x, gridDim.x, wmax, wbad, d)) {
        if (sos_out) *sos_out = st->sos_d;
        if (dt_out) *dt_out = st->dt;
        if (bad_out) *bad_out = (int)st->bad;
    }
}

// ---------------------------------------------------------------------------
// Update_CUDA (GPUkernels.cu:1164-1170): u <- b*R + u, in place
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256)
void k_update(const Geom g, Real* __restrict__ u, const Real* __restrict__ R, const int64_t fs, const Real b)
{
    const int rows = g.ny * g.nz;
    for (int row = blockIdx.x * 4 + threadIdx.y; row < rows; row += gridDim.x * 4) {
        const int e0 = g.idx(0, row % g.ny, row / g.ny);
        for (int x = threadIdx.x; x < g.nx; x += 64) {
            const int e = e0 + x;
#pragma unroll
            for (int q = 0; q < NVAR; ++q) {
                if (kEPair && q == FE) {
                    const double Ed = (double(u[FE*fs + e]) + double(u[FELO*fs + e])) + double(b) * double(R[FE*fs + e]);
                    const Real hi = Real(Ed);
                    u[FE*fs + e] = hi;
                    u[FELO*fs + e] = Real(Ed - double(hi));
                } else {
                    u[q*fs + e] = b*R[q*fs + e] + u[q*fs + e];
                }
            }
        }
    }
}

// E of an E-pair solution <-> a dense host field: to the device, hi = the
// value rounded to binary32 and lo = the remainder (zero for binary32 host
// values); from the device, hi + lo (binary64), rounded once for binary32
template <int TO_PADDED, typename Dense>
__global__ __launch_bounds__(256)
void k_pack_epair(const Geom g, Real* __restrict__ hi, Real* __restrict__ lo, Dense* __restrict__ dense)
{
    const int rows = g.ny * g.nz;
    for (int row = blockIdx.x * 4 + threadIdx.y; row < rows; row += gridDim.x * 4) {
        const int e0 = g.idx(0, row % g.ny, row / g.ny);
        const int64_t d0 = (int64_t)g.nx * row;
        for (int x = threadIdx.x; x < g.nx; x += 64) {
            if (TO_PADDED) {
                const double v = (double)dense[d0 + x];
                const Real h = Real(v);
                hi[e0 + x] = h;
                lo[e0 + x] = Real(v - double(h));
            } else {
                dense[d0 + x] = (Dense)(double(hi[e0 + x]) + double(lo[e0 + x]));
            }
        }
    }
}

// dense reference SoA (NX*NY*NZ per field, Types.h:31) <-> padded layout;
// the dense side may be binary32 or binary64 (converted on the way)
template <int TO_PADDED, typename Dense>
__global__ __launch_bounds__(256)
void k_pack(const Geom g, Real* __restrict__ pad, Dense* __restrict__ dense, const int64_t fs, const int nq)
{
    const int rows = g.ny * g.nz;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    for (int row = blockIdx.x * 4 + threadIdx.y; row < rows; row += gridDim.x * 4) {
        const int y = row % g.ny, z = row / g.ny;
        const int e0 = g.idx(0, y, z);
        const int64_t d0 = (int64_t)g.nx * row;
        for (int x = threadIdx.x; x < g.nx; x += 64)
            for (int q = 0; q < nq; ++q) {
                if (TO_PADDED) pad[q*fs + e0 + x] = (Real)dense[q*n + d0 + x];
                else           dense[q*n + d0 + x] = (Dense)pad[q*fs + e0 + x];
            }
    }
}

// Streamers of the reference's dumps (Types.h:188-280) between the padded
// state and a dense channel-interleaved binary32 array [Z][Y][X][C]
// (HDF5Dumper_MPI.h:118-141):
//   KIND 0 myTensorialStreamer, 9 channels: rho, u, v, w, p, G, P, 0, 0 with
//          u = rho u / rho in the storage type and p = (E - 0.5 |rho u|^2 / rho
//          - P) / G evaluated in binary64 from the stored values (the
//          reference's 0.5 is a double literal)
//   KIND 1 mySaveStreamer, 7 channels: the conserved fields (also read back)
//   KIND 2 myScalarStreamer, 1 channel: rho
template <int KIND, int OUT>
__global__ __launch_bounds__(256)
void k_stream(const Geom g, Real* __restrict__ u, float* __restrict__ dense, const int64_t fs)
{
    constexpr int NC = KIND == 0 ? 9 : KIND == 1 ? 7 : 1;
    const int rows = g.ny * g.nz;
    for (int row = blockIdx.x * 4 + threadIdx.y; row < rows; row += gridDim.x * 4) {
        const int e0 = g.idx(0, row % g.ny, row / g.ny);
        const int64_t d0 = (int64_t)g.nx * row;
        for (int x = threadIdx.x; x < g.nx; x += 64) {
            const int e = e0 + x;
            float* o = dense + (d0 + x) * NC;
            if (!OUT) {
#pragma unroll
                for (int q = 0; q < NVAR; ++q) u[q*fs + e] = (Real)o[q];
                continue;
            }
            const Real r = u[e], ru = u[fs + e], rv = u[2*fs + e], rw = u[3*fs + e];
            const Real E = u[4*fs + e], G = u[5*fs + e], P = u[6*fs + e];
            if (KIND == 2) { o[0] = (float)r; continue; }
            if (KIND == 1) {
                o[0] = (float)r; o[1] = (float)ru; o[2] = (float)rv; o[3] = (float)rw;
                o[4] = (float)E; o[5] = (float)G; o[6] = (float)P;
                continue;
            }
            const Real ss = ru*ru + rv*rv + rw*rw;
            o[0] = (float)r;
            o[1] = (float)(ru / r);
            o[2] = (float)(rv / r);
            o[3] = (float)(rw / r);
            o[4] = (float)(((double)E - 0.5*(double)ss/(double)r - (double)P) / (double)G);
            o[5] = (float)G;
            o[6] = (float)P;
            o[7] = 0.0f;
            o[8] = 0.0f;
        }
    }
}

// device-function test hooks (the counterpart of GPU::TestKernel,
// GPUkernels.cu:1354-1382): run weno_cell / riemann on arrays
__global__ void k_test_weno(const int n, const Real* __restrict__ in5, Real* __restrict__ m, Real* __restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Real* a = in5 + 5*(int64_t)i;
    const Rec r = weno_cell<Real, kFast>(a[0], a[1], a[2], a[3], a[4]);
    m[i] = r.m; p[i] = r.p;
}

// st: n x 7 x 6 conserved values at stencil points i-3..i+2 of face i-1/2;
// out: n x 12 (7 fluxes, vel, Gm, Gp, Pm, Pp)
template <int D>
__global__ void k_test_face(const int n, const Real* __restrict__ st, Real* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Real* s = st + 42*(int64_t)i;
    Prim c[6];
    for (int k = 0; k < 6; ++k)
        c[k] = to_prim<Real, kFast, kP64>(s[0*6+k], s[1*6+k], s[2*6+k], s[3*6+k], s[4*6+k], s[5*6+k], s[6*6+k]);
    const CellRec left = reconstruct<Real, kFast>(c[0], c[1], c[2], c[3], c[4]);   // cell i-1
    const CellRec right = reconstruct<Real, kFast>(c[1], c[2], c[3], c[4], c[5]);  // cell i
    const FaceFlux F = riemann<D, Real, kFast>(left.m, right.p);
    Real* o = out + 12*(int64_t)i;
    for (int q = 0; q < NVAR; ++q) o[q] = F.f[q];
    o[7] = F.vel; o[8] = left.m.G; o[9] = right.p.G; o[10] = left.m.P; o[11] = right.p.P;
}

// ---------------------------------------------------------------------------
// batched ensemble: one launch per kernel kind advances every live case.  A
// flat block index is mapped to (case, local block) by a binary search over
// the per-case prefix offsets; the local block runs the same body as the
// single-case kernel, so every case gets exactly the arithmetic it gets
// alone.  The cases' current-buffer indices live on the device and are
// flipped after each stage, so a whole step can be captured in a graph.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_case(const int* __restrict__ off, const int n, const int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ StageArgs case_args(const CaseDesc& C, const int cur, const Real a, const Real b)
{
    StageArgs A;
    A.sos_next = nullptr; A.sos_cpu = 0;
    A.bcmask = 0;
    for (int f = 0; f < 6; ++f) A.bcmask |= (C.bc[f] == 1) << f;
    A.z0 = 0; A.z1 = C.g.nz;
    A.u = (const Real*)C.u[cur]; A.R = (Real*)C.R; A.acc = (Real*)C.acc; A.uo = (Real*)C.u[1 - cur];
    A.fs = C.g.fstride;
    A.a = a; A.b = b; A.dtinvh = Real(0.0f); A.st = C.st;
    return A;
}

template <int CPU_FORM>
__global__ __launch_bounds__(256)
void kb_sos(const BatchDev B, const double cfl, const double dt_max)
{
    __shared__ SosT wmax[4];
    __shared__ int wbad[4];
    const int c = find_case(B.off_sos, B.n, blockIdx.x);
    const CaseDesc& C = B.cases[c];
    const DtArgs d{1, cfl, C.h, dt_max};
    const int local = blockIdx.x - B.off_sos[c];
    if (C.st->folded) {
        // the previous step's last z sweep folded this state's maximum: only
        // dt remains, done by the case's first block
        if (local == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
            take_folded(C.st);
            __threadfence();
            sos_finalize(C.st, d);
        }
        return;
    }
    sos_body<CPU_FORM>(C.g, (const Real*)C.u[B.cur[c]], C.g.fstride, C.st, local, C.nb_sos, wmax, wbad, d);
}

template <bool FIRST>
__global__ __launch_bounds__(256) MPCF_XY_WAVES
void kb_sweep_xy(const BatchDev B, const Real a, const Real b)
{
    // round-robin over the XCDs as dispatched: the tiles of different cases
    // differ in cost, and an even mix per XCD balances them (contiguous runs
    // per XCD, as in k_sweep_xy, were slower here)
    const int c = find_case(B.off_xy, B.n, blockIdx.x);
    const CaseDesc& C = B.cases[c];
    const int local = blockIdx.x - B.off_xy[c];
    const StageArgs A = case_args(C, B.cur[c], a, b);
    sweep_xy_body<FIRST>(C.g, A, C.chunk_xy, (local % C.nbx_xy) * 4 + threadIdx.y, local / C.nbx_xy);
}

// the z sweep finishes the stage; the last block of each case to finish
// flips the case's current-buffer index (every block has read it before it
// counts itself: no fence is needed, and a device-scope fence here would
// write back and invalidate the XCD's L2 once per block).
// fold != 0 (a step's last stage): the sweep also folds the maximum speed of
// sound of the state it writes, for the next step's kb_sos.  MODE_STAGE_NOR:
// R is not stored (a caller that never reads R after a step)
template <int MODE>
__global__ __launch_bounds__(256)
void kb_sweep_z(const BatchDev B, const Real a, const Real b, const int fold, const int cpu_form)
{
    const int c = find_case(B.off_z, B.n, blockIdx.x);
    const CaseDesc& C = B.cases[c];
    const int local = blockIdx.x - B.off_z[c];
    StageArgs A = case_args(C, B.cur[c], a, b);
    if (fold) { A.sos_next = C.st; A.sos_cpu = cpu_form; }
    sweep_yz_body<2, MODE>(C.g, A, C.chunk_z, (local % C.nbx_z) * 256 + threadIdx.x, local / C.nbx_z);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nblk = (unsigned)(B.off_z[c + 1] - B.off_z[c]);
        if (atomicAdd(&C.st->z_count, 1u) == nblk - 1u) {
            atomicExch(&B.cur[c], 1 - B.cur[c]);
            if (fold) atomicExch(&C.st->folded, 1);
            atomicExch(&C.st->z_count, 0u);
        }
    }
}

// ---------------------------------------------------------------------------
// persistent batched stage: two launches per RK stage for every case, one for
// the x+y items of all cases, then one for their z items (a z item reads the
// R and accumulators the x+y items of its case wrote; the launch boundary
// orders them -- inside one launch that order would take a device-scope
// release and acquire per item, an L2 write-back and invalidation of the XCD,
// which measured 1.7x slower than flat launches).  Each workgroup pulls items
// from a device queue (an atomic counter) until it is empty; an item is one
// workgroup's share of one case's x+y sweep (W = blockDim.x / 64 consecutive
// 58-position tiles over one y-chunk) or z sweep (blockDim.x columns over one
// z-chunk), running the single-case bodies, so every case gets exactly the
// arithmetic it gets alone.  The queue holds the items of the larger cases
// first, so the small ones fill the launch's tail.
// ---------------------------------------------------------------------------
// Thread 0 alone takes an item or counts a finished one; those steps are
// calls of their own, bracketed by barriers.  Written inline, the compiler
// may restructure the item loop around such a branch so that thread 0 leaves
// the loop body on its own while the other threads of its wave go round
// again, which breaks the barriers and lets those threads run an item twice;
// a call is convergent, so the loop body the compiler sees has no branch that
// splits the workgroup.
__device__ __attribute__((noinline)) unsigned take_item(PCtl* ctl)
{
    __shared__ unsigned s_item;
    if (threadIdx.x == 0) s_item = atomicAdd(&ctl->head, 1u);
    __syncthreads();
    const unsigned i = s_item;
    __syncthreads();                              // read everywhere before the next take
    return i;
}

// a z item is finished; the case's last one flips its current-buffer index
// (every z item of the case has read it by then: each reads it before it
// counts itself) and clears the count for the next stage
__device__ __attribute__((noinline)) void count_z(unsigned* done, const unsigned items_z, int* cur_c,
                                                  const int cur, int* folded, const int fold)
{
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(done, 1u) == items_z - 1u) {
        atomicExch(cur_c, 1 - cur);
        if (fold) atomicExch(folded, 1);
        atomicExch(done, 0u);
    }
}

// PHASE 0: the x+y items (queue items [0, seg[nxy].first)); PHASE 1: the z
// items (the rest).  Two kernels, so that each holds only its own sweep's
// registers
template <int PHASE>
__device__ __forceinline__ void kp_stage_body(const PersistDev& P, const Real a, const Real b, const int fold,
                                              const int cpu_form)
{
    constexpr int phase = PHASE;
    const int W = (int)(blockDim.x >> 6);        // waves per item
    const int wib = (int)(threadIdx.x >> 6);     // this wave's place in the item
    const int s0 = phase == 0 ? 0 : P.nxy, s1 = phase == 0 ? P.nxy : P.nseg;
    const unsigned first = P.seg[s0].first, end = P.seg[s1].first;
    for (;;) {
        // the same value in every thread: readfirstlane only makes it scalar
        const unsigned i = first + (unsigned)__builtin_amdgcn_readfirstlane((int)take_item(P.ctl));
        if (i >= end) break;
        int lo = s0, hi = s1 - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (P.seg[mid].first <= i) lo = mid; else hi = mid - 1;
        }
        const PSeg S = P.seg[lo];
        const CaseDesc& C = P.cases[S.c];
        const int j = (int)(i - S.first);
        const int cur = P.cur[S.c];
        StageArgs A = case_args(C, cur, a, b);
        if constexpr (phase == 0) {
            (void)fold; (void)cpu_form;
            sweep_xy_body(C.g, A, C.chunk_xy, (j % C.wxy) * W + wib, j / C.wxy);
        } else {
            if (fold) { A.sos_next = C.st; A.sos_cpu = cpu_form; }
            sweep_yz_body<2, MODE_STAGE>(C.g, A, C.chunk_z, (j % C.wz) * (int)blockDim.x + (int)threadIdx.x, j / C.wz);
            count_z(P.done + S.c, C.items_z, &P.cur[S.c], cur, &C.st->folded, fold);
        }
    }
    // the last workgroup out puts the queue back for the next launch
    if (threadIdx.x == 0 && atomicAdd(&P.ctl->exited, 1u) == P.nblocks - 1u) {
        atomicExch(&P.ctl->head, 0u);
        atomicExch(&P.ctl->exited, 0u);
    }
}

__global__ __launch_bounds__(256) MPCF_XY_WAVES
void kp_stage_xy(const PersistDev P, const Real a, const Real b)
{
    kp_stage_body<0>(P, a, b, 0, 0);
}

__global__ __launch_bounds__(256)
void kp_stage_z(const PersistDev P, const Real a, const Real b, const int fold, const int cpu_form)
{
    kp_stage_body<1>(P, a, b, fold, cpu_form);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static int pick_chunk(const int nd, const int64_t columns, const int mult)
{
    // enough marching threads for ~4 waves per SIMD (256 CUs x 4 SIMDs), times
    // mult: a launch that runs alone wants more (its sweeps hold two waves per
    // SIMD, and the last partial round of waves otherwise idles most of the
    // chip); a slab of a pipelined step shares the chip with the other
    // stream's sweep, which fills that tail
    const int64_t target_threads = 256LL * 4 * 4 * 64 * mult;
    int nchunks = (int)((target_threads + columns - 1) / columns);
    if (nchunks < 1) nchunks = 1;
    int chunk = (nd + nchunks - 1) / nchunks;
    if (chunk < 16) chunk = nd < 16 ? nd : 16;
    return chunk;
}

static int rows_grid(const Geom& g)
{
    const int rows = g.ny * g.nz;
    int blocks = (rows + 3) / 4;
    return blocks < 8192 ? blocks : 8192;
}

void launch_ghosts(const Geom& g, void* uv, const int bc[6], hipStream_t s)
{
    BCSet b;
    for (int i = 0; i < 6; ++i) b.kind[i] = bc[i];
    int tmax = g.ny * g.nz;
    if (g.nx * g.nz > tmax) tmax = g.nx * g.nz;
    if (g.nx * g.ny > tmax) tmax = g.nx * g.ny;
    hipLaunchKernelGGL(k_ghosts, dim3((tmax + 255) / 256, 6), dim3(256), 0, s, g, (Real*)uv, g.fstride, b);
}

// mode 0: export the boundary cells, 1: import into the ghost band, 2: read the ghost band
void launch_halo(const Geom& g, void* uv, void* buf, const int face, const int mode, hipStream_t s)
{
    const int dir = face >> 1;
    const int count = 3 * (dir == 0 ? g.ny * g.nz : dir == 1 ? g.nx * g.nz : g.nx * g.ny);
    const dim3 grid((count + 255) / 256);
    if (mode == 1)
        hipLaunchKernelGGL(k_halo<1>, grid, dim3(256), 0, s, g, (Real*)uv, g.fstride, face, (Real*)buf);
    else if (mode == 2)
        hipLaunchKernelGGL(k_halo<2>, grid, dim3(256), 0, s, g, (Real*)uv, g.fstride, face, (Real*)buf);
    else
        hipLaunchKernelGGL(k_halo<0>, grid, dim3(256), 0, s, g, (Real*)uv, g.fstride, face, (Real*)buf);
}

void launch_sos(const Geom& g, const void* uv, StepState* st, const int cpu_form, const int with_dt, const double cfl,
                const double h, const double dt_max, hipStream_t s)
{
    const Real* u = (const Real*)uv;
    const int nb = rows_grid(g) < 2048 ? rows_grid(g) : 2048;
    const DtArgs d{with_dt, cfl, h, dt_max};
    if (cpu_form) hipLaunchKernelGGL(k_sos<1>, dim3(nb), dim3(64, 4), 0, s, g, u, g.fstride, st, d);
    else          hipLaunchKernelGGL(k_sos<0>, dim3(nb), dim3(64, 4), 0, s, g, u, g.fstride, st, d);
}

void launch_sos_dense(const Geom& g, const void* uv, const int64_t fs, StepState* st, const int cpu_form,
                      const double cfl, const double h, const double dt_max, double* sos_out, double* dt_out,
                      int* bad_out, hipStream_t s)
{
    const Real* u = (const Real*)uv;
    const int rows = (g.ny * g.nz + 3) / 4;
    const int nb = rows < 2048 ? rows : 2048;
    const DtArgs d{1, cfl, h, dt_max};
    if (cpu_form)
        hipLaunchKernelGGL(k_sos_dense<1>, dim3(nb), dim3(64, 4), 0, s, g, u, fs, st, d, sos_out, dt_out, bad_out);
    else
        hipLaunchKernelGGL(k_sos_dense<0>, dim3(nb), dim3(64, 4), 0, s, g, u, fs, st, d, sos_out, dt_out, bad_out);
}

void launch_sos_folded(StepState* st, const double cfl, const double h, const double dt_max, hipStream_t s)
{
    const DtArgs d{1, cfl, h, dt_max};
    hipLaunchKernelGGL(k_sos_folded, dim3(1), dim3(64), 0, s, st, d);
}

// one direction sweep.  dir 0 (x) assigns the accumulators, dir 1 (y)
// accumulates, dir 3 does x and y fused (assigns), dir 2 (z) finishes the
// register update (stage == 0: R only, as Convection; stage == 1: R and the
// new solution, as process_all; stage == 2: the new solution only, R not
// stored).
void launch_sweep(const Geom& g, const int dir, const SweepParams& p, const int stage, hipStream_t s)
{
    StageArgs A;
    A.u = (const Real*)p.u; A.R = (Real*)p.R; A.acc = (Real*)p.acc; A.uo = (Real*)p.uo; A.fs = g.fstride;
    A.a = (Real)p.a; A.b = (Real)p.b; A.dtinvh = (Real)p.dtinvh; A.st = p.st;
    A.bcmask = p.bcmask;
    A.sos_next = p.sos_next; A.sos_cpu = p.sos_cpu;
    A.z0 = p.z0; A.z1 = p.z1 < 0 ? g.nz : p.z1;
    const int nzr = A.z1 - A.z0;
    // a whole-block launch (not a slab) targets four times a slab's waves:
    // at 256^3-512^3 its steps are 0.6-2.7 % faster than with twice
    const int alone = (A.z0 == 0 && A.z1 == g.nz) ? 4 : 1;
    if (dir == 3) {
        const int chunk = pick_chunk(g.ny, (int64_t)g.nx * nzr, alone);
        const int waves = ((g.nx + 6) * nzr + XY_TILE - 1) / XY_TILE;
        const dim3 grid((waves + 3) / 4, (g.ny + chunk - 1) / chunk);
        // a first stage of a folded build does not read R
        if (kFold && A.a == Real(0.0f)) hipLaunchKernelGGL(k_sweep_xy<true>, grid, dim3(64, 4), 0, s, g, A, chunk);
        else                            hipLaunchKernelGGL(k_sweep_xy<false>, grid, dim3(64, 4), 0, s, g, A, chunk);
    } else if (dir == 0) {
        const int npen = g.ny * nzr;
        const int chunk = pick_chunk(g.nx, npen, 1);
        dim3 grid((npen + 63) / 64, (g.nx + chunk - 1) / chunk);
        hipLaunchKernelGGL(k_sweep_x, grid, dim3(64), 0, s, g, A, chunk);
    } else if (dir == 1) {
        const int chunk = pick_chunk(g.ny, (int64_t)g.nx * nzr, 1);
        dim3 grid((g.nx * nzr + 255) / 256, (g.ny + chunk - 1) / chunk);
        hipLaunchKernelGGL((k_sweep_yz<1, MODE_ACC>), grid, dim3(256), 0, s, g, A, chunk);
    } else {
        const int chunk = pick_chunk(nzr, (int64_t)g.nx * g.ny, alone);
        const dim3 grid((g.nx * g.ny + 255) / 256, (nzr + chunk - 1) / chunk);
        if (stage == 2)      hipLaunchKernelGGL((k_sweep_yz<2, MODE_STAGE_NOR>), grid, dim3(256), 0, s, g, A, chunk);
        else if (stage == 1) hipLaunchKernelGGL((k_sweep_yz<2, MODE_STAGE>), grid, dim3(256), 0, s, g, A, chunk);
        else                 hipLaunchKernelGGL((k_sweep_yz<2, MODE_CONV>), grid, dim3(256), 0, s, g, A, chunk);
    }
}

void launch_update(const Geom& g, void* u, const void* R, const double b, hipStream_t s)
{
    hipLaunchKernelGGL(k_update, dim3(rows_grid(g)), dim3(64, 4), 0, s, g, (Real*)u, (const Real*)R, g.fstride, (Real)b);
}

void launch_pack(const Geom& g, void* pad, void* dense, const int to_padded, const int nq, const int dense_f64,
                 hipStream_t s)
{
    Real* P = (Real*)pad;
    const dim3 grid(rows_grid(g)), block(64, 4);
    if (dense_f64) {
        if (to_padded) hipLaunchKernelGGL((k_pack<1, double>), grid, block, 0, s, g, P, (double*)dense, g.fstride, nq);
        else           hipLaunchKernelGGL((k_pack<0, double>), grid, block, 0, s, g, P, (double*)dense, g.fstride, nq);
    } else {
        if (to_padded) hipLaunchKernelGGL((k_pack<1, float>), grid, block, 0, s, g, P, (float*)dense, g.fstride, nq);
        else           hipLaunchKernelGGL((k_pack<0, float>), grid, block, 0, s, g, P, (float*)dense, g.fstride, nq);
    }
}

void launch_pack_epair(const Geom& g, void* hi, void* lo, void* dense, const int to_padded, const int dense_f64,
                       hipStream_t s)
{
    Real* H = (Real*)hi;
    Real* L = (Real*)lo;
    const dim3 grid(rows_grid(g)), block(64, 4);
    if (dense_f64) {
        if (to_padded) hipLaunchKernelGGL((k_pack_epair<1, double>), grid, block, 0, s, g, H, L, (double*)dense);
        else           hipLaunchKernelGGL((k_pack_epair<0, double>), grid, block, 0, s, g, H, L, (double*)dense);
    } else {
        if (to_padded) hipLaunchKernelGGL((k_pack_epair<1, float>), grid, block, 0, s, g, H, L, (float*)dense);
        else           hipLaunchKernelGGL((k_pack_epair<0, float>), grid, block, 0, s, g, H, L, (float*)dense);
    }
}

void launch_stream(const Geom& g, void* pad, float* dense, const int kind, const int out, hipStream_t s)
{
    Real* P = (Real*)pad;
    const dim3 grid(rows_grid(g)), block(64, 4);
    if (!out)           hipLaunchKernelGGL((k_stream<1, 0>), grid, block, 0, s, g, P, dense, g.fstride);
    else if (kind == 0) hipLaunchKernelGGL((k_stream<0, 1>), grid, block, 0, s, g, P, dense, g.fstride);
    else if (kind == 1) hipLaunchKernelGGL((k_stream<1, 1>), grid, block, 0, s, g, P, dense, g.fstride);
    else                hipLaunchKernelGGL((k_stream<2, 1>), grid, block, 0, s, g, P, dense, g.fstride);
}

void launch_test_weno(const int n, const void* in5v, void* mv, void* pv, hipStream_t s)
{
    const Real* in5 = (const Real*)in5v;
    Real* m = (Real*)mv;
    Real* p = (Real*)pv;
    hipLaunchKernelGGL(k_test_weno, dim3((n + 255) / 256), dim3(256), 0, s, n, in5, m, p);
}

void launch_test_face(const int n, const int dir, const void* stv, void* outv, hipStream_t s)
{
    const Real* st = (const Real*)stv;
    Real* out = (Real*)outv;
    dim3 grid((n + 127) / 128), block(128);
    if (dir == 0)      hipLaunchKernelGGL(k_test_face<0>, grid, block, 0, s, n, st, out);
    else if (dir == 1) hipLaunchKernelGGL(k_test_face<1>, grid, block, 0, s, n, st, out);
    else               hipLaunchKernelGGL(k_test_face<2>, grid, block, 0, s, n, st, out);
}

int default_chunk(const int nd, const long long columns) { return pick_chunk(nd, columns, 1); }
// the march length of a whole-block launch that runs alone
int whole_chunk(const int nd, const long long columns) { return pick_chunk(nd, columns, 4); }

// grid of the fused x+y sweep: dims[0] blocks of four waves over the padded
// rows of all planes, dims[2] y-chunks (dims[1] = 1)
void xy_grid(const Geom& g, const int chunk, int dims[3])
{
    const int waves = ((g.nx + 6) * g.nz + XY_TILE - 1) / XY_TILE;
    dims[0] = (waves + 3) / 4; dims[1] = 1; dims[2] = (g.ny + chunk - 1) / chunk;
}

// grid of the z sweep: dims[0] blocks of 256 flattened (x, y) columns, dims[2] z-chunks
void z_grid(const Geom& g, const int chunk, int dims[3])
{
    dims[0] = (g.nx * g.ny + 255) / 256; dims[1] = 1; dims[2] = (g.nz + chunk - 1) / chunk;
}

void batch_sos(const BatchDev& B, const int blocks, const int cpu_form, const double cfl, const double dt_max,
               hipStream_t s)
{
    if (cpu_form) hipLaunchKernelGGL(kb_sos<1>, dim3(blocks), dim3(64, 4), 0, s, B, cfl, dt_max);
    else          hipLaunchKernelGGL(kb_sos<0>, dim3(blocks), dim3(64, 4), 0, s, B, cfl, dt_max);
}

void batch_xy(const BatchDev& B, const int blocks, const double a, const double b, hipStream_t s)
{
    if (kFold && a == 0.0) hipLaunchKernelGGL(kb_sweep_xy<true>, dim3(blocks), dim3(64, 4), 0, s, B, (Real)a, (Real)b);
    else                   hipLaunchKernelGGL(kb_sweep_xy<false>, dim3(blocks), dim3(64, 4), 0, s, B, (Real)a, (Real)b);
}

// storeR == 0: R is not stored (MODE_STAGE_NOR)
void batch_z(const BatchDev& B, const int blocks, const double a, const double b, const int fold, const int cpu_form,
             const int storeR, hipStream_t s)
{
    if (storeR) hipLaunchKernelGGL(kb_sweep_z<MODE_STAGE>, dim3(blocks), dim3(256), 0, s, B, (Real)a, (Real)b, fold, cpu_form);
    else        hipLaunchKernelGGL(kb_sweep_z<MODE_STAGE_NOR>, dim3(blocks), dim3(256), 0, s, B, (Real)a, (Real)b, fold, cpu_form);
}
// the two launches of one stage; blocks[0] workgroups for the x+y items,
// blocks[1] for the z items
void persist_stage(const PersistDev& P, const int blocks[2], const int threads, const double a, const double b,
                   const int fold, const int cpu_form, hipStream_t s)
{
    PersistDev Q = P;
    Q.nblocks = (unsigned)blocks[0];
    hipLaunchKernelGGL(kp_stage_xy, dim3(blocks[0]), dim3(threads), 0, s, Q, (Real)a, (Real)b);
    Q.nblocks = (unsigned)blocks[1];
    hipLaunchKernelGGL(kp_stage_z, dim3(blocks[1]), dim3(threads), 0, s, Q, (Real)a, (Real)b, fold, cpu_form);
}
// resident workgroups of `threads` threads per CU of the x+y (phase 0) or z
// (phase 1) launch
int persist_occupancy(const int threads, const int phase)
{
    int n = 0;
    const hipError_t e = phase == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kp_stage_xy, threads, 0)
                                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kp_stage_z, threads, 0);
    return e == hipSuccess ? n : 0;
}

} // namespace MPCF_NS
} // namespace mpcf
