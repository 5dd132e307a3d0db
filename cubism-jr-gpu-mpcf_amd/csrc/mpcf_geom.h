// mpcf_geom.h -- HBM layout of one node block (host + device).
//
// Every per-cell array (the solution u, the LSRK3 register R, the sweep
// accumulators) uses one padded structure-of-arrays layout:
//
//   field q, cell (x,y,z) with x in [-3, nx+3), y in [-3, ny+3), z in [-3, nz+3)
//   element  = (x + XOFF) + px * ((y + 3) + py * (z + 3))
//   address  = base + q * fstride + element
//
// The three-cell ghost band around the interior replaces the reference's
// separate x/y ghost buffers and z-halo slices (source/Types.h:37-43,
// source/GPU/GPUhousehold.cu:105-134): ghosts are filled on the device and
// every sweep reads its six-point stencil with the same index arithmetic.
// XOFF = 32 puts the first interior cell of every row on a 128-byte L2 line
// boundary (px is a multiple of 32 elements), so a wave's 64 lanes along x
// read exactly two whole lines per field and neighbouring blocks, which run
// on different XCDs with separate L2s, never fetch the same line.
// Only the face-adjacent ghost slabs are ever written or read (the stencils
// are one-directional), so edge and corner padding stays unused.
#pragma once

#include <stdint.h>

#ifndef MPCF_HD
#if defined(__HIPCC__)
#define MPCF_HD __host__ __device__ __forceinline__
#else
#define MPCF_HD inline
#endif
#endif

namespace mpcf {

enum { NVAR = 7, NACC = 10, GHOST = 3, XOFF = 32 };
enum Field { FR = 0, FU, FV, FW, FE, FG, FP };
// accumulator slots of the direction sweeps: 7 flux differences, then the
// cell sums of the reconstructed G and P and the face-velocity difference
enum Acc { A_SUMG = 7, A_SUMP = 8, A_DIVU = 9 };

struct Geom
{
    int nx, ny, nz;       // interior extents
    int px, py, pz;       // padded extents (elements, rows, slices)
    int64_t fstride;      // elements per field

    MPCF_HD int idx(int x, int y, int z) const { return (x + XOFF) + px * ((y + GHOST) + py * (z + GHOST)); }
    MPCF_HD int n(int d) const { return d == 0 ? nx : d == 1 ? ny : nz; }
    // element stride of one step along direction d
    MPCF_HD int stride(int d) const { return d == 0 ? 1 : d == 1 ? px : px * py; }
};

inline Geom make_geom(int nx, int ny, int nz)
{
    Geom g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.px = ((nx + XOFF + GHOST + 31) / 32) * 32;
    g.py = ny + 2 * GHOST;
    g.pz = nz + 2 * GHOST;
    const int64_t e = (int64_t)g.px * g.py * g.pz;
    g.fstride = ((e + 63) / 64) * 64;   // 256-byte aligned field starts
    return g;
}

// device-resident step state: the max speed of sound of the last reduction,
// the time step derived from it and the dt/h every stage reads.  The
// reduction targets and counters must be zero when a block is created.  Builds whose
// speed-of-sound reduction runs in binary64 (the binary64 scheme and the
// selective-binary64 build) use the *_d members; the others the binary32 ones.
struct StepState
{
    double dt;            // binary64, LSRK3_IntegratorMPI.cpp:37-38
    double h, cfl, dt_max;
    float  dtinvh;        // (float)(dt/h), GPUlab.h:258
    float  sos;           // max speed of sound as binary32
    unsigned int sos_bits;// atomicMax target (non-negative floats order as uints)
    int    sos_is_d;      // 1 when the last reduction ran in binary64
    double dtinvh_d;      // dt/h in binary64
    double sos_d;         // max speed of sound as binary64
    unsigned long long sos_bits_d;   // atomicMax target (non-negative doubles order as u64)
    // blocks of the running reduction / batched z sweep that have finished;
    // the last one finalizes and puts the counter back to zero
    unsigned int sos_count;
    unsigned int z_count;
    // maximum speed of sound of the state the last stage of mpcf_step wrote,
    // folded into that stage's z sweep (atomicMax targets, as sos_bits[_d]);
    // the next step turns it into dt instead of reading the state again
    unsigned int next_sos_bits;
    unsigned long long next_sos_bits_d;
    // batched cases: set by the last z block of a step's last stage once
    // next_sos_bits holds the maximum of the state it wrote (every batched
    // step folds; a case's state is zeroed when it joins a batch)
    int folded;
    // non-finite or non-physical cells -- rho <= 0, E <= 0, G <= 0,
    // c^2 <= 0, or a NaN / inf in any field (the assertions of
    // GPUkernels.cu:1171-1173,1218 and MaxSpeedOfSound.cpp:31-48): bad_bits
    // collects the running reduction, bad is the finished reduction's result
    // (the state dt was taken from), next_bad that of the state the last stage
    // of a step wrote (folded with next_sos_bits)
    unsigned int bad_bits;
    unsigned int bad;
    unsigned int next_bad;
};

} // namespace mpcf
