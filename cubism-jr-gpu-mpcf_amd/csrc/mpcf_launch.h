// mpcf_launch.h -- host launchers exported by the five builds of
// mpcf_kernels.hip (namespace parity: bit-exact scheme; fast: FMA-contracted;
// fastp: fast with binary64 pressure recovery and speed-of-sound reduction;
// fastpe: fastp with E stored as a binary32 pair; f64: the scheme in binary64).  Field pointers are untyped: they point to
// the build's storage type (binary32, or binary64 for f64).
#pragma once

#include <hip/hip_runtime.h>
#include "mpcf_geom.h"

namespace mpcf {
// positions of a wave of the fused x+y sweep that it stores (64 lanes less
// the three-cell x-halo on each side)
constexpr int kXYTile = 58;

// arguments of one direction sweep of a stage
struct SweepParams
{
    const void* u;         // stage input solution (ghosts filled)
    void* R;               // LSRK3 register, updated in place by the z sweep
    void* acc;             // accumulators (10 fields)
    void* uo;              // stage output solution (stage == 1)
    double a, b, dtinvh;   // stage coefficients, rounded to the storage type by
                           // the launcher; dtinvh ignored when st != NULL
    const StepState* st;   // device step state carrying dt/h
    StepState* sos_next = nullptr;  // non-null: the z sweep of stage == 1 folds the max speed of
                                    // sound of the state it writes into sos_next->next_sos_bits
    int sos_cpu = 0;                // its form (MPCF_FLAG_SOS_CPU)
    int z0 = 0, z1 = -1;            // z range of the fused x+y sweep's planes and the z sweep's
                                    // rows (a slab of a pipelined step); z1 < 0: the whole block
    int bcmask;            // bit f: face f reflecting; bit 8+f: face f a halo face
                           // (ghost band imported); else absorbing.  The sweeps
                           // apply the boundary conditions as they load
};

// one case of a batched ensemble, as the batch kernels see it
struct CaseDesc
{
    Geom g;
    void* u[2];            // stage input / output buffers (which is input: BatchDev::cur)
    void* R;
    void* acc;
    StepState* st;
    double h;
    int bc[6];
    int chunk_xy, chunk_z; // march lengths of the x+y and z sweeps
    int nbx_xy, nby_xy;    // x+y sweep grid: x tiles, z groups (y chunks = third dim)
    int nbx_z, nby_z;      // z sweep grid: x blocks, y groups (z chunks = third dim)
    int nb_sos;            // speed-of-sound blocks
    // items of the persistent stage kernel (W = workgroup size / 64 waves
    // each): x+y items per y-chunk (W tiles each) and in all; z items per
    // z-chunk (64 W columns each) and in all
    int wxy, wz;
    unsigned items_xy, items_z;
};

// device tables of a batch: per-case descriptors, the current-buffer index of
// each case (flipped on the device after every stage) and, per kernel kind,
// the prefix offsets of the cases' blocks in the flat launch grid
struct BatchDev
{
    const CaseDesc* cases;
    int* cur;
    int n;
    const int* off_xy;
    const int* off_z;
    const int* off_sos;
};

// the work queue of a persistent batched stage: a run of the x+y (kind 0) or
// the z (kind 1) items of one case, starting at item `first` of the queue
// (the x+y items of all cases come first)
struct PSeg
{
    int c;
    int kind;
    unsigned first;
};
// queue head and the workgroups that found it empty
struct PCtl
{
    unsigned head, exited;
};
struct PersistDev
{
    const CaseDesc* cases;
    int* cur;
    const PSeg* seg;       // nseg + 1 entries: the x+y segments [0, nxy), then the z
                           // segments; seg[nseg].first = items in all
    int nseg, nxy;
    unsigned* done;        // per case: finished z items of the running stage
    PCtl* ctl;
    unsigned nblocks;      // workgroups of the launch (set per launch)
};
}

#define MPCF_DECLARE_LAUNCHERS(NS)                                                                         \
namespace mpcf { namespace NS {                                                                            \
void launch_ghosts(const Geom& g, void* u, const int bc[6], hipStream_t s);                                \
void launch_halo(const Geom& g, void* u, void* buf, int face, int mode, hipStream_t s);                    \
void launch_sos_folded(StepState* st, double cfl, double h, double dt_max, hipStream_t s);                 \
void launch_sos_dense(const Geom& g, const void* u, int64_t fs, StepState* st, int cpu_form, double cfl,   \
                      double h, double dt_max, double* sos_out, double* dt_out, int* bad_out, hipStream_t s); \
void launch_sos(const Geom& g, const void* u, StepState* st, int cpu_form, int with_dt, double cfl,          \
                double h, double dt_max, hipStream_t s);                                                   \
void launch_sweep(const Geom& g, int dir, const SweepParams& p, int stage, hipStream_t s);                \
void launch_update(const Geom& g, void* u, const void* R, double b, hipStream_t s);                        \
void launch_pack(const Geom& g, void* pad, void* dense, int to_padded, int nq, int dense_f64, hipStream_t s); \
void launch_pack_epair(const Geom& g, void* hi, void* lo, void* dense, int to_padded, int dense_f64,         \
                       hipStream_t s);                                                                     \
void launch_stream(const Geom& g, void* pad, float* dense, int kind, int out, hipStream_t s);             \
void launch_test_weno(int n, const void* in5, void* m, void* p, hipStream_t s);                            \
void launch_test_face(int n, int dir, const void* st, void* out, hipStream_t s);                           \
void batch_sos(const BatchDev& B, int blocks, int cpu_form, double cfl, double dt_max, hipStream_t s);    \
void batch_xy(const BatchDev& B, int blocks, double a, double b, hipStream_t s);                           \
void batch_z(const BatchDev& B, int blocks, double a, double b, int fold, int cpu_form, int storeR,         \
             hipStream_t s);                                                                               \
void persist_stage(const PersistDev& P, const int blocks[2], int threads, double a, double b, int fold,     \
                   int cpu_form, hipStream_t s);                                                           \
int persist_occupancy(int threads, int phase);                                                             \
void xy_grid(const Geom& g, int chunk, int dims[3]);                                                       \
void z_grid(const Geom& g, int chunk, int dims[3]);                                                        \
int default_chunk(int nd, long long columns);                                                              \
int whole_chunk(int nd, long long columns);                                                                \
} }

MPCF_DECLARE_LAUNCHERS(parity)
MPCF_DECLARE_LAUNCHERS(fast)
MPCF_DECLARE_LAUNCHERS(fastp)
MPCF_DECLARE_LAUNCHERS(fastpe)
MPCF_DECLARE_LAUNCHERS(f64)
