// mpcf_launch.h -- host launchers exported by the two builds of
// mpcf_kernels.hip (namespace parity: bit-exact scheme; fast: FMA-contracted).
#pragma once

#include <hip/hip_runtime.h>
#include "mpcf_geom.h"

namespace mpcf {
// arguments of one direction sweep of a stage
struct SweepParams
{
    const float* u;        // stage input solution (ghosts filled)
    float* R;              // LSRK3 register, updated in place by the z sweep
    float* acc;            // accumulators (10 fields)
    float* uo;             // stage output solution (stage == 1)
    float a, b, dtinvh;    // stage coefficients; dtinvh ignored when st != NULL
    const StepState* st;   // device step state carrying dt/h
};
}

#define MPCF_DECLARE_LAUNCHERS(NS)                                                                         \
namespace mpcf { namespace NS {                                                                            \
void launch_ghosts(const Geom& g, float* u, const int bc[6], hipStream_t s);                               \
void launch_sos(const Geom& g, const float* u, StepState* st, int cpu_form, hipStream_t s);                 \
void launch_dt(StepState* st, double cfl, double h, double dt_max, hipStream_t s);                         \
void launch_sweep(const Geom& g, int dir, const SweepParams& p, int stage, hipStream_t s);                \
void launch_update(const Geom& g, float* u, const float* R, float b, hipStream_t s);                       \
void launch_pack(const Geom& g, float* pad, float* dense, int to_padded, int nq, hipStream_t s);           \
void launch_test_weno(int n, const float* in5, float* m, float* p, hipStream_t s);                         \
void launch_test_face(int n, int dir, const float* st, float* out, hipStream_t s);                         \
} }

MPCF_DECLARE_LAUNCHERS(parity)
MPCF_DECLARE_LAUNCHERS(fast)
