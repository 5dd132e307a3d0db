// mpcf_reduce.h -- binary64 reductions over node blocks (mpcf_reduce.hip).
#pragma once

#include <hip/hip_runtime.h>
#include "mpcf_geom.h"

namespace mpcf {
namespace reduce {
// doubles of device scratch the two reductions need for geometry g
int scratch_doubles(const Geom& g);
// out28[4q + k], field q: k=0 sum|a-b|, 1 max|a-b|, 2 sum|b|, 3 max|b| (device pointer);
// *_pair: that block holds E as a binary32 pair (low part in field 7)
void diff(const Geom& g, const void* a, int a_f64, int64_t fsa, int a_pair, const void* b, int b_f64, int64_t fsb,
          int b_pair, double* scratch, double* out28, hipStream_t s);
// out7[q] = sum over interior cells of field q (device pointer)
void sums(const Geom& g, const void* u, int f64, int64_t fs, int pair, double* scratch, double* out7, hipStream_t s);
}
}
