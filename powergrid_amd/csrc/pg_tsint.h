// pg_tsint.h -- device part of the min-max time-segmentation interpolator.
#pragma once

#include "pg_internal.h"

namespace pgt {

// AA[l*M + i + sh*Mo] = conj(sum_j W[l*N + j] exp(i fm_j tv_i)) for i < Mo,
// every shot sh < shots (M = Mo*shots); fm, tv, W in FP64
void minmax_weights(const double* fm, const double* tv, const double2* W, int N, int Mo, int L,
                    int shots, float2* AA, hipStream_t s);

}  // namespace pgt
