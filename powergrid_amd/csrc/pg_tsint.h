// pg_tsint.h -- device part of the min-max time-segmentation interpolator.
#pragma once

#include "pg_internal.h"

namespace pgt {

// AA[l*M + i + sh*Mo] = conj(sum_j W[l*N + j] exp(i fm_j tv_i)) for i < Mo,
// every shot sh < shots (M = Mo*shots); fm, tv, W in FP64
void minmax_weights(const double* fm, const double* tv, const double2* W, int N, int Mo, int L,
                    int shots, float2* AA, hipStream_t s);

// Histogram interpolator (interptype 3): with the field map binned to nb
// frequencies om_n of weight h_n and z_n = exp(-i om_n tau),
//   c_k(t) = sum_n h_n exp(i om_n t) z_n^k,  k = 0..L,
//   AA[l*M + i + sh*Mo] = conj(sum_k iG[l*(L+1) + k] c_k(tv_i)),  l < L;
// nb <= kHistBins, L + 1 <= kHistCoef; all in FP64
constexpr int kHistBins = 256;
constexpr int kHistCoef = 17;
void histogram_weights(const double* h, const double* om, const double2* z, int nb,
                       const double* tv, int Mo, const double2* iG, int L, int shots, float2* AA,
                       hipStream_t s);

}  // namespace pgt
