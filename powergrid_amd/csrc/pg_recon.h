// pg_recon.h -- device kernels of the direct (non-iterative) reconstruction
// support: the density-compensation iteration, the per-coil weighting of
// gridCoilImages and the coil-combined magnitude image
// (reconSupport/directRecon.cpp).
#pragma once

#include "pg_internal.h"

namespace pgr {

// d[i] = d[i] / real(c[i])   (directRecon.cpp:50, complex / real)
void dcf_update(float2* d, const float2* c, long long M, hipStream_t s);
// out[i] = |d[i] / ||d||_2|   (directRecon.cpp:54-56); part: kMaxParts doubles
void dcf_finish(const float2* d, long long M, double* part, float* out, hipStream_t s);
// out[c*M + i] = dcf[i] * data[c*M + i]   (directRecon.cpp:123, dcf % data)
void weight_columns(float2* out, const float* dcf, const float2* data, long long M, int nc,
                    hipStream_t s);
// out[v] = sqrt(sum_c |x[v + c*nvox]|)   (directRecon.cpp:63-73)
void sum_of_magnitudes(const float2* x, long long nvox, int nc, float* out, hipStream_t s);

// mag[i] = |x[i]|, phs[i] = arg(x[i])   (processNIFTI.hpp:110-111)
void mag_phase(const float2* x, long long n, float* mag, float* phs, hipStream_t s);

constexpr int kMaxParts = 256;

}  // namespace pgr
