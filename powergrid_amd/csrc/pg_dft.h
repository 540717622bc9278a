// pg_dft.h -- exact field-corrected DFT launchers (pg_dft.hip).
#pragma once

#include "pg_internal.h"

namespace pgd {

struct DftData {
    int M, N, precision;
    float *kx, *ky, *kz, *t;   // M
    float *ix, *iy, *iz, *fm;  // N (fm in rad/s)
    // GdftR2 (null otherwise): field-map difference maps (N each) and the
    // image dimensions of the sinc factors
    float *gx = nullptr, *gy = nullptr, *gz = nullptr;
    int numx = 1, numy = 1, numz = 1;
};

// partial-sum buffer size (float2 elements) needed by forward/adjoint
size_t part_elems(int M, int N, int nc, bool adjoint, int nshot = 1);
// out[c*ostride + i] for c < nc (ostride -1: M); S may be null (nc must then be 1).
// With nshot > 1 (M a multiple of nshot) the samples are nshot shots of
// M/nshot, shot-major, and shot s's maps are S_c + s*s_shot.
void forward(const DftData& d, const float2* x, const float2* S, int nc, float2* out, float2* part,
             hipStream_t s, long long ostride = -1, int nshot = 1, long long s_shot = 0);
// out[j] (+)= sum_c conj(S_c[j]) (E^H y_c)[j], y_c at y + c*ystride (-1: M);
// accumulate adds to out.  With nshot > 1 as in forward: shot s's samples
// use its own maps, and the shots' sums are added in shot order.
void adjoint(const DftData& d, const float2* y, const float2* S, int nc, float2* out, float2* part,
             hipStream_t s, long long ystride = -1, bool accumulate = false, int nshot = 1,
             long long s_shot = 0);

// The allocation-free entry point's partial-sum buffer: a file-scope device
// array of static_scratch_elems() float2 (64 MiB) in pg_dft.hip, one copy per
// device for the life of the process.  static_scratch() returns its device
// address on the current device, resolved with hipGetSymbolAddress on the
// first call per device and kept.
size_t static_scratch_elems();
float2* static_scratch();

}  // namespace pgd
