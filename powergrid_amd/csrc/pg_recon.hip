// pg_recon.hip -- kernels of the direct reconstruction support
// (reconSupport/directRecon.cpp): density-compensation update and
// normalisation, dcf weighting of coil data, coil-combined magnitude image.
// All are single passes over HBM; the norm is a two-stage FP64 reduction in
// a fixed order (deterministic).
#include "pg_recon.h"

#include <algorithm>

namespace {

constexpr int kT = 256;

__global__ void __launch_bounds__(kT) k_dcf_update(float2* __restrict__ d,
                                                    const float2* __restrict__ c, long long M)
{
    const long long i = (long long)blockIdx.x * kT + threadIdx.x;
    if (i >= M) return;
    const float r = c[i].x;
    const float2 v = d[i];
    d[i] = make_float2(v.x / r, v.y / r);
}

// part[b] = sum over the block's fixed slice of |d|^2 in FP64
__global__ void __launch_bounds__(kT) k_norm_parts(const float2* __restrict__ d, long long M,
                                                    double* __restrict__ part)
{
    __shared__ double sh[kT / 64];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < M;
         i += (long long)gridDim.x * kT) {
        const float2 v = d[i];
        acc += (double)v.x * v.x + (double)v.y * v.y;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kT / 64; ++w) s += sh[w];
        part[blockIdx.x] = s;
    }
}

// out[i] = |d[i] / n|, n = sqrt(sum of the np partials), summed by every
// block in the same order
__global__ void __launch_bounds__(kT) k_dcf_finish(const float2* __restrict__ d, long long M,
                                                    const double* __restrict__ part, int np,
                                                    float* __restrict__ out)
{
    __shared__ float s_n;
    if (threadIdx.x < 64) {
        double s = 0.0;
        for (int k = threadIdx.x; k < np; k += 64) s += part[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (threadIdx.x == 0) s_n = (float)sqrt(s);
    }
    __syncthreads();
    const long long i = (long long)blockIdx.x * kT + threadIdx.x;
    if (i >= M) return;
    const float n = s_n;
    const float2 v = d[i];
    out[i] = hypotf(v.x / n, v.y / n);
}

__global__ void __launch_bounds__(kT) k_weight_columns(float2* __restrict__ out,
                                                        const float* __restrict__ dcf,
                                                        const float2* __restrict__ data,
                                                        long long M, int nc)
{
    const long long i = (long long)blockIdx.x * kT + threadIdx.x;
    if (i >= M) return;
    const float w = dcf[i];
    for (int c = 0; c < nc; ++c) {
        const float2 v = data[(long long)c * M + i];
        out[(long long)c * M + i] = make_float2(w * v.x, w * v.y);
    }
}

// two voxels per thread (16-byte loads) when nvox is even and x 16-byte
// aligned; the coil sum runs in coil order in FP32
template <bool VEC>
__global__ void __launch_bounds__(kT) k_sum_of_magnitudes(const float2* __restrict__ x,
                                                           long long nvox, int nc,
                                                           float* __restrict__ out)
{
    const long long t = (long long)blockIdx.x * kT + threadIdx.x;
    if (VEC) {
        const long long v = 2 * t;
        if (v >= nvox) return;
        float a = 0.f, b = 0.f;
        for (int c = 0; c < nc; ++c) {
            const float4 q = *reinterpret_cast<const float4*>(x + (long long)c * nvox + v);
            a += hypotf(q.x, q.y);
            b += hypotf(q.z, q.w);
        }
        *reinterpret_cast<float2*>(out + v) = make_float2(sqrtf(a), sqrtf(b));
    } else {
        if (t >= nvox) return;
        float a = 0.f;
        for (int c = 0; c < nc; ++c) {
            const float2 q = x[(long long)c * nvox + t];
            a += hypotf(q.x, q.y);
        }
        out[t] = sqrtf(a);
    }
}

// magnitude and phase of a complex image (processNIFTI.hpp:110-111: abs, arg)
__global__ void __launch_bounds__(kT) k_mag_phase(const float2* __restrict__ x, long long n,
                                                   float* __restrict__ mag,
                                                   float* __restrict__ phs)
{
    const long long i = (long long)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const float2 v = x[i];
    mag[i] = hypotf(v.x, v.y);
    phs[i] = atan2f(v.y, v.x);
}

unsigned nblocks(long long n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

namespace pgr {

void dcf_update(float2* d, const float2* c, long long M, hipStream_t s)
{
    hipLaunchKernelGGL(k_dcf_update, dim3(nblocks(M)), dim3(kT), 0, s, d, c, M);
}

void dcf_finish(const float2* d, long long M, double* part, float* out, hipStream_t s)
{
    const int np = (int)std::min<long long>(kMaxParts, std::max<long long>(1, (M + kT - 1) / kT));
    hipLaunchKernelGGL(k_norm_parts, dim3(np), dim3(kT), 0, s, d, M, part);
    hipLaunchKernelGGL(k_dcf_finish, dim3(nblocks(M)), dim3(kT), 0, s, d, M, (const double*)part,
                       np, out);
}

void weight_columns(float2* out, const float* dcf, const float2* data, long long M, int nc,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_weight_columns, dim3(nblocks(M)), dim3(kT), 0, s, out, dcf, data, M, nc);
}

void mag_phase(const float2* x, long long n, float* mag, float* phs, hipStream_t s)
{
    hipLaunchKernelGGL(k_mag_phase, dim3(nblocks(n)), dim3(kT), 0, s, x, n, mag, phs);
}

void sum_of_magnitudes(const float2* x, long long nvox, int nc, float* out, hipStream_t s)
{
    const bool vec = (nvox % 2 == 0) && ((uintptr_t)x % 16 == 0) && ((uintptr_t)out % 8 == 0);
    if (vec)
        hipLaunchKernelGGL(k_sum_of_magnitudes<true>, dim3(nblocks((nvox + 1) / 2)), dim3(kT), 0,
                           s, x, nvox, nc, out);
    else
        hipLaunchKernelGGL(k_sum_of_magnitudes<false>, dim3(nblocks(nvox)), dim3(kT), 0, s, x,
                           nvox, nc, out);
}

}  // namespace pgr
