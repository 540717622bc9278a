// pg_tsint.hip -- the data-dependent part of the exact least-squares min-max
// time-segmentation interpolator (TimeSegmentation.cpp:143-149):
//   AA[l][i] = conj( sum_j W[l][j] exp(i fm_j t_i) ),  i < Mo, l < L,
// with W = inv(G^H G) G^H formed on the host.  One (sample, voxel) pair
// costs one FP64 sincos and L complex FP64 multiply-adds; it runs once per
// plan, in FP64 (the interpolator's Gram matrix is ill-conditioned, so its
// weights are formed in double and stored as FP32, like every other table).
#include "pg_tsint.h"

namespace {

constexpr int kT = 256;   // threads (samples) per block
constexpr int kTile = 128;  // voxels staged per step
constexpr int kLC = 8;    // segments per pass

__global__ void __launch_bounds__(kT) k_minmax_weights(const double* __restrict__ fm,
                                                        const double* __restrict__ tv,
                                                        const double2* __restrict__ W, int N,
                                                        int Mo, int L, int l0,
                                                        float2* __restrict__ AA, long long M,
                                                        int shots)
{
    __shared__ double s_fm[kTile];
    __shared__ double2 s_w[kLC][kTile];
    const int i = blockIdx.x * kT + threadIdx.x;
    const double ti = (i < Mo) ? tv[i] : 0.0;
    const int nl = min(kLC, L - l0);
    double ar[kLC], ai[kLC];
#pragma unroll
    for (int l = 0; l < kLC; ++l) ar[l] = ai[l] = 0.0;
    for (int j0 = 0; j0 < N; j0 += kTile) {
        const int cnt = min(kTile, N - j0);
        __syncthreads();
        for (int e = threadIdx.x; e < kTile; e += kT) s_fm[e] = (e < cnt) ? fm[j0 + e] : 0.0;
        for (int e = threadIdx.x; e < kLC * kTile; e += kT) {
            const int l = e / kTile, j = e - l * kTile;
            s_w[l][j] = (l < nl && j < cnt) ? W[(long long)(l0 + l) * N + j0 + j]
                                            : make_double2(0.0, 0.0);
        }
        __syncthreads();
        for (int j = 0; j < cnt; ++j) {
            double sn, cs;
            sincos(s_fm[j] * ti, &sn, &cs);
#pragma unroll
            for (int l = 0; l < kLC; ++l) {
                const double2 w = s_w[l][j];
                ar[l] += w.x * cs - w.y * sn;
                ai[l] += w.x * sn + w.y * cs;
            }
        }
    }
    if (i >= Mo) return;
    for (int l = 0; l < nl; ++l) {
        const float2 v = make_float2((float)ar[l], (float)-ai[l]);
        for (int sh = 0; sh < shots; ++sh) AA[(long long)(l0 + l) * M + i + (long long)sh * Mo] = v;
    }
}

}  // namespace

namespace pgt {

void minmax_weights(const double* fm, const double* tv, const double2* W, int N, int Mo, int L,
                    int shots, float2* AA, hipStream_t s)
{
    const long long M = (long long)Mo * shots;
    for (int l0 = 0; l0 < L; l0 += kLC)
        hipLaunchKernelGGL(k_minmax_weights, dim3((unsigned)((Mo + kT - 1) / kT)), dim3(kT), 0, s,
                           fm, tv, W, N, Mo, L, l0, AA, M, shots);
}

}  // namespace pgt
