// pg_tsint.hip -- the per-sample parts of the min-max time-segmentation
// interpolators.  The exact least-squares one (TimeSegmentation.cpp:143-149):
//   AA[l][i] = conj( sum_j W[l][j] exp(i fm_j t_i) ),  i < Mo, l < L,
// with W = inv(G^H G) G^H formed on the host.  One (sample, voxel) pair
// costs one FP64 sincos and L complex FP64 multiply-adds; it runs once per
// plan, in FP64 (the interpolator's Gram matrix is ill-conditioned, so its
// weights are formed in double and stored as FP32, like every other table).
// The histogram one (interptype 3) sums over the field map's 256 histogram
// bins instead of its voxels.
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

// TimeSegmentation.cpp:196-203 (interptype 3) per sample: the L+1 low
// outputs of the KK-point FFT of h_n exp(i n dwn t), shifted by
// exp(i (min + dwn/2)(t - k tau)), are c_k(t) = sum_n h_n exp(i om_n (t - k tau))
// with om_n = min + (n + 1/2) dwn; the powers z_n^k come by repeated
// multiplication (L <= 16 steps).  One thread per sample.
__global__ void __launch_bounds__(kT) k_histogram_weights(
    const double* __restrict__ h, const double* __restrict__ om, const double2* __restrict__ z,
    int nb, const double* __restrict__ tv, int Mo, const double2* __restrict__ iG, int L,
    float2* __restrict__ AA, long long M, int shots)
{
    constexpr int NB = pgt::kHistBins, NC = pgt::kHistCoef;
    __shared__ double s_h[NB], s_om[NB];
    __shared__ double2 s_z[NB];
    __shared__ double2 s_g[NC * NC];
    for (int e = threadIdx.x; e < NB; e += kT) {
        s_h[e] = (e < nb) ? h[e] : 0.0;
        s_om[e] = (e < nb) ? om[e] : 0.0;
        s_z[e] = (e < nb) ? z[e] : make_double2(1.0, 0.0);
    }
    const int nc = L + 1;
    for (int e = threadIdx.x; e < NC * NC; e += kT)
        s_g[e] = (e < nc * nc) ? iG[e] : make_double2(0.0, 0.0);
    __syncthreads();
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= Mo) return;
    const double ti = tv[i];
    double cr[NC], ci[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) cr[k] = ci[k] = 0.0;
    for (int n = 0; n < nb; ++n) {
        double sn, cs;
        sincos(s_om[n] * ti, &sn, &cs);
        double br = s_h[n] * cs, bi = s_h[n] * sn;
        const double2 zz = s_z[n];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (k < nc) {
                cr[k] += br;
                ci[k] += bi;
                const double nr = br * zz.x - bi * zz.y;
                bi = br * zz.y + bi * zz.x;
                br = nr;
            }
        }
    }
    for (int l = 0; l < L; ++l) {
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (k < nc) {
                const double2 g = s_g[l * nc + k];
                ar += g.x * cr[k] - g.y * ci[k];
                ai += g.x * ci[k] + g.y * cr[k];
            }
        }
        const float2 v = make_float2((float)ar, (float)-ai);
        for (int sh = 0; sh < shots; ++sh) AA[(long long)l * M + i + (long long)sh * Mo] = v;
    }
}

}  // namespace

namespace pgt {

void histogram_weights(const double* h, const double* om, const double2* z, int nb,
                       const double* tv, int Mo, const double2* iG, int L, int shots, float2* AA,
                       hipStream_t s)
{
    if (nb < 1 || nb > kHistBins || L < 1 || L + 1 > kHistCoef || Mo < 1) return;
    const long long M = (long long)Mo * shots;
    hipLaunchKernelGGL(k_histogram_weights, dim3((unsigned)((Mo + kT - 1) / kT)), dim3(kT), 0, s,
                       h, om, z, nb, tv, Mo, iG, L, AA, M, shots);
}

void minmax_weights(const double* fm, const double* tv, const double2* W, int N, int Mo, int L,
                    int shots, float2* AA, hipStream_t s)
{
    const long long M = (long long)Mo * shots;
    for (int l0 = 0; l0 < L; l0 += kLC)
        hipLaunchKernelGGL(k_minmax_weights, dim3((unsigned)((Mo + kT - 1) / kT)), dim3(kT), 0, s,
                           fm, tv, W, N, Mo, L, l0, AA, M, shots);
}

}  // namespace pgt
