// pg_dft.hip -- exact field-corrected DFT (Gdft, Gdft.cpp:52-136 /
// ftCpu.cpp:94-215), batched over SENSE coils.
//   forward: y_c[i] = sum_j S_c[j] x[j] exp(-i (2 pi k_i . r_j + fm_j t_i))
//   adjoint: x[j]   = sum_c conj(S_c[j]) sum_i y_c[i] exp(+i (...))
// The phase of a (sample, pixel) pair is computed once and used for every
// coil.  Work is split over (output block, input slice); slices are summed
// by a second kernel in a fixed order.
// pcSENSE's shots (pcSENSE.cpp:108-153) run as one launch: the M samples
// are nshot shots of M/nshot, shot-major, and shot s has its own nc maps at
// S + s * s_shot.  Forward sample tiles and adjoint sample slices never
// straddle a shot, so a block applies one shot's maps.
//
// precision 0 -- the reference's arithmetic class (FP32 phase, FP32 trig):
//   the phase is formed in revolutions, p = k.r + fm t / (2 pi), reduced to
//   [-1/2, 1/2] and fed to the hardware v_sin_f32 / v_cos_f32 (which take
//   revolutions); FP32 accumulation.
// precision 1 -- FP64 phase, FP64 sin/cos, FP64 accumulation (the exact
//   NUDFT; used to synthesise data and as an accuracy reference).
// GdftR2 (GdftR2.cpp, ftCpuWithGrads.cpp:98-212) multiplies every pair by
//   b = sinc(kx/Nx + Gx t) sinc(ky/Ny + Gy t) sinc(kz/Nz + Gz t),
//   sinc(x) = sin(pi x)/(pi x) (1 below 1e-4), with the field-map difference
//   maps G staged beside the pixel coordinates.
#include "pg_dft.h"

#include <mutex>

namespace {

// partial sums of pg_dft_apply_static: every element a launch reads was
// written earlier in the same launch sequence, so nothing is cleared
constexpr size_t kScratchElems = (size_t)8 << 20;  // 8M float2 = 64 MiB
__device__ __align__(256) float2 scratch[kScratchElems];

constexpr int kT = 64;  // threads per block (one wavefront) = tile width

// sin(pi x)/(pi x) in FP32 with the hardware sine (which takes revolutions:
// sin(pi x) = sin(2 pi (x/2)), reduced to [-1/2, 1/2] turns first)
__device__ __forceinline__ float sinc_f(float x)
{
    if (fabsf(x) < 1e-4f) return 1.0f;
    float r = 0.5f * x;
    r = r - rintf(r);
    return __builtin_amdgcn_sinf(r) / (3.14159265358979f * x);
}

__device__ __forceinline__ double sinc_d(double x)
{
    if (fabs(x) < 1e-4) return 1.0;
    return sin(M_PI * x) / (M_PI * x);
}

struct R2Args {
    const float *gx, *gy, *gz;
    float inx, iny, inz;  // 1/Nx, 1/Ny, 1/Nz
};

template <int NCB, bool F64, bool R2>
__global__ void __launch_bounds__(kT) k_gdft_fwd(
    const float* __restrict__ kx, const float* __restrict__ ky, const float* __restrict__ kz,
    const float* __restrict__ t, const float* __restrict__ ix, const float* __restrict__ iy,
    const float* __restrict__ iz, const float* __restrict__ fm, const float2* __restrict__ x,
    const float2* __restrict__ S, int nc, int M, int N, int slice_len, int shot_len,
    long long s_shot, float2* __restrict__ part, R2Args r2)
{
    __shared__ float s_r[kT][4];
    __shared__ float s_g[R2 ? kT : 1][3];
    __shared__ float2 s_v[kT][NCB];
    // ceil(shot_len / kT) tiles per shot
    const int tps = (shot_len + kT - 1) / kT;
    const int sh = blockIdx.x / tps;
    const int il = (blockIdx.x - sh * tps) * kT + threadIdx.x;
    const bool live = il < shot_len;
    const int i = sh * shot_len + il;
    if (S) S += sh * s_shot;
    const int c0 = blockIdx.z * NCB;
    const int j0 = blockIdx.y * slice_len;
    const int j1 = min(N, j0 + slice_len);
    const float tw = 6.283185307179586f;
    float ki[3] = {0, 0, 0};
    float ti = 0;
    if (live) {
        ki[0] = kx[i];
        ki[1] = ky[i];
        ki[2] = kz[i];
        ti = t[i];
    }
    double ar[NCB], ai[NCB];
#pragma unroll
    for (int c = 0; c < NCB; ++c) ar[c] = ai[c] = 0.0;
    float fr[NCB], fi[NCB];
#pragma unroll
    for (int c = 0; c < NCB; ++c) fr[c] = fi[c] = 0.f;

    for (int jb = j0; jb < j1; jb += kT) {
        __syncthreads();
        const int j = jb + threadIdx.x;
        if (j < j1) {
            s_r[threadIdx.x][0] = ix[j];
            s_r[threadIdx.x][1] = iy[j];
            s_r[threadIdx.x][2] = iz[j];
            s_r[threadIdx.x][3] = F64 ? fm[j] : (float)((double)fm[j] / (2.0 * M_PI));
            if (R2) {
                s_g[threadIdx.x][0] = r2.gx[j];
                s_g[threadIdx.x][1] = r2.gy[j];
                s_g[threadIdx.x][2] = r2.gz[j];
            }
            const float2 xv = x[j];
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                float2 v = xv;
                if (S && c0 + c < nc) {
                    const float2 s = S[(long long)(c0 + c) * N + j];
                    v = make_float2(s.x * xv.x - s.y * xv.y, s.x * xv.y + s.y * xv.x);
                }
                s_v[threadIdx.x][c] = v;
            }
        }
        __syncthreads();
        const int cnt = min(kT, j1 - jb);
        for (int q = 0; q < cnt; ++q) {
            float cs, sn;
            if (F64) {
                const double ph = 2.0 * M_PI * ((double)ki[0] * s_r[q][0] + (double)ki[1] * s_r[q][1] +
                                                (double)ki[2] * s_r[q][2]) +
                                  (double)s_r[q][3] * ti;
                double sd, cd;
                sincos(ph, &sd, &cd);
                if (R2) {
                    const double b = sinc_d((double)ki[0] * r2.inx + (double)s_g[q][0] * ti) *
                                     sinc_d((double)ki[1] * r2.iny + (double)s_g[q][1] * ti) *
                                     sinc_d((double)ki[2] * r2.inz + (double)s_g[q][2] * ti);
                    sd *= b;
                    cd *= b;
                }
#pragma unroll
                for (int c = 0; c < NCB; ++c) {
                    const float2 v = s_v[q][c];
                    ar[c] += cd * v.x + sd * v.y;
                    ai[c] += -sd * v.x + cd * v.y;
                }
            } else {
                float p = ki[0] * s_r[q][0] + ki[1] * s_r[q][1] + ki[2] * s_r[q][2] + s_r[q][3] * ti;
                p = p - rintf(p);
                sn = __builtin_amdgcn_sinf(p);
                cs = __builtin_amdgcn_cosf(p);
                if (R2) {
                    const float b = sinc_f(ki[0] * r2.inx + s_g[q][0] * ti) *
                                    sinc_f(ki[1] * r2.iny + s_g[q][1] * ti) *
                                    sinc_f(ki[2] * r2.inz + s_g[q][2] * ti);
                    sn *= b;
                    cs *= b;
                }
#pragma unroll
                for (int c = 0; c < NCB; ++c) {
                    const float2 v = s_v[q][c];
                    fr[c] += cs * v.x + sn * v.y;
                    fi[c] += -sn * v.x + cs * v.y;
                }
            }
        }
    }
    (void)tw;
    if (live) {
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
            if (c0 + c >= nc) break;
            const float2 r = F64 ? make_float2((float)ar[c], (float)ai[c]) : make_float2(fr[c], fi[c]);
            part[((long long)blockIdx.y * nc + c0 + c) * M + i] = r;
        }
    }
}

template <int NCB, bool F64, bool R2>
__global__ void __launch_bounds__(kT) k_gdft_adj(
    const float* __restrict__ kx, const float* __restrict__ ky, const float* __restrict__ kz,
    const float* __restrict__ t, const float* __restrict__ ix, const float* __restrict__ iy,
    const float* __restrict__ iz, const float* __restrict__ fm, const float2* __restrict__ y,
    long long ystride, const float2* __restrict__ S, int nc, int M, int N, int slice_len,
    int shot_len, long long s_shot, float2* __restrict__ part, R2Args r2)
{
    __shared__ float s_k[kT][4];
    __shared__ float2 s_v[kT][NCB];
    const int j = blockIdx.x * kT + threadIdx.x;
    // gridDim.y / nshot slices per shot, each within its shot
    const int sps = (int)gridDim.y / (M / shot_len);
    const int sh = blockIdx.y / sps;
    const int i0 = sh * shot_len + (blockIdx.y - sh * sps) * slice_len;
    const int i1 = min(sh * shot_len + shot_len, i0 + slice_len);
    if (S) S += sh * s_shot;
    float rj[3] = {0, 0, 0};
    float fj = 0;
    float gj[3] = {0, 0, 0};
    if (j < N) {
        rj[0] = ix[j];
        rj[1] = iy[j];
        rj[2] = iz[j];
        fj = F64 ? fm[j] : (float)((double)fm[j] / (2.0 * M_PI));
        if (R2) {
            gj[0] = r2.gx[j];
            gj[1] = r2.gy[j];
            gj[2] = r2.gz[j];
        }
    }
    double xr = 0, xi = 0;
    float gr = 0, gi = 0;
    for (int c0 = 0; c0 < nc; c0 += NCB) {
        double ar[NCB], ai[NCB];
        float fr[NCB], fi[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
            ar[c] = ai[c] = 0;
            fr[c] = fi[c] = 0;
        }
        for (int ib = i0; ib < i1; ib += kT) {
            __syncthreads();
            const int i = ib + threadIdx.x;
            if (i < i1) {
                s_k[threadIdx.x][0] = kx[i];
                s_k[threadIdx.x][1] = ky[i];
                s_k[threadIdx.x][2] = kz[i];
                s_k[threadIdx.x][3] = t[i];
#pragma unroll
                for (int c = 0; c < NCB; ++c)
                    s_v[threadIdx.x][c] = (c0 + c < nc) ? y[(long long)(c0 + c) * ystride + i]
                                                        : make_float2(0.f, 0.f);
            }
            __syncthreads();
            const int cnt = min(kT, i1 - ib);
            for (int q = 0; q < cnt; ++q) {
                if (F64) {
                    const double ph = 2.0 * M_PI * ((double)s_k[q][0] * rj[0] + (double)s_k[q][1] * rj[1] +
                                                    (double)s_k[q][2] * rj[2]) +
                                      (double)fj * s_k[q][3];
                    double sd, cd;
                    sincos(ph, &sd, &cd);
                    if (R2) {
                        const double tq = s_k[q][3];
                        const double b = sinc_d((double)s_k[q][0] * r2.inx + (double)gj[0] * tq) *
                                         sinc_d((double)s_k[q][1] * r2.iny + (double)gj[1] * tq) *
                                         sinc_d((double)s_k[q][2] * r2.inz + (double)gj[2] * tq);
                        sd *= b;
                        cd *= b;
                    }
#pragma unroll
                    for (int c = 0; c < NCB; ++c) {
                        const float2 v = s_v[q][c];
                        ar[c] += cd * v.x - sd * v.y;
                        ai[c] += sd * v.x + cd * v.y;
                    }
                } else {
                    float p = s_k[q][0] * rj[0] + s_k[q][1] * rj[1] + s_k[q][2] * rj[2] + fj * s_k[q][3];
                    p = p - rintf(p);
                    float sn = __builtin_amdgcn_sinf(p);
                    float cs = __builtin_amdgcn_cosf(p);
                    if (R2) {
                        const float tq = s_k[q][3];
                        const float b = sinc_f(s_k[q][0] * r2.inx + gj[0] * tq) *
                                        sinc_f(s_k[q][1] * r2.iny + gj[1] * tq) *
                                        sinc_f(s_k[q][2] * r2.inz + gj[2] * tq);
                        sn *= b;
                        cs *= b;
                    }
#pragma unroll
                    for (int c = 0; c < NCB; ++c) {
                        const float2 v = s_v[q][c];
                        fr[c] += cs * v.x - sn * v.y;
                        fi[c] += sn * v.x + cs * v.y;
                    }
                }
            }
        }
        if (j < N) {
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                if (c0 + c >= nc) break;
                double vr = F64 ? ar[c] : fr[c], vi = F64 ? ai[c] : fi[c];
                if (S) {
                    const float2 s = S[(long long)(c0 + c) * N + j];
                    const double tr = vr * s.x + vi * s.y;  // v * conj(s)
                    const double ti = vi * s.x - vr * s.y;
                    vr = tr;
                    vi = ti;
                }
                if (F64) {
                    xr += vr;
                    xi += vi;
                } else {
                    gr += (float)vr;
                    gi += (float)vi;
                }
            }
        }
    }
    if (j < N) part[(long long)blockIdx.y * N + j] = F64 ? make_float2((float)xr, (float)xi) : make_float2(gr, gi);
}

// out[(i / row) * ostride + i % row] (+)= sum of the slices of element i:
// each run of `group` consecutive slices is summed in slice order, and the
// runs are added in order
__global__ void k_sum_slices(const float2* part, int nslice, int group, long long n,
                             long long row, long long ostride, int acc, float2* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 s = make_float2(0.f, 0.f);
    for (int g0 = 0; g0 < nslice; g0 += group) {
        float2 t = part[(long long)g0 * n + i];
        for (int k = g0 + 1; k < g0 + group; ++k) {
            const float2 p = part[(long long)k * n + i];
            t.x += p.x;
            t.y += p.y;
        }
        s = g0 ? make_float2(s.x + t.x, s.y + t.y) : t;
    }
    const long long r = i / row;
    const long long o = r * ostride + (i - r * row);
    if (acc) {
        const float2 a = out[o];
        s = make_float2(a.x + s.x, a.y + s.y);
    }
    out[o] = s;
}

R2Args r2_args(const pgd::DftData& d)
{
    return R2Args{d.gx, d.gy, d.gz, 1.0f / (float)d.numx, 1.0f / (float)d.numy,
                  1.0f / (float)d.numz};
}

// slices per shot: aim for >= 4096 wavefronts in flight, each slice at
// least 64 inputs (small problems, such as pcSENSE's DFTs of ~1k samples a
// shot, would otherwise leave most of the 1024 SIMDs idle)
int slices_for(long long outputs, long long inputs, int nshot = 1)
{
    long long waves = (outputs + kT - 1) / kT;
    int s = 1;
    while (waves * nshot * s < 4096 && inputs / (s * 2) >= 64 && nshot * s < 64) s *= 2;
    return s;
}

}  // namespace

namespace pgd {

size_t part_elems(int M, int N, int nc, bool adjoint, int nshot)
{
    if (adjoint) return (size_t)slices_for(N, M / nshot, nshot) * nshot * N;
    return (size_t)slices_for((long long)M, N) * nc * M;
}

void forward(const DftData& d, const float2* x, const float2* S, int nc, float2* out, float2* part,
             hipStream_t s, long long ostride, int nshot, long long s_shot)
{
    if (ostride < 0) ostride = d.M;
    if (nshot < 1 || d.M % nshot)
        throw pg::Error{PG_ERR_ARG, "the DFT's shots must split its samples evenly"};
    const int shot_len = d.M / nshot;
    const int ns = slices_for(d.M, d.N);
    const int slen = (d.N + ns - 1) / ns;
    constexpr int NCB = 8;
    dim3 grid(nshot * ((shot_len + kT - 1) / kT), ns, (nc + NCB - 1) / NCB);
    const R2Args r2 = r2_args(d);
#define PG_GDFT_FWD(F64, R2)                                                                     \
    hipLaunchKernelGGL((k_gdft_fwd<NCB, F64, R2>), grid, dim3(kT), 0, s, d.kx, d.ky, d.kz, d.t,  \
                       d.ix, d.iy, d.iz, d.fm, x, S, nc, d.M, d.N, slen, shot_len, s_shot, part, \
                       r2)
    if (d.precision == 1) {
        if (d.gx) PG_GDFT_FWD(true, true);
        else PG_GDFT_FWD(true, false);
    } else {
        if (d.gx) PG_GDFT_FWD(false, true);
        else PG_GDFT_FWD(false, false);
    }
#undef PG_GDFT_FWD
    const long long n = (long long)nc * d.M;
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, ns, ns,
                       n, (long long)d.M, ostride, 0, out);
}

void adjoint(const DftData& d, const float2* y, const float2* S, int nc, float2* out, float2* part,
             hipStream_t s, long long ystride, bool accumulate, int nshot, long long s_shot)
{
    if (ystride < 0) ystride = d.M;
    if (nshot < 1 || d.M % nshot)
        throw pg::Error{PG_ERR_ARG, "the DFT's shots must split its samples evenly"};
    const int shot_len = d.M / nshot;
    const int sps = slices_for(d.N, shot_len, nshot);
    const int slen = (shot_len + sps - 1) / sps;
    constexpr int NCB = 8;
    dim3 grid((d.N + kT - 1) / kT, nshot * sps, 1);
    const R2Args r2 = r2_args(d);
#define PG_GDFT_ADJ(F64, R2)                                                                     \
    hipLaunchKernelGGL((k_gdft_adj<NCB, F64, R2>), grid, dim3(kT), 0, s, d.kx, d.ky, d.kz, d.t,  \
                       d.ix, d.iy, d.iz, d.fm, y, ystride, S, nc, d.M, d.N, slen, shot_len,      \
                       s_shot, part, r2)
    if (d.precision == 1) {
        if (d.gx) PG_GDFT_ADJ(true, true);
        else PG_GDFT_ADJ(true, false);
    } else {
        if (d.gx) PG_GDFT_ADJ(false, true);
        else PG_GDFT_ADJ(false, false);
    }
#undef PG_GDFT_ADJ
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((d.N + 255) / 256)), dim3(256), 0, s, part,
                       nshot * sps, sps, (long long)d.N, (long long)d.N, (long long)d.N,
                       accumulate ? 1 : 0, out);
}

size_t static_scratch_elems() { return kScratchElems; }

float2* static_scratch()
{
    constexpr int kMaxDev = 64;
    static std::mutex mu;
    static float2* addr[kMaxDev] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev)
        throw pg::Error{PG_ERR_HIP, "no current device for the static scratch"};
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw pg::Error{PG_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                            hipGetErrorString(e)};
        addr[dev] = (float2*)p;
    }
    return addr[dev];
}

}  // namespace pgd
