// pg_fft.hip -- oversampled-grid FFT passes of the gridding NUFFT for gfx950.
//
// The reference pipeline (gridding.cpp:584-727 forward, :418-581 adjoint) is
//   forward: deapodize -> zero_pad -> fftshift -> FFT -> ifftshift -> interp
//   adjoint: spread -> ifftshift -> IFFT -> fftshift -> crop -> deapodize
// with unnormalised transforms (fftCPU.cpp:48-106) and even grid sizes, so
// every shift is a circular shift by G/2 (griddingSupport.cpp:599-627).
// Here the shifts, pad and crop are index arithmetic inside the FFT passes,
// and only the non-zero lines of the padded grid are transformed:
//   forward: [prep + FFT_y on the Nx(*Nz) image columns]  -> T
//            [FFT_x on rows of T (Nx non-zero inputs)]     -> grid (2-D)
//            (3-D: FFT_x -> U, then FFT_z (Nz non-zero inputs) -> grid)
//   adjoint: [IFFT_x (keep the Nx cropped outputs)]        -> T   (3-D: IFFT_z first)
//            [IFFT_y + crop + deapod + WoH + sum_l + conj(S)] -> image
// Each 1-D transform is a mixed radix-4/2/3 Stockham FFT staged in LDS.
#include "pg_internal.h"
#include "pg_gnufft.h"

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// In-LDS Stockham FFT of `nlines` lines of length f.n (line i at A + i*pitch).
// Returns the buffer holding the result (A or B).  SIGN = -1: forward
// (e^{-i}), +1: backward (e^{+i}); unnormalised in both directions.
template <int SIGN>
__device__ float2* lds_fft(float2* A, float2* B, int pitch, int nlines, const pg::FftArgs& f)
{
    const int n = f.n;
    int Ns = 1;
    for (int st = 0; st < f.nst; ++st) {
        const int R = f.radix[st];
        const int nb = n / R;
        const int total = nb * nlines;
        const int tstride = n / (Ns * R);
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            const int line = t / nb;
            const int j = t - line * nb;
            const int k = j % Ns;
            const float2* in = A + line * pitch;
            float2* out = B + line * pitch;
            float2 v0 = in[j], v1 = in[j + nb], v2, v3;
            float2 w1 = f.tw[k * tstride];
            if (SIGN > 0) w1.y = -w1.y;
            v1 = cmul(v1, w1);
            if (R >= 3) {
                float2 w2 = f.tw[2 * k * tstride];
                if (SIGN > 0) w2.y = -w2.y;
                v2 = cmul(in[j + 2 * nb], w2);
            }
            if (R == 4) {
                float2 w3 = f.tw[3 * k * tstride];
                if (SIGN > 0) w3.y = -w3.y;
                v3 = cmul(in[j + 3 * nb], w3);
            }
            const int base = (j - k) * R + k;
            if (R == 4) {
                float2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
                float2 t2 = cadd(v1, v3), d = csub(v1, v3);
                // t3 = (SIGN * i) * (v1 - v3)
                float2 t3 = make_float2(-SIGN * d.y, SIGN * d.x);
                out[base] = cadd(t0, t2);
                out[base + Ns] = cadd(t1, t3);
                out[base + 2 * Ns] = csub(t0, t2);
                out[base + 3 * Ns] = csub(t1, t3);
            } else if (R == 2) {
                out[base] = cadd(v0, v1);
                out[base + Ns] = csub(v0, v1);
            } else {  // R == 3
                const float h = 0.86602540378443864676f;
                float2 t1 = cadd(v1, v2);
                float2 t2 = make_float2(v0.x - 0.5f * t1.x, v0.y - 0.5f * t1.y);
                float2 u = make_float2(SIGN * h * (v1.x - v2.x), SIGN * h * (v1.y - v2.y));
                out[base] = cadd(v0, t1);
                out[base + Ns] = make_float2(t2.x - u.y, t2.y + u.x);
                out[base + 2 * Ns] = make_float2(t2.x + u.y, t2.y - u.x);
            }
        }
        __syncthreads();
        float2* tmp = A;
        A = B;
        B = tmp;
        Ns *= R;
    }
    return A;
}

// ------------------------------------------------------------------------
// Forward prep + FFT along y for CPW image columns (x, z) of one grid b.
//   v = deapod( Wo_l * (S_c * x) )    (SENSE.cpp:59-63, TimeSegmentation.cpp:266-270,
//                                      griddingSupport.cpp:262-274)
// placed at gy = (y + offY + Gy/2) % Gy (zero_pad + fftshift), FFT_y, and
// written at (k + Gy/2) % Gy (ifftshift) into T[b][col][.].
__global__ void __launch_bounds__(256) k_fwd_prep_colfft(
    const float2* __restrict__ img, long long img_cstride, const float2* __restrict__ S,
    const float2* __restrict__ Wo, const float* __restrict__ dfx, const float* __restrict__ dfy,
    const float* __restrict__ dfz, float inv_osd, int Nx, int Ny, int Nz, int offY, int L,
    int CPW, pg::FftArgs fy, float2* __restrict__ T)
{
    extern __shared__ float2 lds[];
    const int Gy = fy.n;
    const int pitch = Gy + 1;
    float2* A = lds;
    float2* B = lds + CPW * pitch;
    const int b = blockIdx.y;
    const int c = b / L, l = b - c * L;
    const int ncol = Nx * Nz;
    const int col0 = blockIdx.x * CPW;
    const long long N = (long long)Nx * Ny * Nz;

    for (int t = threadIdx.x; t < CPW * pitch; t += blockDim.x) A[t] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int t = threadIdx.x; t < CPW * Ny; t += blockDim.x) {
        const int cl = t / Ny, y = t - cl * Ny;
        const int col = col0 + cl;
        if (col >= ncol) continue;
        const int x = col % Nx, z = col / Nx;
        const long long nidx = (long long)y + (long long)col * Ny;
        float2 v = img[c * img_cstride + nidx];
        if (S) v = cmul(S[(long long)c * N + nidx], v);
        if (Wo) v = cmul(Wo[(long long)l * N + nidx], v);
        float kern = dfx[x] * dfy[y];
        if (Nz > 1) kern = kern * dfz[z];
        if (isnan(kern)) {
            v = make_float2(0.f, 0.f);
        } else {
            v.x = (v.x / kern) * inv_osd;
            v.y = (v.y / kern) * inv_osd;
        }
        int gy = y + offY + Gy / 2;
        if (gy >= Gy) gy -= Gy;
        A[cl * pitch + gy] = v;
    }
    __syncthreads();
    float2* R = lds_fft<-1>(A, B, pitch, CPW, fy);
    const long long Tb = (long long)b * ncol * Gy;
    for (int t = threadIdx.x; t < CPW * Gy; t += blockDim.x) {
        const int cl = t / Gy, k = t - cl * Gy;
        const int col = col0 + cl;
        if (col >= ncol) continue;
        int o = k + Gy / 2;
        if (o >= Gy) o -= Gy;
        T[Tb + (long long)col * Gy + o] = R[cl * pitch + k];
    }
}

// ------------------------------------------------------------------------
// Generic FFT along a strided axis.  Element (outer, i, inner) of the source
// is src[outer*src_os + i*inner_n + inner]; input entry i (i < nin) is placed
// at (i + in_off) % n, output entry o (o < nout) is taken from (o + out_off) % n
// and stored at dst[outer*dst_os + o*inner_n + inner].  A block handles RPW
// consecutive inner indices of one outer index.
template <int SIGN>
__global__ void __launch_bounds__(256) k_line_fft(const float2* __restrict__ src,
                                                  float2* __restrict__ dst, long long inner_n,
                                                  long long src_os, long long dst_os, int nin,
                                                  int in_off, int nout, int out_off, int RPW,
                                                  pg::FftArgs f)
{
    extern __shared__ float2 lds[];
    const int n = f.n;
    const int pitch = n + 1;
    float2* A = lds;
    float2* B = lds + RPW * pitch;
    const long long outer = blockIdx.y;
    const long long inner0 = (long long)blockIdx.x * RPW;
    const float2* s = src + outer * src_os + inner0;
    float2* d = dst + outer * dst_os + inner0;
    const int nv = (int)((inner_n - inner0) < RPW ? (inner_n - inner0) : RPW);

    if (nin < n)
        for (int t = threadIdx.x; t < RPW * pitch; t += blockDim.x) A[t] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int t = threadIdx.x; t < RPW * nin; t += blockDim.x) {
        const int j = t % RPW, i = t / RPW;
        if (j >= nv) continue;
        int p = i + in_off;
        p %= n;
        A[j * pitch + p] = s[(long long)i * inner_n + j];
    }
    __syncthreads();
    float2* R = lds_fft<SIGN>(A, B, pitch, RPW, f);
    for (int t = threadIdx.x; t < RPW * nout; t += blockDim.x) {
        const int j = t % RPW, o = t / RPW;
        if (j >= nv) continue;
        int p = o + out_off;
        p %= n;
        d[(long long)o * inner_n + j] = R[j * pitch + p];
    }
}

// ------------------------------------------------------------------------
// Adjoint tail: IFFT along y of CPW columns of T (inputs placed at
// (ny + Gy/2) % Gy, the ifftshift), crop rows y -> (y + offY + Gy/2) % Gy
// (fftshift + crop_center_region), deapodize, multiply by WoH_l, sum over
// the L segments of coil c, multiply by conj(S_c) and store.
// (gridding.cpp:505-576, TimeSegmentation.cpp:313-323, SENSE.cpp:84-92)
__global__ void __launch_bounds__(256) k_adj_colfft_post(
    const float2* __restrict__ T, const float2* __restrict__ S, const float2* __restrict__ WoH,
    const float* __restrict__ dfx, const float* __restrict__ dfy, const float* __restrict__ dfz,
    float inv_osd, int Nx, int Ny, int Nz, int offY, int L, int CPW, pg::FftArgs fy,
    float2* __restrict__ out, long long out_cstride)
{
    extern __shared__ float2 lds[];
    const int Gy = fy.n;
    const int pitch = Gy + 1;
    float2* A = lds;
    float2* B = lds + CPW * pitch;
    const int c = blockIdx.y;
    const int ncol = Nx * Nz;
    const int col0 = blockIdx.x * CPW;
    const long long N = (long long)Nx * Ny * Nz;
    constexpr int MAXO = 4;  // outputs per thread: CPW * Ny <= 4 * blockDim
    float2 acc[MAXO];
#pragma unroll
    for (int r = 0; r < MAXO; ++r) acc[r] = make_float2(0.f, 0.f);

    for (int l = 0; l < L; ++l) {
        const long long b = (long long)c * L + l;
        const float2* Tb = T + b * ncol * Gy;
        for (int t = threadIdx.x; t < CPW * Gy; t += blockDim.x) {
            const int cl = t / Gy, m = t - cl * Gy;
            const int col = col0 + cl;
            int p = m + Gy / 2;
            if (p >= Gy) p -= Gy;
            A[cl * pitch + p] = (col < ncol) ? Tb[(long long)col * Gy + m] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        float2* R = lds_fft<+1>(A, B, pitch, CPW, fy);
#pragma unroll
        for (int r = 0; r < MAXO; ++r) {
            const int t = threadIdx.x + r * blockDim.x;
            if (t >= CPW * Ny) break;
            const int cl = t / Ny, y = t - cl * Ny;
            const int col = col0 + cl;
            if (col >= ncol) continue;
            const int x = col % Nx, z = col / Nx;
            int p = y + offY + Gy / 2;
            if (p >= Gy) p -= Gy;
            float2 q = R[cl * pitch + p];
            float kern = dfx[x] * dfy[y];
            if (Nz > 1) kern = kern * dfz[z];
            if (isnan(kern)) {
                q = make_float2(0.f, 0.f);
            } else {
                q.x = (q.x / kern) * inv_osd;
                q.y = (q.y / kern) * inv_osd;
            }
            const long long nidx = (long long)y + (long long)col * Ny;
            if (WoH) q = cmul(q, WoH[(long long)l * N + nidx]);
            acc[r] = cadd(acc[r], q);
        }
        __syncthreads();  // R is reused as the next segment's load buffer
    }
#pragma unroll
    for (int r = 0; r < MAXO; ++r) {
        const int t = threadIdx.x + r * blockDim.x;
        if (t >= CPW * Ny) break;
        const int cl = t / Ny, y = t - cl * Ny;
        const int col = col0 + cl;
        if (col >= ncol) continue;
        const long long nidx = (long long)y + (long long)col * Ny;
        float2 v = acc[r];
        if (S) {
            float2 s = S[(long long)c * N + nidx];
            s.y = -s.y;
            v = cmul(v, s);
        }
        out[c * out_cstride + nidx] = v;
    }
}

int pick_cpw(int Gy, int Ny)
{
    // enough butterflies for 256 threads, and CPW*Ny <= 4*256 outputs
    int cpw = 1;
    while (cpw * (Gy / 4) < 256 && cpw * 2 * Ny <= 1024) cpw *= 2;
    return cpw;
}

int pick_rpw(int n) { return n <= 256 ? 16 : 8; }

}  // namespace

namespace pgk {

void launch_fwd_prep_fft(const pg::Gnufft& G, const float2* img, long long img_cstride,
                         WeightArgs w, int nc, int L, float2* tbuf, float2* tbuf2, float2* grids,
                         hipStream_t s)
{
    const pg::Geometry& g = G.geo;
    const int B = nc * L;
    const int ncol = g.Nx * g.Nz;
    const int cpw = pick_cpw(g.Gy, g.Ny);
    const size_t lds1 = sizeof(float2) * 2 * cpw * (g.Gy + 1);
    hipLaunchKernelGGL(k_fwd_prep_colfft, dim3((ncol + cpw - 1) / cpw, B), dim3(256), lds1, s, img,
                       img_cstride, w.S, w.Wo, G.dfx, G.dfy, G.dfz, G.inv_osd, g.Nx, g.Ny, g.Nz,
                       g.offY, L, cpw, pg::fft_args(G.py), tbuf);
    // FFT along x: rows (b, z, gy), Nx non-zero inputs at (x + offX + Gx/2) % Gx
    const int rpw = pick_rpw(g.Gx);
    const size_t ldsx = sizeof(float2) * 2 * rpw * (g.Gx + 1);
    float2* xdst = (g.dims == 2) ? grids : tbuf2;
    hipLaunchKernelGGL(k_line_fft<-1>, dim3((g.Gy + rpw - 1) / rpw, (unsigned)B * g.Nz), dim3(256),
                       ldsx, s, tbuf, xdst, (long long)g.Gy, (long long)g.Nx * g.Gy,
                       (long long)g.Gx * g.Gy, g.Nx, g.offX + g.Gx / 2, g.Gx, g.Gx / 2, rpw,
                       pg::fft_args(G.px));
    if (g.dims == 3) {
        // FFT along z: lines (b, gx, gy), Nz non-zero inputs at (z + offZ + Gz/2) % Gz
        const int rpz = pick_rpw(g.Gz);
        const size_t ldsz = sizeof(float2) * 2 * rpz * (g.Gz + 1);
        const long long inner = (long long)g.Gx * g.Gy;
        hipLaunchKernelGGL(k_line_fft<-1>, dim3((unsigned)((inner + rpz - 1) / rpz), B), dim3(256),
                           ldsz, s, tbuf2, grids, inner, (long long)g.Nz * inner,
                           (long long)g.Gz * inner, g.Nz, g.offZ + g.Gz / 2, g.Gz, g.Gz / 2, rpz,
                           pg::fft_args(G.pz));
    }
}

void launch_adj_fft_post(const pg::Gnufft& G, float2* grids, float2* tbuf, float2* tbuf2,
                         WeightArgs w, int nc, int L, float2* out, long long out_cstride,
                         hipStream_t s)
{
    const pg::Geometry& g = G.geo;
    const int B = nc * L;
    const float2* xsrc = grids;
    if (g.dims == 3) {
        // IFFT along z on the spread grid (ifftshift on load), keep Nz planes
        const int rpz = pick_rpw(g.Gz);
        const size_t ldsz = sizeof(float2) * 2 * rpz * (g.Gz + 1);
        const long long inner = (long long)g.Gx * g.Gy;
        hipLaunchKernelGGL(k_line_fft<+1>, dim3((unsigned)((inner + rpz - 1) / rpz), B), dim3(256),
                           ldsz, s, grids, tbuf2, inner, (long long)g.Gz * inner,
                           (long long)g.Nz * inner, g.Gz, g.Gz / 2, g.Nz, g.offZ + g.Gz / 2, rpz,
                           pg::fft_args(G.pz));
        xsrc = tbuf2;
    }
    // IFFT along x (ifftshift on load), keep the Nx cropped columns
    const int rpw = pick_rpw(g.Gx);
    const size_t ldsx = sizeof(float2) * 2 * rpw * (g.Gx + 1);
    hipLaunchKernelGGL(k_line_fft<+1>, dim3((g.Gy + rpw - 1) / rpw, (unsigned)B * g.Nz), dim3(256),
                       ldsx, s, xsrc, tbuf, (long long)g.Gy, (long long)g.Gx * g.Gy,
                       (long long)g.Nx * g.Gy, g.Gx, g.Gx / 2, g.Nx, g.offX + g.Gx / 2, rpw,
                       pg::fft_args(G.px));
    const int cpw = pick_cpw(g.Gy, g.Ny);
    const size_t lds1 = sizeof(float2) * 2 * cpw * (g.Gy + 1);
    const int ncol = g.Nx * g.Nz;
    hipLaunchKernelGGL(k_adj_colfft_post, dim3((ncol + cpw - 1) / cpw, nc), dim3(256), lds1, s,
                       tbuf, w.S, w.WoH, G.dfx, G.dfy, G.dfz, G.inv_osd, g.Nx, g.Ny, g.Nz, g.offY, L,
                       cpw, pg::fft_args(G.py), out, out_cstride);
}

}  // namespace pgk
