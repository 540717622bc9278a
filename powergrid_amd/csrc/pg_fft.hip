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
// Power-of-two lengths 16..1024 use the register-resident two-level FFT of
// pg_fftreg.h (one LDS exchange per line); other lengths whose prime factors
// are at most 13 use a mixed radix-4/2/3/5/7/11/13 Stockham FFT staged in LDS.
#include "pg_internal.h"
#include "pg_gnufft.h"
#include "pg_fftreg.h"

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// One radix-R butterfly of a Stockham stage for an odd prime R (5, 7, 11,
// 13): the R inputs in[r * nb] times the stage twiddles w^(r k) (tw:
// exp(-2 pi i m / n), m < n; kt = k n / (Ns R)), then the R-point DFT
//   out[q Ns] = sum_r v_r exp(SIGN 2 pi i q r / R)
// whose roots are the table entries at multiples of n / R, the exponent q r
// reduced mod R as it is stepped.  Scalar loops over runtime R with no
// per-thread arrays; the inputs are re-read from LDS for every output.
template <int SIGN>
__device__ __forceinline__ void odd_dft(const float2* in, int nb, float2* out, int Ns, int R,
                                        const float2* __restrict__ tw, int kt, int n)
{
    const int step = n / R;
    for (int q = 0; q < R; ++q) {
        float2 acc = in[0];
        int e = 0;
        for (int r = 1; r < R; ++r) {
            e += q;
            if (e >= R) e -= R;
            float2 w = tw[r * kt];
            float2 ro = tw[e * step];
            if (SIGN > 0) {
                w.y = -w.y;
                ro.y = -ro.y;
            }
            acc = cadd(acc, cmul(cmul(in[r * nb], w), ro));
        }
        out[q * Ns] = acc;
    }
}

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
            const int base = (j - k) * R + k;
            if (R >= 5) {
                odd_dft<SIGN>(in + j, nb, out + base, Ns, R, f.tw, k * tstride, n);
                continue;
            }
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

// ------------------------------------------------------------------------
// Register-FFT kernels for power-of-two grid lengths (pg_fftreg.h).  Same
// index conventions as the LDS Stockham kernels above.

// Forward prep + FFT_y: one line per R2 threads (a wavefront holds 64/R2
// whole columns, so the exchange is wavefront-local).  Grid: (column groups,
// B); 64*wpb threads per block (wpb <= 4).  Every load is issued
// unconditionally at a clamped in-bounds address and masked afterwards, so
// that all loads of a thread are in flight together.  dap[n] =
// 1/(os^d psi(n)), 0 where psi is NaN (griddingSupport.cpp:262-274).
// HAS: bit 0 = coil maps S, bit 1 = segment phases Wo, bit 2 = no
// deapodization (a plain zero-padded FFT of each image).
// HALF: the zero-padded layout of gridos 2 (Ny = N/2 image rows placed from
// row 3N/4 on, cyclically): entry m of a thread is an image row exactly when
// (m - 3 R1/4) mod R1 < R1/2, whatever the thread, so the other half of the
// loads are dropped at compile time.
template <int R1>
__host__ __device__ constexpr bool half_live(int m)
{
    return ((m - 3 * R1 / 4 + R1) % R1) < R1 / 2;
}

template <int R1, int R2, int HAS, bool HALF>
__global__ void __launch_bounds__(256) k_fwd_col_reg(
    const float2* __restrict__ img, long long img_cstride, const float2* __restrict__ S,
    const float2* __restrict__ Wo, const float* __restrict__ dap, int Nx, int Ny, int Nz,
    int offY, int L, const float2* __restrict__ tw, float2* __restrict__ T)
{
    constexpr int N = R1 * R2;            // Gy
    constexpr int LPW = 64 / R2;          // columns per wavefront
    constexpr int PITCH = R1 * (R2 + 1);  // LDS float2 per column
    __shared__ float2 lds[4 * LPW * PITCH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int t = lane % R2, jl = lane / R2;
    const int b = blockIdx.y;
    const int c = b / L, l = b - c * L;
    const int ncol = Nx * Nz;
    const int col = (blockIdx.x * (blockDim.x >> 6) + wv) * LPW + jl;
    const bool valid = col < ncol;
    const long long Nimg = (long long)Nx * Ny * Nz;
    const int shift = (offY + N / 2) % N;  // image row y sits at (y + shift) % N
    const long long cofs = (long long)(valid ? col : 0) * Ny;
    const float2* im = img + c * img_cstride + cofs;
    const float2* Sc = S + (long long)c * Nimg + cofs;
    const float2* Wl = Wo + (long long)l * Nimg + cofs;
    const float* dp = dap + cofs;

    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) {
        if (HALF && !half_live<R1>(m)) {
            v[m] = make_float2(0.f, 0.f);
            continue;
        }
        const int y = (t + R2 * m - shift) & (N - 1);
        const bool ok = valid && y < Ny;
        const int yc = ok ? y : 0;
        float2 a = im[yc];
        const float d = (HAS & 4) ? 1.f : dp[yc];
        if (HAS & 1) a = pgf::cmul(Sc[yc], a);
        if (HAS & 2) a = pgf::cmul(Wl[yc], a);
        v[m] = ok ? make_float2(a.x * d, a.y * d) : make_float2(0.f, 0.f);
    }
    float2* my = lds + (wv * LPW + jl) * PITCH;
    pgf::line_fft<R1, R2, -1, false>(v, t, my, [](int k1, int q) { return k1 * (R2 + 1) + q; },
                                     tw);
    if (!valid) return;
    float2* dst = T + ((long long)b * ncol + col) * N;
#pragma unroll
    for (int r = 0; r < R1 / R2; ++r)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
            const int k = t + R2 * r + R1 * k2;
            dst[(k + N / 2) & (N - 1)] = v[r * R2 + k2];
        }
}

// FFT along a strided axis (same arguments as k_line_fft).  A block holds
// LPB consecutive inner lines x R2 threads, inner index fastest across lanes
// so that every global access is a run of LPB consecutive elements.  Length
// 256 takes 32 lines (256-byte runs, 512 threads): against 16 lines the x
// FFTs of the 3-D workload's class planes took 1.85 ms instead of 2.13
// (forward) and 2.02 (adjoint); 64 lines (one 1024-thread block per CU)
// slowed the adjoint to 1.97 ms.
template <int R1, int R2>
constexpr int line_lpb() { return R1 * R2 >= 1024 ? 8 : (R1 * R2 == 256 ? 32 : 16); }

template <int R1, int R2, int SIGN, bool HALF>
__global__ void __launch_bounds__((line_lpb<R1, R2>() * R2)) k_line_reg(const float2* __restrict__ src,
                                                  float2* __restrict__ dst, long long inner_n,
                                                  long long src_os, long long dst_os, int nin,
                                                  int in_off, int nout, int out_off,
                                                  const float2* __restrict__ tw)
{
    constexpr int N = R1 * R2;
    constexpr int LPB = line_lpb<R1, R2>();
    constexpr int KST = R2 * LPB + 16;  // LDS stride between k1 rows (padded)
    __shared__ float2 lds[R1 * KST];
    const int j = threadIdx.x % LPB, t = threadIdx.x / LPB;
    const long long outer = blockIdx.y;
    const long long inner = (long long)blockIdx.x * LPB + j;
    const bool valid = inner < inner_n;
    const float2* s = src + outer * src_os + (valid ? inner : 0);
    float2* d = dst + outer * dst_os + inner;

    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) {
        if (HALF && !half_live<R1>(m)) {
            v[m] = make_float2(0.f, 0.f);
            continue;
        }
        const int i = (t + R2 * m - in_off) & (N - 1);  // input entry at position p
        const bool ok = valid && i < nin;
        const float2 a = s[(long long)(ok ? i : 0) * inner_n];
        v[m] = ok ? a : make_float2(0.f, 0.f);
    }
    pgf::line_fft<R1, R2, SIGN, true>(
        v, t, lds, [j](int k1, int q) { return k1 * KST + q * LPB + j; }, tw);
    if (!valid) return;
#pragma unroll
    for (int r = 0; r < R1 / R2; ++r)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
            const int k = t + R2 * r + R1 * k2;
            const int o = (k - out_off) & (N - 1);  // output entry taken from position k
            if (o < nout) d[(long long)o * inner_n] = v[r * R2 + k2];
        }
}

// IFFT along y of the columns of every plane b (ifftshift on load), keeping
// the Ny rows of the crop (fftshift + crop_center_region), unweighted:
//   out[b * out_pstride + col * Ny + y]
// The same line layout as k_fwd_col_reg: one line per R2 threads, LPW lines
// per wavefront, four wavefronts per block.
template <int R1, int R2>
__global__ void __launch_bounds__(256) k_col_ifft_crop(const float2* __restrict__ T, int ncol,
                                                       int Ny, int offY,
                                                       const float2* __restrict__ tw,
                                                       float2* __restrict__ out,
                                                       long long out_pstride)
{
    constexpr int N = R1 * R2;
    constexpr int LPW = 64 / R2;
    constexpr int PITCH = R1 * (R2 + 1);
    __shared__ float2 lds[4 * LPW * PITCH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int t = lane % R2, jl = lane / R2;
    const long long b = blockIdx.y;
    const int col = (blockIdx.x * 4 + wv) * LPW + jl;
    const bool valid = col < ncol;
    const float2* src = T + (b * ncol + (valid ? col : 0)) * N;
    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) v[m] = src[(t + R2 * m + N / 2) & (N - 1)];
    float2* my = lds + (wv * LPW + jl) * PITCH;
    pgf::line_fft<R1, R2, 1, false>(v, t, my, [](int k1, int q) { return k1 * (R2 + 1) + q; }, tw);
    if (!valid) return;
    const int shift = (offY + N / 2) % N;  // image row y sits at (y + shift) % N
    float2* dst = out + b * out_pstride + (long long)col * Ny;
#pragma unroll
    for (int r = 0; r < R1 / R2; ++r)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
            const int y = (t + R2 * r + R1 * k2 - shift) & (N - 1);
            if (y < Ny) dst[y] = v[r * R2 + k2];
        }
}

// z stage of a plan on z-class planes (3-D stacks of 2-D trajectories).
// Forward: for coil c, image column col = x Ny + y (Nxy of them) and segment
// l, with v_z = dxy(col) Wo_l S_c x(z, col) over the Nz image planes
// (SENSE.cpp:59-63, TimeSegmentation.cpp:266-270, the x-y deapodization),
//   H[(k B + c L + l) Nxy + col] = sum_z zm[k][z] v_z
// zm folds the zero padding, shifts, FFT and deapodization along z with
// class k's z taps.  Block: 64 consecutive columns x ng wavefronts; the
// wavefront g takes classes g, g + ng, ... (at most kZsKpt), so every load
// and store of a wavefront is 64 consecutive columns; z in chunks of 8.
constexpr int kZsKpt = 8;

__global__ void __launch_bounds__(1024) k_zstage_fwd(
    const float2* __restrict__ img, long long img_cstride, const float2* __restrict__ S,
    const float2* __restrict__ Wo, const float* __restrict__ dxy, int Nxy, int Nz, int L, int B,
    int nzc, const float2* __restrict__ zm, float2* __restrict__ H)
{
    extern __shared__ float2 szm[];  // nzc x Nz
    for (int e = threadIdx.x; e < nzc * Nz; e += blockDim.x) szm[e] = zm[e];
    __syncthreads();
    const int ng = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.y;
    const int col = blockIdx.x * 64 + lane;
    if (col >= Nxy) return;
    const long long N = (long long)Nxy * Nz;
    const float2* xc = img + c * img_cstride + col;
    const float2* Sc = S ? S + (long long)c * N + col : nullptr;
    const float d = dxy[col];
    for (int l = 0; l < L; ++l) {
        const float2* Wl = Wo ? Wo + (long long)l * N + col : nullptr;
        float2 acc[kZsKpt];
#pragma unroll
        for (int i = 0; i < kZsKpt; ++i) acc[i] = make_float2(0.f, 0.f);
        for (int z0 = 0; z0 < Nz; z0 += 8) {
            float2 v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int z = min(z0 + q, Nz - 1);
                float2 a = xc[(long long)z * Nxy];
                if (Sc) a = pgf::cmul(Sc[(long long)z * Nxy], a);
                if (Wl) a = pgf::cmul(Wl[(long long)z * Nxy], a);
                v[q] = (z0 + q < Nz) ? make_float2(a.x * d, a.y * d) : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int i = 0; i < kZsKpt; ++i) {
                const int k = g + i * ng;
                if (k >= nzc) break;
                const float2* zr = szm + k * Nz;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float2 w = zr[min(z0 + q, Nz - 1)];
                    acc[i].x += w.x * v[q].x - w.y * v[q].y;
                    acc[i].y += w.x * v[q].y + w.y * v[q].x;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kZsKpt; ++i) {
            const int k = g + i * ng;
            if (k >= nzc) break;
            H[((long long)k * B + (long long)c * L + l) * Nxy + col] = acc[i];
        }
    }
}

// The forward z stage for Nz <= NZ (8, 16 or 32): one block of four
// wavefronts per (64 image columns, coil c), coil fastest in the grid, so
// that the blocks of one column chunk run together and each XCD reads the
// chunk's segment phases Wo once.  The block forms S_c x(z) for its columns
// once, in LDS; wavefront g then takes the segments l = g, g + 4, ...:
//   v_z = dxy Wo_l(z) S_c x(z)  (z < Nz, in registers),
//   H[(k B + c L + l) Nxy + col] = sum_z zm[k][z] v_z  for every class k,
// with zm read at uniform addresses (scalar loads) and one 512-byte store
// per (k, l).  The products and the z sum run in the order of k_zstage_fwd.
template <int NZ>
__global__ void __launch_bounds__(256) k_zstage_fwd_reg(
    const float2* __restrict__ img, long long img_cstride, const float2* __restrict__ S,
    const float2* __restrict__ Wo, const float* __restrict__ dxy, int Nxy, int Nz, int L, int B,
    int nc, int nzc, const float2* __restrict__ zm, float2* __restrict__ H)
{
    constexpr int NW = 4;
    __shared__ float2 su[NZ * 64];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.x % nc;
    const int col = (blockIdx.x / nc) * 64 + lane;
    const bool ok = col < Nxy;
    const int colc = ok ? col : Nxy - 1;
    const long long N = (long long)Nxy * Nz;
    for (int z = g; z < Nz; z += NW) {
        const long long o = (long long)z * Nxy + colc;
        float2 a = img[c * img_cstride + o];
        if (S) a = pgf::cmul(S[(long long)c * N + o], a);
        su[z * 64 + lane] = a;
    }
    __syncthreads();
    const float d = dxy[colc];
    for (int l = g; l < L; l += NW) {
        float2 v[NZ];
        if (Wo) {
            float2 w[NZ];
#pragma unroll
            for (int z = 0; z < NZ; ++z)
                w[z] = Wo[(long long)l * N + (long long)min(z, Nz - 1) * Nxy + colc];
#pragma unroll
            for (int z = 0; z < NZ; ++z) v[z] = pgf::cmul(w[z], su[min(z, Nz - 1) * 64 + lane]);
        } else {
#pragma unroll
            for (int z = 0; z < NZ; ++z) v[z] = su[min(z, Nz - 1) * 64 + lane];
        }
#pragma unroll
        for (int z = 0; z < NZ; ++z)
            v[z] = z < Nz ? make_float2(v[z].x * d, v[z].y * d) : make_float2(0.f, 0.f);
        float2* hp = H + ((long long)c * L + l) * Nxy + colc;
        for (int k = 0; k < nzc; ++k) {
            const float2* zr = zm + (long long)k * Nz;
            float ar = 0.f, ai = 0.f;
#pragma unroll
            for (int z = 0; z < NZ; ++z) {
                const float2 w = zr[min(z, Nz - 1)];
                ar += w.x * v[z].x - w.y * v[z].y;
                ai += w.x * v[z].y + w.y * v[z].x;
            }
            if (ok) hp[(long long)k * B * Nxy] = make_float2(ar, ai);
        }
    }
}

// The forward z stage on the matrix cores (Nz <= 32, nzc <= 32): the grid
// and staging of k_zstage_fwd_reg, with the class sums as a real product
//   [Re H; Im H] (64 x 64 columns) = [[Zr, -Zi], [Zi, Zr]] (64 x 64) [vr; vi]
// (the -Zi block applied as -vi) in v_mfma_f32_16x16x4f32 steps (A[i][k] in
// lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16 (i / 4),
// register i % 4).  Lane (p, q) =
// (lane % 16, lane / 16) keeps the class matrix entries of rows 16 rt + p
// and z = 4 s + q in registers for the whole kernel, and per segment and
// 16-column tile forms v_z = dxy Wo_l(z) S_c x(z) for its column 16 ct + p
// and those z.  Same products as k_zstage_fwd_reg; the z sums run in the
// matrix cores' order (four terms per step).
typedef float pgz_f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256, 4) k_zstage_fwd_mfma(
    const float2* __restrict__ img, long long img_cstride, const float2* __restrict__ S,
    const float2* __restrict__ Wo, const float* __restrict__ dxy, int Nxy, int Nz, int L, int B,
    int nc, int nzc, const float2* __restrict__ zm, float2* __restrict__ H)
{
    constexpr int NW = 4;
    __shared__ float2 su[32 * 64];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.x % nc;
    const int col0 = (blockIdx.x / nc) * 64;
    const long long N = (long long)Nxy * Nz;
    {
        const int col = col0 + lane;
        const int colc = col < Nxy ? col : Nxy - 1;
        for (int z = g; z < Nz; z += NW) {
            const long long o = (long long)z * Nxy + colc;
            float2 a = img[c * img_cstride + o];
            if (S) a = pgf::cmul(S[(long long)c * N + o], a);
            su[z * 64 + lane] = a;
        }
    }
    const int p = lane & 15, q = lane >> 4;
    // class matrix entries: row k = 16 rt + p, z = 4 s + q (zero outside)
    float zr[2][8], zi[2][8];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 16 * rt + p, z = 4 * s + q;
            const float2 w = (k < nzc && z < Nz) ? zm[(long long)k * Nz + z] : make_float2(0.f, 0.f);
            zr[rt][s] = w.x;
            zi[rt][s] = w.y;
        }
    __syncthreads();
    // tasks (segment l = g + NW (tau / 4), column tile ct = tau % 4) of this
    // wavefront; the next task's segment phases are requested before this
    // task's products
    const int ntask = ((L - g + NW - 1) / NW) * 4;
    float2 wn[8];
    auto load_wo = [&](int tau, float2(&w)[8]) {
        const int l = g + NW * (tau >> 2);
        const int col = col0 + 16 * (tau & 3) + p;
        const int colc = col < Nxy ? col : Nxy - 1;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int zc = min(4 * s + q, Nz - 1);
            w[s] = Wo ? Wo[(long long)l * N + (long long)zc * Nxy + colc] : make_float2(1.f, 0.f);
        }
    };
    if (ntask > 0) load_wo(0, wn);
#pragma unroll 1
    for (int tau = 0; tau < ntask; ++tau) {
        const int l = g + NW * (tau >> 2), ct = tau & 3;
        float2 wc[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) wc[s] = wn[s];
        if (tau + 1 < ntask) load_wo(tau + 1, wn);
        float2* hp = H + ((long long)c * L + l) * Nxy;
        {
            const int cl = 16 * ct + p;  // column in the block's chunk
            const int col = col0 + cl;
            const int colc = col < Nxy ? col : Nxy - 1;
            const float d = dxy[colc];
            float2 v[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int z = 4 * s + q;
                const int zc = z < Nz ? z : Nz - 1;
                float2 a = su[zc * 64 + cl];
                if (Wo) a = pgf::cmul(wc[s], a);
                v[s] = z < Nz ? make_float2(a.x * d, a.y * d) : make_float2(0.f, 0.f);
            }
            pgz_f4 hr[2], hi[2];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
                hr[rt] = pgz_f4{0.f, 0.f, 0.f, 0.f};
                hi[rt] = pgz_f4{0.f, 0.f, 0.f, 0.f};
            }
            // Re H = Zr vr + Zi (-vi), Im H = Zi vr + Zr vi
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const float b = s < 8 ? v[s].x : v[s - 8].y;
                const float br = s < 8 ? b : -b;
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    const float ar = s < 8 ? zr[rt][s] : zi[rt][s - 8];
                    const float ai = s < 8 ? zi[rt][s] : zr[rt][s - 8];
                    hr[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, hr[rt], 0, 0, 0);
                    hi[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b, hi[rt], 0, 0, 0);
                }
            }
            if (col < Nxy)
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int k = 16 * rt + 4 * q + r;
                        if (k < nzc)
                            hp[(long long)k * B * Nxy + col] = make_float2(hr[rt][r], hi[rt][r]);
                    }
        }
    }
}

// The forward z stage fused with the y FFT of the class planes (gridos 2,
// Ny = Gy / 2 = 64 CT, the zero-padded half layout of k_fwd_col_reg, Nz and
// the class count at most 32).  One block of 256 threads per (image column
// x, coil c), c fastest in the grid (the blocks of one x share its segment
// phases in L2).  Wavefront wv takes the CT 16-row tiles
// y = 16 (wv CT + ct) + p; lane (p, q) = (lane % 16, lane / 16) keeps
// k_zstage_fwd_mfma's class matrix entries in registers, and S_c x(z) and
// dxy of its rows and z = 4 s + q for the whole kernel.  For each segment l:
//   1. the class sums H_k(y) of grid b = c L + l as k_zstage_fwd_mfma forms
//      them (same products, same order), stored in LDS as nzc columns of Ny;
//   2. all four wavefronts transform the class columns along y with
//      k_fwd_col_reg's zero padding, shift and register FFT, and write them
//      to T, plane k B + b, where k_fwd_col_reg writes the class planes' FFTs.
// So the result is bit-identical to the two kernels, without the class
// planes in HBM.  HS, HW: S and Wo are given (compile-time, so that their
// loads are issued together rather than each behind a branch and a wait).
// The next segment's phases are loaded during this segment's
// products.  Static LDS (class columns, exchange rows): 66 KB at Gy = 256,
// two blocks per CU.
template <int R1, int R2, int CT, bool HS, bool HW>
__global__ void __launch_bounds__(256, 2) k_zfwd_yfft_mfma(
    const float2* __restrict__ img, long long img_cstride, const float2* __restrict__ S,
    const float2* __restrict__ Wo, const float* __restrict__ dxy, int Nx, int Ny, int Nz,
    int offY, int L, int B, int nc, int nzc, const float2* __restrict__ zm,
    const float2* __restrict__ tw, float2* __restrict__ T)
{
    constexpr int N = R1 * R2;  // Gy
    constexpr int NY = N / 2;
    constexpr int LPW = 64 / R2;
    constexpr int PITCH = R1 * (R2 + 1);
    constexpr int NRD = (32 + 4 * LPW - 1) / (4 * LPW);
    static_assert(NY == 64 * CT, "Ny = 64 CT");
    __shared__ float2 sh[32 * NY];
    __shared__ float2 ex[4 * LPW * PITCH];
    const int tid = threadIdx.x;
    const int c = blockIdx.x % nc, x = blockIdx.x / nc;
    const int Nxy = Nx * Ny;
    const long long Nimg = (long long)Nxy * Nz;
    const int lane = tid & 63, wv = tid >> 6;
    const int t = lane % R2, jl = lane / R2;
    const int shift = (offY + N / 2) % N;  // image row y sits at (y + shift) % N
    float2* my = ex + (wv * LPW + jl) * PITCH;
    const int p = lane & 15, q = lane >> 4;
    // class matrix entries: row k = 16 rt + p, z = 4 s + q (zero outside)
    float zr[2][8], zi[2][8];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 16 * rt + p, z = 4 * s + q;
            const float2 w = (k < nzc && z < Nz) ? zm[(long long)k * Nz + z] : make_float2(0.f, 0.f);
            zr[rt][s] = w.x;
            zi[rt][s] = w.y;
        }
    // S_c x(z) and dxy of the lane's rows
    float2 u0[CT][8];
    float dd[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const long long o = (long long)x * Ny + 16 * (wv * CT + ct) + p;
        dd[ct] = dxy[o];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int zc = min(4 * s + q, Nz - 1);
            float2 a = img[c * img_cstride + (long long)zc * Nxy + o];
            if (HS) a = pgf::cmul(S[(long long)c * Nimg + (long long)zc * Nxy + o], a);
            u0[ct][s] = a;
        }
    }
    float2 wn[CT][8];
    auto load_wo = [&](int l, float2(&w)[CT][8]) {
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const long long o = (long long)x * Ny + 16 * (wv * CT + ct) + p;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int zc = min(4 * s + q, Nz - 1);
                w[ct][s] = HW ? Wo[(long long)l * Nimg + (long long)zc * Nxy + o]
                              : make_float2(1.f, 0.f);
            }
        }
    };
    load_wo(0, wn);
#pragma unroll 1
    for (int l = 0; l < L; ++l) {
        float2 wc[CT][8];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int s = 0; s < 8; ++s) wc[ct][s] = wn[ct][s];
        if (l + 1 < L) load_wo(l + 1, wn);
        pgz_f4 hr[CT][2], hi[CT][2];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            float2 v[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                float2 a = u0[ct][s];
                if (HW) a = pgf::cmul(wc[ct][s], a);
                v[s] = 4 * s + q < Nz ? make_float2(a.x * dd[ct], a.y * dd[ct])
                                      : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
                hr[ct][rt] = pgz_f4{0.f, 0.f, 0.f, 0.f};
                hi[ct][rt] = pgz_f4{0.f, 0.f, 0.f, 0.f};
            }
            // Re H = Zr vr + Zi (-vi), Im H = Zi vr + Zr vi
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const float b = s < 8 ? v[s].x : v[s - 8].y;
                const float br = s < 8 ? b : -b;
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    const float ar = s < 8 ? zr[rt][s] : zi[rt][s - 8];
                    const float ai = s < 8 ? zi[rt][s] : zr[rt][s - 8];
                    hr[ct][rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, hr[ct][rt], 0, 0, 0);
                    hi[ct][rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b, hi[ct][rt], 0, 0, 0);
                }
            }
        }
        if (l > 0) __syncthreads();  // the previous segment's columns are read
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int y = 16 * (wv * CT + ct) + p;
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = 16 * rt + 4 * q + r;
                    if (k < nzc) sh[k * NY + y] = make_float2(hr[ct][rt][r], hi[ct][rt][r]);
                }
        }
        __syncthreads();
        const int b = c * L + l;
#pragma unroll 1
        for (int rd = 0; rd < NRD; ++rd) {
            const int k = (rd * 4 + wv) * LPW + jl;
            const bool valid = k < nzc;
            const float2* col = sh + (valid ? k : 0) * NY;
            float2 v[R1];
#pragma unroll
            for (int m = 0; m < R1; ++m) {
                if (!half_live<R1>(m)) {
                    v[m] = make_float2(0.f, 0.f);
                    continue;
                }
                const int y = (t + R2 * m - shift) & (N - 1);
                const float2 a = col[y < NY ? y : 0];
                v[m] = valid && y < NY ? a : make_float2(0.f, 0.f);
            }
            pgf::line_fft<R1, R2, -1, false>(
                v, t, my, [](int k1, int qq) { return k1 * (R2 + 1) + qq; }, tw);
            if (!valid) continue;
            float2* dst = T + (((long long)k * B + b) * Nx + x) * N;
#pragma unroll
            for (int r = 0; r < R1 / R2; ++r)
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) {
                    const int kk = t + R2 * r + R1 * k2;
                    dst[(kk + N / 2) & (N - 1)] = v[r * R2 + k2];
                }
        }
    }
}

constexpr int kZfFftWaves = 2;

// Step 1 of the fused adjoint below: the first kZfFftWaves wavefronts
// transform the class columns of grid b along y (ifftshift on load, as
// k_col_ifft_crop; lane = (line jl, thread t of the line's register FFT))
// and keep the Ny cropped rows in sh (nzc x Ny).
template <int R1, int R2, int FW = kZfFftWaves>
__device__ __forceinline__ void zf_class_columns(const float2* __restrict__ T, float2* sh,
                                                 float2* my, int wv, int t, int jl, int nzc,
                                                 int B, int b, int Nx, int x, int Ny, int shift,
                                                 const float2* __restrict__ tw)
{
    constexpr int N = R1 * R2;
    constexpr int LPW = 64 / R2;
    if (wv >= FW) return;
#pragma unroll 1
    for (int k0 = 0; k0 < nzc; k0 += FW * LPW) {
        const int k = k0 + wv * LPW + jl;
        const bool valid = k < nzc;
        const float2* src = T + (((long long)(valid ? k : 0) * B + b) * Nx + x) * N;
        // a copy of t the compiler cannot see through, so that the
        // per-element indices and twiddles are formed in the loop instead
        // of being held across it
        int tt = t;
        asm volatile("" : "+v"(tt));
        float2 v[R1];
#pragma unroll
        for (int m = 0; m < R1; ++m) v[m] = src[(tt + R2 * m + N / 2) & (N - 1)];
        pgf::line_fft<R1, R2, 1, false>(
            v, tt, my, [](int k1, int q) { return k1 * (R2 + 1) + q; }, tw);
        if (!valid) continue;
#pragma unroll
        for (int r = 0; r < R1 / R2; ++r)
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                const int yy = (tt + R2 * r + R1 * k2 - shift) & (N - 1);
                if (yy < Ny) sh[k * Ny + yy] = v[r * R2 + k2];
            }
    }
}

// The same step split in two for a register pipeline over segments: the
// loads of grid b's class columns (NRD rounds of 4 LPW lines, all four
// wavefronts), then their transforms into sh.
template <int R1, int R2, int NRD>
__device__ __forceinline__ void zf_load_columns(const float2* __restrict__ T, int wv, int t,
                                                int jl, int nzc, int B, int b, int Nx, int x,
                                                float2 (&v)[NRD][R1])
{
    constexpr int N = R1 * R2;
    constexpr int LPW = 64 / R2;
#pragma unroll
    for (int rd = 0; rd < NRD; ++rd) {
        const int k = (rd * 4 + wv) * LPW + jl;
        const float2* src = T + (((long long)(k < nzc ? k : 0) * B + b) * Nx + x) * N;
#pragma unroll
        for (int m = 0; m < R1; ++m) v[rd][m] = src[(t + R2 * m + N / 2) & (N - 1)];
    }
}

template <int R1, int R2, int NRD>
__device__ __forceinline__ void zf_transform_columns(float2 (&v)[NRD][R1], float2* sh, float2* my,
                                                     int wv, int t, int jl, int nzc, int Ny,
                                                     int shift, const float2* __restrict__ tw)
{
    constexpr int N = R1 * R2;
    constexpr int LPW = 64 / R2;
#pragma unroll
    for (int rd = 0; rd < NRD; ++rd) {
        const int k = (rd * 4 + wv) * LPW + jl;
        pgf::line_fft<R1, R2, 1, false>(
            v[rd], t, my, [](int k1, int q) { return k1 * (R2 + 1) + q; }, tw);
        if (k < nzc) {
#pragma unroll
            for (int r = 0; r < R1 / R2; ++r)
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) {
                    const int yy = (t + R2 * r + R1 * k2 - shift) & (N - 1);
                    if (yy < Ny) sh[k * Ny + yy] = v[rd][r * R2 + k2];
                }
        }
    }
}

// The adjoint y IFFT of the class planes fused with the adjoint z stage
// (gridos 2, 64 <= Ny = Gy/2 <= 256).  One block of 256 threads per (image column
// x, coil c), c fastest in the grid (the blocks of one x share its WoH in
// L2).  For each segment l:
//   1. the first kZfFftWaves wavefronts transform the class columns of grid
//      b = c L + l along y (ifftshift on load, as k_col_ifft_crop) and keep
//      the Ny cropped rows in LDS;
//   2. thread (group g, row y) forms u_z = sum_k conj(Zm[k][z]) h_k(y) for
//      its ZW image planes z = g ZW, ... and adds WoH_l(z) u_z to its sums.
// Then out = conj(S_c) dxy sum_l (...), as k_zstage_adj, whose sums it
// forms in the same order, without the class planes in HBM.
template <int R1, int R2, int ZW>
__global__ void __launch_bounds__(256) k_yifft_zadj(
    const float2* __restrict__ T, const float2* __restrict__ S, const float2* __restrict__ WoH,
    const float* __restrict__ dxy, int Nx, int Ny, int Nz, int offY, int L, int B, int nc,
    int nzc, const float2* __restrict__ zm, const float2* __restrict__ tw,
    float2* __restrict__ out, long long out_cstride)
{
    constexpr int N = R1 * R2;  // Gy
    constexpr int LPW = 64 / R2;
    constexpr int PITCH = R1 * (R2 + 1);
    extern __shared__ float2 sm[];
    float2* sh = sm;                     // nzc x Ny cropped class columns
    float2* ex = sm + (size_t)nzc * Ny;  // kZfFftWaves * LPW line exchange rows
    const int tid = threadIdx.x;
    const int c = blockIdx.x % nc, x = blockIdx.x / nc;
    const int Nxy = Nx * Ny;
    const long long Nimg = (long long)Nxy * Nz;
    const int lane = tid & 63, wv = tid >> 6;
    const int t = lane % R2, jl = lane / R2;
    const int shift = (offY + N / 2) % N;  // image row y sits at (y + shift) % N
    float2* my = ex + (wv * LPW + jl) * PITCH;
    // Ny >= 64: the rows of a wavefront share their group (uniform z0)
    const int y = tid % Ny, g = __builtin_amdgcn_readfirstlane(tid / Ny);
    const int z0 = g * ZW;
    const long long o = (long long)x * Ny + y;
    float2 acc[ZW];
#pragma unroll
    for (int q = 0; q < ZW; ++q) acc[q] = make_float2(0.f, 0.f);
#pragma unroll 1
    for (int l = 0; l < L; ++l) {
        zf_class_columns<R1, R2>(T, sh, my, wv, t, jl, nzc, B, c * L + l, Nx, x, Ny, shift, tw);
        __syncthreads();
        float2 u[ZW];
#pragma unroll
        for (int q = 0; q < ZW; ++q) u[q] = make_float2(0.f, 0.f);
#pragma unroll 1
        for (int k = 0; k < nzc; ++k) {
            const float2 h = sh[k * Ny + y];
            const float2* zr = zm + (long long)k * Nz;
#pragma unroll
            for (int q = 0; q < ZW; ++q) {
                const float2 w = zr[min(z0 + q, Nz - 1)];
                u[q].x += w.x * h.x + w.y * h.y;  // conj(w) h
                u[q].y += w.x * h.y - w.y * h.x;
            }
        }
        if (WoH) {
            const float2* wl = WoH + (long long)l * Nimg + o;
#pragma unroll
            for (int q = 0; q < ZW; ++q) {
                const int z = min(z0 + q, Nz - 1);
                u[q] = pgf::cmul(u[q], wl[(long long)z * Nxy]);
            }
        }
#pragma unroll
        for (int q = 0; q < ZW; ++q) {
            acc[q].x += u[q].x;
            acc[q].y += u[q].y;
        }
        __syncthreads();  // the next segment's columns replace these
    }
    const float d = dxy[o];
#pragma unroll
    for (int q = 0; q < ZW; ++q) {
        const int z = z0 + q;
        if (z >= Nz) break;
        float2 v = make_float2(acc[q].x * d, acc[q].y * d);
        if (S) {
            float2 sc = S[(long long)c * Nimg + (long long)z * Nxy + o];
            sc.y = -sc.y;
            v = pgf::cmul(v, sc);
        }
        out[c * out_cstride + (long long)z * Nxy + o] = v;
    }
}

// The fused adjoint for Ny = Gy / 2 = 64 CT (CT = 1, 2), Nz and nzc at most
// 32, with all four wavefronts on the y IFFTs and the z stage on the matrix
// cores: after step 1, wavefront wv takes the CT 16-row tiles
// y = 16 (wv CT + ct) + p of the cropped class columns and forms
//   [Re u; Im u] (z x y) = [[Zr^T, Zi^T] [hr; hi], [Zr^T, Zi^T] [hi; -hr]]
// (u_z = sum_k conj(Zm[k][z]) h_k) in v_mfma_f32_16x16x4f32 steps: lane
// (p, q) = (lane % 16, lane / 16) reads Zm[k][z] for z = 16 rt + p,
// k = 4 s + q from an LDS copy of the class matrix, supplies h_k(y) for
// k = 4 s + q, and receives u_z(y) for z = 16 rt + 4 q + r (register r),
// which it multiplies by WoH_l(z) and adds to its sums.  The epilogue is
// k_yifft_zadj's.  Same products; the k sums in the matrix cores' order.
// HS: S is given (compile-time, so that the epilogue's loads of S are issued
// together rather than each behind a branch and a wait).
// Static LDS (class columns, four wavefronts' exchange rows, Zm): 75 KB at
// Gy = 256, two blocks per CU.
template <int R1, int R2, int CT, bool HS>
__global__ void __launch_bounds__(256, 2) k_yifft_zadj_mfma(
    const float2* __restrict__ T, const float2* __restrict__ S, const float2* __restrict__ WoH,
    const float* __restrict__ dxy, int Nx, int Ny, int Nz, int offY, int L, int B, int nc,
    int nzc, const float2* __restrict__ zm, const float2* __restrict__ tw,
    float2* __restrict__ out, long long out_cstride)
{
    constexpr int N = R1 * R2;  // Gy
    constexpr int NY = N / 2;
    constexpr int LPW = 64 / R2;
    constexpr int PITCH = R1 * (R2 + 1);
    static_assert(NY == 64 * CT, "Ny = 64 CT");
    __shared__ float2 sh[32 * NY];
    __shared__ float2 ex[4 * LPW * PITCH];
    __shared__ float2 zs[32 * 32];
    const int tid = threadIdx.x;
    const int c = blockIdx.x % nc, x = blockIdx.x / nc;
    const int Nxy = Nx * Ny;
    const long long Nimg = (long long)Nxy * Nz;
    const int lane = tid & 63, wv = tid >> 6;
    const int t = lane % R2, jl = lane / R2;
    const int shift = (offY + N / 2) % N;
    float2* my = ex + (wv * LPW + jl) * PITCH;
    const int p = lane & 15, q = lane >> 4;
    for (int e = tid; e < nzc * Nz; e += 256) zs[e] = zm[e];
    float2 acc[CT][2][4];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[ct][rt][r] = make_float2(0.f, 0.f);
    // segment l + 1's class columns are loaded while segment l's z stage
    // runs on the matrix cores
    constexpr int NRD = (32 + 4 * LPW - 1) / (4 * LPW);
    float2 vc[NRD][R1];
    zf_load_columns<R1, R2, NRD>(T, wv, t, jl, nzc, B, c * L, Nx, x, vc);
#pragma unroll 1
    for (int l = 0; l < L; ++l) {
        zf_transform_columns<R1, R2, NRD>(vc, sh, my, wv, t, jl, nzc, Ny, shift, tw);
        __syncthreads();
        if (l + 1 < L) zf_load_columns<R1, R2, NRD>(T, wv, t, jl, nzc, B, c * L + l + 1, Nx, x, vc);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int y = 16 * (wv * CT + ct) + p;
            pgz_f4 ur[2], ui[2];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
                ur[rt] = pgz_f4{0.f, 0.f, 0.f, 0.f};
                ui[rt] = pgz_f4{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll 2
            for (int s = 0; s < 8; ++s) {
                const int k = 4 * s + q;
                const float2 h = k < nzc ? sh[k * Ny + y] : make_float2(0.f, 0.f);
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    const int z = 16 * rt + p;
                    const float2 w =
                        (k < nzc && z < Nz) ? zs[k * Nz + z] : make_float2(0.f, 0.f);
                    ur[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.x, h.x, ur[rt], 0, 0, 0);
                    ur[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.y, h.y, ur[rt], 0, 0, 0);
                    ui[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.x, h.y, ui[rt], 0, 0, 0);
                    ui[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.y, -h.x, ui[rt], 0, 0, 0);
                }
            }
            const long long o = (long long)x * Ny + y;
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int z = min(16 * rt + 4 * q + r, Nz - 1);
                    float2 u = make_float2(ur[rt][r], ui[rt][r]);
                    if (WoH) u = pgf::cmul(u, WoH[(long long)l * Nimg + (long long)z * Nxy + o]);
                    acc[ct][rt][r].x += u.x;
                    acc[ct][rt][r].y += u.y;
                }
        }
        __syncthreads();  // the next segment's columns replace these
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int y = 16 * (wv * CT + ct) + p;
        const long long o = (long long)x * Ny + y;
        const float d = dxy[o];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int z = 16 * rt + 4 * q + r;
                if (z >= Nz) continue;
                float2 v = make_float2(acc[ct][rt][r].x * d, acc[ct][rt][r].y * d);
                if (HS) {
                    float2 sc = S[(long long)c * Nimg + (long long)z * Nxy + o];
                    sc.y = -sc.y;
                    v = pgf::cmul(v, sc);
                }
                out[c * out_cstride + (long long)z * Nxy + o] = v;
            }
    }
}

// lines per wavefront and LDS exchange row of the register FFT of length n
int zf_lpw(int n)
{
    switch (n) {
    case 16: return 64 / pgf::Split<16>::R2;
    case 32: return 64 / pgf::Split<32>::R2;
    case 64: return 64 / pgf::Split<64>::R2;
    case 128: return 64 / pgf::Split<128>::R2;
    case 256: return 64 / pgf::Split<256>::R2;
    case 512: return 64 / pgf::Split<512>::R2;
    default: return 64 / pgf::Split<1024>::R2;
    }
}
int zf_pitch(int n)
{
    switch (n) {
    case 16: return pgf::Split<16>::R1 * (pgf::Split<16>::R2 + 1);
    case 32: return pgf::Split<32>::R1 * (pgf::Split<32>::R2 + 1);
    case 64: return pgf::Split<64>::R1 * (pgf::Split<64>::R2 + 1);
    case 128: return pgf::Split<128>::R1 * (pgf::Split<128>::R2 + 1);
    case 256: return pgf::Split<256>::R1 * (pgf::Split<256>::R2 + 1);
    case 512: return pgf::Split<512>::R1 * (pgf::Split<512>::R2 + 1);
    default: return pgf::Split<1024>::R1 * (pgf::Split<1024>::R2 + 1);
    }
}

// Adjoint z stage: for coil c and image column col, over this wavefront's 8
// image planes z,
//   P_c(z, col) = conj(S_c) dxy(col) sum_l WoH_l(z, col) sum_k conj(zm[k][z]) Hp[(k B + c L + l) Nxy + col]
// (TimeSegmentation.cpp:313-323, SENSE.cpp:84-92).  Block: 64 columns x
// ceil(Nz / 8) wavefronts.
__global__ void __launch_bounds__(1024) k_zstage_adj(
    const float2* __restrict__ Hp, const float2* __restrict__ S, const float2* __restrict__ WoH,
    const float* __restrict__ dxy, int Nxy, int Nz, int L, int B, int nzc,
    const float2* __restrict__ zm, float2* __restrict__ out, long long out_cstride)
{
    extern __shared__ float2 szm[];  // nzc x Nz
    for (int e = threadIdx.x; e < nzc * Nz; e += blockDim.x) szm[e] = zm[e];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int z0 = 8 * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.y;
    const int col = blockIdx.x * 64 + lane;
    if (col >= Nxy) return;
    const long long N = (long long)Nxy * Nz;
    float2 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = make_float2(0.f, 0.f);
    for (int l = 0; l < L; ++l) {
        const float2* hl = Hp + ((long long)c * L + l) * Nxy + col;
        float2 u[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) u[q] = make_float2(0.f, 0.f);
        for (int k = 0; k < nzc; ++k) {
            const float2 h = hl[(long long)k * B * Nxy];
            const float2* zr = szm + k * Nz;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float2 w = zr[min(z0 + q, Nz - 1)];
                u[q].x += w.x * h.x + w.y * h.y;  // conj(w) h
                u[q].y += w.x * h.y - w.y * h.x;
            }
        }
        if (WoH) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int z = min(z0 + q, Nz - 1);
                u[q] = pgf::cmul(u[q], WoH[(long long)l * N + (long long)z * Nxy + col]);
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            acc[q].x += u[q].x;
            acc[q].y += u[q].y;
        }
    }
    const float d = dxy[col];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int z = z0 + q;
        if (z >= Nz) break;
        float2 v = make_float2(acc[q].x * d, acc[q].y * d);
        if (S) {
            float2 sc = S[(long long)c * N + (long long)z * Nxy + col];
            sc.y = -sc.y;
            v = pgf::cmul(v, sc);
        }
        out[c * out_cstride + (long long)z * Nxy + col] = v;
    }
}

// Adjoint tail for coil c: IFFT_y of the L segment columns, crop, deapod,
// x WoH_l, sum over l, x conj(S_c).  Block = NW wavefronts over the same
// LPW columns; wavefront w takes segments l = w, w+NW, ...; the NW partial
// sums are added in a fixed order through LDS (deterministic).
template <int R1>
constexpr int post_waves() { return R1 >= 32 ? 4 : 8; }

template <int R1, int R2, int HAS>
__global__ void __launch_bounds__(64 * post_waves<R1>()) k_adj_post_reg(
    const float2* __restrict__ T, const float2* __restrict__ S, const float2* __restrict__ WoH,
    const float* __restrict__ dap, int Nx, int Ny, int Nz, int offY, int L,
    const float2* __restrict__ tw, float2* __restrict__ out, long long out_cstride)
{
    constexpr int N = R1 * R2;  // Gy
    constexpr int LPW = 64 / R2;
    constexpr int PITCH = R1 * (R2 + 1);
    constexpr int NW = post_waves<R1>();
    __shared__ float2 lds[NW * LPW * PITCH];
    // wavefront index made scalar: the segment loop below is uniform
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = lane % R2, jl = lane / R2;
    const int c = blockIdx.y;
    const int ncol = Nx * Nz;
    const int col = blockIdx.x * LPW + jl;
    const bool valid = col < ncol;
    const long long Nimg = (long long)Nx * Ny * Nz;
    const int shift = (offY + N / 2) % N;  // image row y sits at (y + shift) % N
    const long long cofs = (long long)(valid ? col : 0) * Ny;
    const float* dp = dap + cofs;

    // the image rows held by this thread after the transform, and their factors
    int yk[R1];
    float dk[R1];
#pragma unroll
    for (int r = 0; r < R1 / R2; ++r)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
            const int y = (t + R2 * r + R1 * k2 - shift) & (N - 1);
            const bool ok = valid && y < Ny;
            yk[r * R2 + k2] = ok ? y : -1;
            const float d = dp[ok ? y : 0];
            dk[r * R2 + k2] = ok ? d : 0.f;
        }

    float2 acc[R1];
#pragma unroll
    for (int i = 0; i < R1; ++i) acc[i] = make_float2(0.f, 0.f);
    float2* my = lds + (wv * LPW + jl) * PITCH;
    for (int l = wv; l < L; l += NW) {
        const float2* src = T + (((long long)c * L + l) * ncol + (valid ? col : 0)) * N;
        float2 v[R1];
#pragma unroll
        for (int m = 0; m < R1; ++m) v[m] = src[(t + R2 * m + N / 2) & (N - 1)];
        pgf::line_fft<R1, R2, 1, false>(v, t, my, [](int k1, int q) { return k1 * (R2 + 1) + q; },
                                        tw);
        const float2* Wl = WoH + (long long)l * Nimg + cofs;
#pragma unroll
        for (int i = 0; i < R1; ++i) {
            float2 q = make_float2(v[i].x * dk[i], v[i].y * dk[i]);
            if (HAS & 2) q = pgf::cmul(q, Wl[yk[i] < 0 ? 0 : yk[i]]);
            acc[i].x += q.x;
            acc[i].y += q.y;
        }
    }
    // cross-wavefront sum: red[w][jl][y]
    __syncthreads();
    float2* red = lds;
#pragma unroll
    for (int i = 0; i < R1; ++i)
        if (yk[i] >= 0) red[(wv * LPW + jl) * Ny + yk[i]] = acc[i];
    __syncthreads();
    for (int e = threadIdx.x; e < LPW * Ny; e += blockDim.x) {
        const int jj = e / Ny, y = e - jj * Ny;
        const int cc = blockIdx.x * LPW + jj;
        if (cc >= ncol) continue;
        float2 sum = red[jj * Ny + y];
        for (int w = 1; w < NW; ++w) {
            const float2 a = red[(w * LPW + jj) * Ny + y];
            sum.x += a.x;
            sum.y += a.y;
        }
        const long long nidx = (long long)y + (long long)cc * Ny;
        if (HAS & 1) {
            float2 sc = S[(long long)c * Nimg + nidx];
            sc.y = -sc.y;
            sum = pgf::cmul(sum, sc);
        }
        out[c * out_cstride + nidx] = sum;
    }
}

bool reg_len(int n) { return n >= 16 && n <= 1024 && (n & (n - 1)) == 0; }

template <int N>
void fwd_col_reg_n(const pg::Gnufft& G, const float2* img, long long img_cstride,
                   const pgk::WeightArgs& w, int B, int L, float2* tbuf, hipStream_t s)
{
    constexpr int R1 = pgf::Split<N>::R1, R2 = pgf::Split<N>::R2;
    const pg::Geometry& g = G.geo;
    const int ncol = g.Nx * g.Nz;
    const int lpw = 64 / R2;
    const int waves = (ncol + lpw - 1) / lpw;
    const int wpb = (waves * B >= 1024) ? 4 : 1;
    const int has = (w.S ? 1 : 0) | (w.Wo ? 2 : 0);
    const bool half = 2 * g.Ny == N && (g.offY + N / 2) % N == 3 * N / 4;
    const dim3 grid((waves + wpb - 1) / wpb, B), block(64 * wpb);
#define PG_COL(H, HF)                                                                             \
    hipLaunchKernelGGL((k_fwd_col_reg<R1, R2, H, HF>), grid, block, 0, s, img, img_cstride, w.S,   \
                       w.Wo, G.dap, g.Nx, g.Ny, g.Nz, g.offY, L, (const float2*)G.py.tw, tbuf)
    switch (has + (half ? 4 : 0)) {
    case 0: PG_COL(0, false); break;
    case 1: PG_COL(1, false); break;
    case 2: PG_COL(2, false); break;
    case 3: PG_COL(3, false); break;
    case 4: PG_COL(0, true); break;
    case 5: PG_COL(1, true); break;
    case 6: PG_COL(2, true); break;
    default: PG_COL(3, true); break;
    }
#undef PG_COL
}

template <int N, int SIGN>
void line_reg_n(const float2* src, float2* dst, long long inner_n, long long src_os,
                long long dst_os, int nin, int in_off, int nout, int out_off, unsigned outer,
                const pg::FftPlan1D& p, hipStream_t s)
{
    constexpr int R1 = pgf::Split<N>::R1, R2 = pgf::Split<N>::R2;
    constexpr int LPB = line_lpb<R1, R2>();
    const dim3 grid((unsigned)((inner_n + LPB - 1) / LPB), outer), block(LPB * R2);
    if (2 * nin == N && in_off % N == 3 * N / 4)
        hipLaunchKernelGGL((k_line_reg<R1, R2, SIGN, true>), grid, block, 0, s, src, dst, inner_n,
                           src_os, dst_os, nin, in_off % N, nout, out_off % N,
                           (const float2*)p.tw);
    else
        hipLaunchKernelGGL((k_line_reg<R1, R2, SIGN, false>), grid, block, 0, s, src, dst, inner_n,
                           src_os, dst_os, nin, in_off % N, nout, out_off % N,
                           (const float2*)p.tw);
}

template <int N>
void adj_post_reg_n(const pg::Gnufft& G, const float2* tbuf, const pgk::WeightArgs& w, int nc,
                    int L, float2* out, long long out_cstride, hipStream_t s)
{
    constexpr int R1 = pgf::Split<N>::R1, R2 = pgf::Split<N>::R2;
    const pg::Geometry& g = G.geo;
    const int ncol = g.Nx * g.Nz;
    const int lpw = 64 / R2;
    const int has = (w.S ? 1 : 0) | (w.WoH ? 2 : 0);
    const dim3 grid((ncol + lpw - 1) / lpw, nc), block(64 * post_waves<R1>());
#define PG_POST(H)                                                                               \
    hipLaunchKernelGGL((k_adj_post_reg<R1, R2, H>), grid, block, 0, s, tbuf, w.S, w.WoH, G.dap,   \
                       g.Nx, g.Ny, g.Nz, g.offY, L, (const float2*)G.py.tw, out, out_cstride)
    switch (has) {
    case 0: PG_POST(0); break;
    case 1: PG_POST(1); break;
    case 2: PG_POST(2); break;
    default: PG_POST(3); break;
    }
#undef PG_POST
}

#define PG_REG_SWITCH(n, CALL) \
    switch (n) {               \
    case 16: CALL(16); break;  \
    case 32: CALL(32); break;  \
    case 64: CALL(64); break;  \
    case 128: CALL(128); break; \
    case 256: CALL(256); break; \
    case 512: CALL(512); break; \
    default: CALL(1024); break; \
    }

// strided line FFT: register kernel for power-of-two lengths, Stockham otherwise
template <int SIGN>
void line_fft_any(const float2* src, float2* dst, long long inner_n, long long src_os,
                  long long dst_os, int nin, int in_off, int nout, int out_off, unsigned outer,
                  const pg::FftPlan1D& p, hipStream_t s)
{
    if (reg_len(p.n)) {
#define PG_CALL_LINE(NN) \
    line_reg_n<NN, SIGN>(src, dst, inner_n, src_os, dst_os, nin, in_off, nout, out_off, outer, p, s)
        PG_REG_SWITCH(p.n, PG_CALL_LINE)
#undef PG_CALL_LINE
        return;
    }
    const int rpw = pick_rpw(p.n);
    const size_t lds = sizeof(float2) * 2 * rpw * (p.n + 1);
    hipLaunchKernelGGL(k_line_fft<SIGN>, dim3((unsigned)((inner_n + rpw - 1) / rpw), outer),
                       dim3(256), lds, s, src, dst, inner_n, src_os, dst_os, nin, in_off, nout,
                       out_off, rpw, pg::fft_args(p));
}

// the y FFT of Bp zero-padded Nx x Ny planes (image layout, y fastest),
// unweighted: planes[b][x][y] -> T[b][x][.] (k_fwd_col_reg without maps,
// phases or deapodization)
template <int N>
void fwd_col_planes_n(const pg::Gnufft& G, const float2* planes, int Bp, float2* T, hipStream_t s)
{
    constexpr int R1 = pgf::Split<N>::R1, R2 = pgf::Split<N>::R2;
    const pg::Geometry& g = G.geo;
    const int lpw = 64 / R2;
    const int waves = (g.Nx + lpw - 1) / lpw;
    const int wpb = (waves * (long long)Bp >= 1024) ? 4 : 1;
    const bool half = 2 * g.Ny == N && (g.offY + N / 2) % N == 3 * N / 4;
    const dim3 grid((waves + wpb - 1) / wpb, Bp), block(64 * wpb);
    const long long nxy = (long long)g.Nx * g.Ny;
    if (half)
        hipLaunchKernelGGL((k_fwd_col_reg<R1, R2, 4, true>), grid, block, 0, s, planes, nxy,
                           (const float2*)nullptr, (const float2*)nullptr, (const float*)nullptr,
                           g.Nx, g.Ny, 1, g.offY, 1, (const float2*)G.py.tw, T);
    else
        hipLaunchKernelGGL((k_fwd_col_reg<R1, R2, 4, false>), grid, block, 0, s, planes, nxy,
                           (const float2*)nullptr, (const float2*)nullptr, (const float*)nullptr,
                           g.Nx, g.Ny, 1, g.offY, 1, (const float2*)G.py.tw, T);
}

template <int N>
void col_ifft_crop_n(const pg::Gnufft& G, const float2* T, int Bp, float2* planes, hipStream_t s)
{
    constexpr int R1 = pgf::Split<N>::R1, R2 = pgf::Split<N>::R2;
    const pg::Geometry& g = G.geo;
    const int lpw = 64 / R2;
    const dim3 grid((g.Nx + 4 * lpw - 1) / (4 * lpw), Bp), block(256);
    hipLaunchKernelGGL((k_col_ifft_crop<R1, R2>), grid, block, 0, s, T, g.Nx, g.Ny, g.offY,
                       (const float2*)G.py.tw, planes, (long long)g.Nx * g.Ny);
}

}  // namespace

namespace pgk {

void launch_fwd_prep_fft(const pg::Gnufft& G, const float2* img, long long img_cstride,
                         WeightArgs w, int nc, int L, float2* tbuf, float2* tbuf2, float2* grids,
                         hipStream_t s, bool planes)
{
    const pg::Geometry& g = G.geo;
    const int B = nc * L;
    const int ncol = g.Nx * g.Nz;
    if (planes) {
        // z-class planes: the z stage (maps, phases, deapodization, z FFT and
        // z taps) into B*nzc image-sized planes (tbuf2), their y FFT (tbuf)
        // and x FFT (grids, plane (k*B + b))
        PG_REQUIRE(zclass_active(G, B) && reg_len(g.Gy), PG_ERR_STATE,
                   "no z-class planes for this batch");
        const int Bp = B * G.nzc;
        const int nxy = g.Nx * g.Ny;
        const unsigned nch = (unsigned)((nxy + 63) / 64);
        const int N = g.Gy;
        const bool half = 2 * g.Ny == N && (g.offY + N / 2) % N == 3 * N / 4;
        const bool mfma = g_zmfma.load(std::memory_order_relaxed) && g.Nz <= 32 && G.nzc <= 32;
        if (mfma && g_zfwd.load(std::memory_order_relaxed) && half &&
            (g.Ny == 64 || g.Ny == 128)) {
            // the z stage and the y FFT in one kernel, straight into tbuf
            const unsigned nb = (unsigned)((long long)nc * g.Nx);
#define PG_ZFY(NN, CT, HS, HW)                                                                \
    hipLaunchKernelGGL((k_zfwd_yfft_mfma<pgf::Split<NN>::R1, pgf::Split<NN>::R2, CT, HS, HW>), \
                       dim3(nb), dim3(256), 0, s, img, img_cstride, w.S, w.Wo, G.dapxy, g.Nx,     \
                       g.Ny, g.Nz, g.offY, L, B, nc, G.nzc, G.zmat, (const float2*)G.py.tw, tbuf)
#define PG_ZFY_N(NN, CT)                                                                      \
    do {                                                                                       \
        if (w.S && w.Wo)                                                                       \
            PG_ZFY(NN, CT, true, true);                                                        \
        else if (w.S)                                                                          \
            PG_ZFY(NN, CT, true, false);                                                       \
        else if (w.Wo)                                                                         \
            PG_ZFY(NN, CT, false, true);                                                       \
        else                                                                                   \
            PG_ZFY(NN, CT, false, false);                                                      \
    } while (0)
            if (g.Ny == 64)
                PG_ZFY_N(128, 1);
            else
                PG_ZFY_N(256, 2);
#undef PG_ZFY_N
#undef PG_ZFY
            line_fft_any<-1>(tbuf, grids, (long long)g.Gy, (long long)g.Nx * g.Gy,
                             (long long)g.Gx * g.Gy, g.Nx, (g.offX + g.Gx / 2) % g.Gx, g.Gx,
                             g.Gx / 2, (unsigned)Bp, G.px, s);
            return;
        }
#define PG_ZSF(NZ)                                                                              \
    hipLaunchKernelGGL(k_zstage_fwd_reg<NZ>, dim3(nch * nc), dim3(256), 0, s, img, img_cstride, \
                       w.S, w.Wo, G.dapxy, nxy, g.Nz, L, B, nc, G.nzc, G.zmat, tbuf2)
        if (mfma) {
            hipLaunchKernelGGL(k_zstage_fwd_mfma, dim3(nch * nc), dim3(256), 0, s, img,
                               img_cstride, w.S, w.Wo, G.dapxy, nxy, g.Nz, L, B, nc, G.nzc,
                               G.zmat, tbuf2);
        } else if (g.Nz <= 8) {
            PG_ZSF(8);
        } else if (g.Nz <= 16) {
            PG_ZSF(16);
        } else if (g.Nz <= 32) {
            PG_ZSF(32);
        } else {
            const int ng = (G.nzc + kZsKpt - 1) / kZsKpt;
            const size_t lds = sizeof(float2) * (size_t)G.nzc * g.Nz;
            hipLaunchKernelGGL(k_zstage_fwd, dim3(nch, nc), dim3(64 * ng), lds, s, img,
                               img_cstride, w.S, w.Wo, G.dapxy, nxy, g.Nz, L, B, G.nzc, G.zmat,
                               tbuf2);
        }
#undef PG_ZSF
#define PG_CALL_COLP(NN) fwd_col_planes_n<NN>(G, tbuf2, Bp, tbuf, s)
        PG_REG_SWITCH(g.Gy, PG_CALL_COLP)
#undef PG_CALL_COLP
        line_fft_any<-1>(tbuf, grids, (long long)g.Gy, (long long)g.Nx * g.Gy,
                         (long long)g.Gx * g.Gy, g.Nx, (g.offX + g.Gx / 2) % g.Gx, g.Gx, g.Gx / 2,
                         (unsigned)Bp, G.px, s);
        return;
    }
    if (reg_len(g.Gy)) {
#define PG_CALL_COL(NN) fwd_col_reg_n<NN>(G, img, img_cstride, w, B, L, tbuf, s)
        PG_REG_SWITCH(g.Gy, PG_CALL_COL)
#undef PG_CALL_COL
    } else {
        const int cpw = pick_cpw(g.Gy, g.Ny);
        const size_t lds1 = sizeof(float2) * 2 * cpw * (g.Gy + 1);
        hipLaunchKernelGGL(k_fwd_prep_colfft, dim3((ncol + cpw - 1) / cpw, B), dim3(256), lds1, s,
                           img, img_cstride, w.S, w.Wo, G.dfx, G.dfy, G.dfz, G.inv_osd, g.Nx, g.Ny,
                           g.Nz, g.offY, L, cpw, pg::fft_args(G.py), tbuf);
    }
    // FFT along x: rows (b, z, gy), Nx non-zero inputs at (x + offX + Gx/2) % Gx
    float2* xdst = (g.dims == 2) ? grids : tbuf2;
    line_fft_any<-1>(tbuf, xdst, (long long)g.Gy, (long long)g.Nx * g.Gy, (long long)g.Gx * g.Gy,
                     g.Nx, (g.offX + g.Gx / 2) % g.Gx, g.Gx, g.Gx / 2, (unsigned)B * g.Nz, G.px, s);
    if (g.dims == 3) {
        // FFT along z: lines (b, gx, gy), Nz non-zero inputs at (z + offZ + Gz/2) % Gz
        const long long inner = (long long)g.Gx * g.Gy;
        line_fft_any<-1>(tbuf2, grids, inner, (long long)g.Nz * inner, (long long)g.Gz * inner,
                         g.Nz, (g.offZ + g.Gz / 2) % g.Gz, g.Gz, g.Gz / 2, (unsigned)B, G.pz, s);
    }
}

void launch_adj_fft_post(const pg::Gnufft& G, float2* grids, float2* tbuf, float2* tbuf2,
                         WeightArgs w, int nc, int L, float2* out, long long out_cstride,
                         hipStream_t s, bool planes)
{
    const pg::Geometry& g = G.geo;
    const int B = nc * L;
    const float2* xsrc = grids;
    if (planes) {
        // z-class planes: x IFFT of every plane keeping the Nx cropped
        // columns (tbuf), y IFFT and crop (tbuf2: B*nzc image-sized planes),
        // then the z stage into the coils' partial images
        PG_REQUIRE(zclass_active(G, B) && reg_len(g.Gy), PG_ERR_STATE,
                   "no z-class planes for this batch");
        const int Bp = B * G.nzc;
        const int nxy = g.Nx * g.Ny;
        line_fft_any<+1>(grids, tbuf, (long long)g.Gy, (long long)g.Gx * g.Gy,
                         (long long)g.Nx * g.Gy, g.Gx, g.Gx / 2, g.Nx, (g.offX + g.Gx / 2) % g.Gx,
                         (unsigned)Bp, G.px, s);
        const int N = g.Gy;
        const bool half = 2 * g.Ny == N && (g.offY + N / 2) % N == 3 * N / 4;
        const size_t fused_lds =
            sizeof(float2) * ((size_t)G.nzc * g.Ny + (size_t)kZfFftWaves * zf_lpw(N) * zf_pitch(N));
        const int zgroups = 256 / std::max(g.Ny, 1);
        if (g_zfuse.load(std::memory_order_relaxed) && g_zamfma.load(std::memory_order_relaxed) &&
            half && (g.Ny == 64 || g.Ny == 128) && g.Nz <= 32 && G.nzc <= 32) {
            // y IFFT on four wavefronts, crop and the z stage (matrix cores)
            const unsigned nb = (unsigned)((long long)nc * g.Nx);
#define PG_YZM(NN, CT, HS)                                                                      \
    hipLaunchKernelGGL((k_yifft_zadj_mfma<pgf::Split<NN>::R1, pgf::Split<NN>::R2, CT, HS>),      \
                       dim3(nb), dim3(256), 0, s, tbuf, w.S, w.WoH, G.dapxy, g.Nx, g.Ny, g.Nz,      \
                       g.offY, L, B, nc, G.nzc, G.zmat, (const float2*)G.py.tw, out, out_cstride)
            if (g.Ny == 64 && w.S)
                PG_YZM(128, 1, true);
            else if (g.Ny == 64)
                PG_YZM(128, 1, false);
            else if (w.S)
                PG_YZM(256, 2, true);
            else
                PG_YZM(256, 2, false);
#undef PG_YZM
            return;
        }
        if (g_zfuse.load(std::memory_order_relaxed) && half && g.Ny >= 64 && g.Ny <= 256 &&
            g.Nz <= 32 * zgroups && fused_lds <= 64 * 1024) {
            // y IFFT, crop and the z stage in one kernel
            const int zw = (g.Nz + zgroups - 1) / zgroups;
            const unsigned nb = (unsigned)((long long)nc * g.Nx);
#define PG_YZA(NN, ZW)                                                                          \
    hipLaunchKernelGGL((k_yifft_zadj<pgf::Split<NN>::R1, pgf::Split<NN>::R2, ZW>), dim3(nb),      \
                       dim3(256), fused_lds, s, tbuf, w.S, w.WoH, G.dapxy, g.Nx, g.Ny, g.Nz,        \
                       g.offY, L, B, nc, G.nzc, G.zmat, (const float2*)G.py.tw, out, out_cstride)
#define PG_YZA_N(NN)                                                                            \
    do {                                                                                         \
        if (zw <= 8)                                                                             \
            PG_YZA(NN, 8);                                                                       \
        else if (zw <= 16)                                                                       \
            PG_YZA(NN, 16);                                                                      \
        else                                                                                     \
            PG_YZA(NN, 32);                                                                      \
    } while (0)
            PG_REG_SWITCH(g.Gy, PG_YZA_N)
#undef PG_YZA_N
#undef PG_YZA
            return;
        }
#define PG_CALL_CROP(NN) col_ifft_crop_n<NN>(G, tbuf, Bp, tbuf2, s)
        PG_REG_SWITCH(g.Gy, PG_CALL_CROP)
#undef PG_CALL_CROP
        const int zg = (g.Nz + 7) / 8;
        const size_t lds = sizeof(float2) * (size_t)G.nzc * g.Nz;
        hipLaunchKernelGGL(k_zstage_adj, dim3((nxy + 63) / 64, nc), dim3(64 * zg), lds, s, tbuf2,
                           w.S, w.WoH, G.dapxy, nxy, g.Nz, L, B, G.nzc, G.zmat, out, out_cstride);
        return;
    }
    if (g.dims == 3) {
        // IFFT along z on the spread grid (ifftshift on load), keep Nz planes
        const long long inner = (long long)g.Gx * g.Gy;
        line_fft_any<+1>(grids, tbuf2, inner, (long long)g.Gz * inner, (long long)g.Nz * inner,
                         g.Gz, g.Gz / 2, g.Nz, (g.offZ + g.Gz / 2) % g.Gz, (unsigned)B, G.pz, s);
        xsrc = tbuf2;
    }
    // IFFT along x (ifftshift on load), keep the Nx cropped columns
    line_fft_any<+1>(xsrc, tbuf, (long long)g.Gy, (long long)g.Gx * g.Gy, (long long)g.Nx * g.Gy,
                     g.Gx, g.Gx / 2, g.Nx, (g.offX + g.Gx / 2) % g.Gx, (unsigned)B * g.Nz, G.px, s);
    if (reg_len(g.Gy)) {
#define PG_CALL_POST(NN) adj_post_reg_n<NN>(G, tbuf, w, nc, L, out, out_cstride, s)
        PG_REG_SWITCH(g.Gy, PG_CALL_POST)
#undef PG_CALL_POST
        return;
    }
    const int cpw = pick_cpw(g.Gy, g.Ny);
    const size_t lds1 = sizeof(float2) * 2 * cpw * (g.Gy + 1);
    const int ncol = g.Nx * g.Nz;
    hipLaunchKernelGGL(k_adj_colfft_post, dim3((ncol + cpw - 1) / cpw, nc), dim3(256), lds1, s,
                       tbuf, w.S, w.WoH, G.dfx, G.dfy, G.dfz, G.inv_osd, g.Nx, g.Ny, g.Nz, g.offY, L,
                       cpw, pg::fft_args(G.py), out, out_cstride);
}

}  // namespace pgk
