// pg_fftreg.h -- register-resident FFT building blocks for gfx950.
//
// A line of length N = R1*R2 (both powers of two, R1 = R2 or R1 = 2*R2) is
// transformed by R2 threads:
//   1. thread t holds v[m] = x[t + R2*m], m < R1, and does a length-R1 DFT in
//      registers (compile-time twiddles, trivial ones folded away);
//   2. multiplies entry k1 by w_N^(t*k1) (FP64-computed table, stored FP32);
//   3. one exchange through LDS (padded rows: conflict-free both ways);
//   4. thread q does R1/R2 length-R2 DFTs over t, for k1 = q + R2*r.
// On return thread q holds v[r*R2 + k2] = X[(q + R2*r) + R1*k2].
// This is the Cooley-Tukey split n = t + R2 m, k = k1 + R1 k2 of the DFT
// X[k] = sum_n x[n] exp(SIGN 2 pi i n k / N), unnormalised (fftCPU.cpp:48-106
// computes the same sums with FFTW).
#pragma once

#include <hip/hip_runtime.h>

namespace pgf {

constexpr double kPi = 3.14159265358979323846264338327950288;

// sin / cos of 2*pi*k/n in double, evaluated at compile time (Taylor series
// after reduction to [-pi/4, pi/4] by octant symmetry).
constexpr double taylor_sin(double x)
{
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x)
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
// returns {cos, sin} of 2 pi k / n
struct CS {
    double c, s;
};
constexpr CS cexp_turn(long k, long n)
{
    // angle = 2 pi k / n; octant o = round(8 k / n), remainder r in [-1/16, 1/16] turns
    k %= n;
    if (k < 0) k += n;
    const long o = (8 * k + n / 2) / n;  // nearest octant, 0..8
    const double r = 2.0 * kPi * ((double)(8 * k - o * n) / (8.0 * (double)n));
    const double sr = taylor_sin(r), cr = taylor_cos(r);
    const double h = 0.70710678118654752440084436210484904;
    switch (o & 7) {
        case 0: return {cr, sr};
        case 1: return {h * (cr - sr), h * (cr + sr)};
        case 2: return {-sr, cr};
        case 3: return {-h * (cr + sr), h * (cr - sr)};
        case 4: return {-cr, -sr};
        case 5: return {-h * (cr - sr), -h * (cr + sr)};
        case 6: return {sr, -cr};
        default: return {h * (cr + sr), -h * (cr - sr)};
    }
}

// exp(-2 pi i k / R) for k < R
template <int R>
struct Tw {
    float c[R];
    float s[R];
    constexpr Tw() : c(), s()
    {
        for (int k = 0; k < R; ++k) {
            const CS v = cexp_turn(k, R);
            c[k] = (float)v.c;
            s[k] = (float)(-v.s);
        }
    }
};

template <int R>
constexpr int bitrev(int i)
{
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// b * exp(SIGN 2 pi i k / R)
template <int SIGN, int R>
__host__ __device__ __forceinline__ float2 twid(float2 b, int k, const Tw<R>& tw)
{
    if (k == 0) return b;
    if (2 * k == R) return make_float2(-b.x, -b.y);
    if (4 * k == R) return SIGN < 0 ? make_float2(b.y, -b.x) : make_float2(-b.y, b.x);
    if (4 * k == 3 * R) return SIGN < 0 ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
    const float c = tw.c[k];
    const float s = SIGN < 0 ? tw.s[k] : -tw.s[k];
    return make_float2(b.x * c - b.y * s, b.x * s + b.y * c);
}

// In-register DFT of length R (power of two), natural order in and out:
// radix-2 decimation in time after a compile-time bit reversal.
template <int R, int SIGN>
__host__ __device__ __forceinline__ void reg_dft(float2 (&v)[R])
{
    constexpr Tw<R> tw{};
    float2 u[R];
#pragma unroll
    for (int i = 0; i < R; ++i) u[i] = v[bitrev<R>(i)];
#pragma unroll
    for (int half = 1; half < R; half *= 2) {
#pragma unroll
        for (int i = 0; i < R; i += 2 * half) {
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const float2 a = u[i + j];
                const float2 b = twid<SIGN, R>(u[i + j + half], j * (R / (2 * half)), tw);
                u[i + j] = make_float2(a.x + b.x, a.y + b.y);
                u[i + j + half] = make_float2(a.x - b.x, a.y - b.y);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = u[i];
}

// Split of a supported power-of-two length into R1 * R2.
template <int N>
struct Split;
template <> struct Split<16> { static constexpr int R1 = 4, R2 = 4; };
template <> struct Split<32> { static constexpr int R1 = 8, R2 = 4; };
template <> struct Split<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct Split<128> { static constexpr int R1 = 16, R2 = 8; };
template <> struct Split<256> { static constexpr int R1 = 16, R2 = 16; };
template <> struct Split<512> { static constexpr int R1 = 32, R2 = 16; };
template <> struct Split<1024> { static constexpr int R1 = 32, R2 = 32; };

// Ordering of the LDS exchange between the threads of one line.
//   wave_sync: the threads of a line are in one wavefront.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Steps 1-4 above for one line.  `at(k1, t)` gives the LDS slot of entry k1
// of thread t (a permutation chosen by the caller for conflict-free access);
// BLOCK_SYNC selects a block barrier (lines spread over wavefronts) or a
// wavefront-level one.  tw: exp(-2 pi i k / N), k < N.
template <int R1, int R2, int SIGN, bool BLOCK_SYNC, class At>
__device__ __forceinline__ void line_fft(float2 (&v)[R1], int t, float2* lds, At at,
                                         const float2* __restrict__ tw)
{
    constexpr int N = R1 * R2;
    static_assert(R1 == R2 || R1 == 2 * R2, "unsupported split");
    reg_dft<R1, SIGN>(v);
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) {
        float2 w = tw[(t * k1) & (N - 1)];
        if (SIGN > 0) w.y = -w.y;
        v[k1] = cmul(v[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) lds[at(k1, t)] = v[k1];
    if (BLOCK_SYNC)
        __syncthreads();
    else
        wave_lds_sync();
#pragma unroll
    for (int r = 0; r < R1 / R2; ++r) {
        float2 u[R2];
        const int k1 = t + R2 * r;
#pragma unroll
        for (int q = 0; q < R2; ++q) u[q] = lds[at(k1, q)];
        reg_dft<R2, SIGN>(u);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[r * R2 + k2] = u[k2];
    }
    // the caller may reuse the LDS rows right away
    if (BLOCK_SYNC)
        __syncthreads();
    else
        wave_lds_sync();
}

}  // namespace pgf
