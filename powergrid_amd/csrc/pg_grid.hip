// pg_grid.hip -- Kaiser-Bessel interpolation (forward gather) and spreading
// (adjoint scatter) on the oversampled grid, batched over coils x segments.
//
// Weights follow the reference exactly (gridding.cpp:64-97):
//   s  = os * (k + N/2)
//   taps n in [max(0, ceil(s - W*os/2)), min(os*N - 1, floor(s + W*os/2))]
//        (clamped, no wrap-around)
//   w  = LUT[(int)(d*d*4/W^2 * S)], d = |s - n| / os, 0 if the index >= S
// A window holds ceil(W*os) taps; the extra tap that appears when s - W*os/2
// is an integer sits at distance exactly W/2 and has weight 0, so it is not
// visited.
//
// Spreading is output-stationary: each block owns a tile of grid points
// (16x16 in 2-D, 8x8x4 in 3-D) for a chunk of BC grids, walks the samples
// whose window meets the tile (built once per trajectory on the host, in
// sample order), and accumulates in registers.  No atomics, no zero-fill of
// the grid, and the per-point summation order is the reference's serial
// sample order.
#include "pg_gnufft.h"
#include "pg_internal.h"

namespace {

struct Win {
    float s;
    int lo, hi;
};

__device__ __forceinline__ Win window(float k, int N, float os, float halfw)
{
    Win w;
    w.s = os * (k + (float)N / 2.0f);
    w.lo = (int)ceilf(w.s - halfw);  // unclamped start; taps < 0 are skipped
    w.hi = (int)fminf(os * (float)N - 1.0f, floorf(w.s + halfw));
    return w;
}

__device__ __forceinline__ float lut_w(float s, int n, float os, float W2, const float* lut, int S)
{
    float d = fabsf(s - (float)n) / os;
    int k0 = (int)((d * d * 4.0f / W2) * (float)S);
    return (k0 >= S) ? 0.0f : lut[k0];
}

template <int NT>
__device__ __forceinline__ void taps(const Win& w, float os, float W2, const float* lut, int S,
                                     float (&wt)[NT])
{
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int n = w.lo + j;
        wt[j] = (n >= 0 && n <= w.hi) ? lut_w(w.s, n, os, W2, lut, S) : 0.0f;
    }
}

// Forward interpolation, one thread per (sample i, coil c); sums the L
// segments with the AA weights (TimeSegmentation.cpp:279-295).
template <int NT, bool D3>
__global__ void __launch_bounds__(256) k_interp(
    const float* __restrict__ kx, const float* __restrict__ ky, const float* __restrict__ kz,
    int M, int Nx, int Ny, int Nz, int Gx, int Gy, float os, float halfw, float W2,
    const float* __restrict__ lut, int S, const float2* __restrict__ grids, long long G, int L,
    const float* __restrict__ AA, float2* __restrict__ out, long long out_cstride)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= M) return;
    const Win wx_ = window(kx[i], Nx, os, halfw);
    const Win wy_ = window(ky[i], Ny, os, halfw);
    float wx[NT], wy[NT], wz[D3 ? NT : 1];
    taps<NT>(wx_, os, W2, lut, S, wx);
    taps<NT>(wy_, os, W2, lut, S, wy);
    Win wz_{0.f, 0, 0};
    if (D3) {
        wz_ = window(kz[i], Nz, os, halfw);
        float t[NT];
        taps<NT>(wz_, os, W2, lut, S, t);
#pragma unroll
        for (int j = 0; j < NT; ++j) wz[j % (D3 ? NT : 1)] = t[j];
    }
    float2 tot = make_float2(0.f, 0.f);
    for (int l = 0; l < L; ++l) {
        const float2* g = grids + ((long long)c * L + l) * G;
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int jz = 0; jz < (D3 ? NT : 1); ++jz) {
            const int nz = D3 ? wz_.lo + jz : 0;
            if (D3 && (nz < 0 || nz > wz_.hi)) continue;
            const float2* gz = g + (long long)nz * Gx * Gy;
#pragma unroll
            for (int jx = 0; jx < NT; ++jx) {
                const int nx = wx_.lo + jx;
                if (nx < 0 || nx > wx_.hi) continue;
                const float2* row = gz + (long long)nx * Gy;
#pragma unroll
                for (int jy = 0; jy < NT; ++jy) {
                    const int ny = wy_.lo + jy;
                    if (ny < 0 || ny > wy_.hi) continue;
                    float w = D3 ? wx[jx] * wy[jy] * wz[jz] : wx[jx] * wy[jy];
                    const float2 v = row[ny];
                    ar += w * v.x;
                    ai += w * v.y;
                }
            }
        }
        if (AA) {
            const float a = AA[(long long)l * M + i];
            ar = a * ar;
            ai = a * ai;
        }
        tot.x += ar;
        tot.y += ai;
    }
    out[c * out_cstride + i] = tot;
}

constexpr int kChunk = 64;  // samples staged in LDS per pass

template <int NT, bool D3, int BC>
__global__ void __launch_bounds__(256) k_spread(
    const float* __restrict__ kx, const float* __restrict__ ky, const float* __restrict__ kz,
    int M, int Nx, int Ny, int Nz, int Gx, int Gy, int Gz, float os, float halfw, float W2,
    const float* __restrict__ lut, int S, const int* __restrict__ tile_off,
    const int* __restrict__ tile_idx, int ntx, int nty, const float2* __restrict__ in,
    long long in_cstride, const float2* __restrict__ resid_y, const float* __restrict__ Wd,
    int L, const float* __restrict__ AA, int B, float2* __restrict__ grids, long long G)
{
    __shared__ int s_lo[kChunk][3];
    __shared__ float s_w[kChunk][3][NT];
    __shared__ float2 s_v[kChunk][BC];

    const int tile = blockIdx.x;
    const int b0 = blockIdx.y * BC;
    const int tid = threadIdx.x;
    int px, py, pz = 0;
    if (!D3) {
        const int tx = tile % ntx, ty = tile / ntx;
        px = tx * 16 + (tid >> 4);
        py = ty * 16 + (tid & 15);
    } else {
        const int txy = ntx * nty;
        const int tz = tile / txy, r = tile - tz * txy;
        const int tx = r % ntx, ty = r / ntx;
        px = tx * 8 + ((tid >> 3) & 7);
        py = ty * 8 + (tid & 7);
        pz = tz * 4 + (tid >> 6);
    }
    float2 acc[BC];
#pragma unroll
    for (int bb = 0; bb < BC; ++bb) acc[bb] = make_float2(0.f, 0.f);

    const int beg = tile_off[tile], end = tile_off[tile + 1];
    for (int base = beg; base < end; base += kChunk) {
        const int cnt = min(kChunk, end - base);
        __syncthreads();
        if (tid < cnt) {
            const int i = tile_idx[base + tid];
            const Win a = window(kx[i], Nx, os, halfw);
            const Win b = window(ky[i], Ny, os, halfw);
            float t[NT];
            taps<NT>(a, os, W2, lut, S, t);
#pragma unroll
            for (int j = 0; j < NT; ++j) s_w[tid][0][j] = t[j];
            taps<NT>(b, os, W2, lut, S, t);
#pragma unroll
            for (int j = 0; j < NT; ++j) s_w[tid][1][j] = t[j];
            s_lo[tid][0] = a.lo;
            s_lo[tid][1] = b.lo;
            if (D3) {
                const Win z = window(kz[i], Nz, os, halfw);
                taps<NT>(z, os, W2, lut, S, t);
#pragma unroll
                for (int j = 0; j < NT; ++j) s_w[tid][2][j] = t[j];
                s_lo[tid][2] = z.lo;
            } else {
                s_lo[tid][2] = 0;
            }
        }
        for (int t = tid; t < cnt * BC; t += blockDim.x) {
            const int sl = t / BC, bb = t - sl * BC;
            const int b = b0 + bb;
            float2 v = make_float2(0.f, 0.f);
            if (b < B) {
                const int i = tile_idx[base + sl];
                const int c = b / L, l = b - c * L;
                const long long di = c * in_cstride + i;
                v = in[di];
                if (resid_y) {
                    // W % (yi - Ax)   (solve_pwls_pcg.hpp:83)
                    const float2 yv = resid_y[di];
                    v = make_float2(yv.x - v.x, yv.y - v.y);
                    if (Wd) {
                        const float wv = Wd[di];
                        v = make_float2(wv * v.x, wv * v.y);
                    }
                }
                if (AA) {
                    // conj(AA) % d with AA real  (TimeSegmentation.cpp:304-310)
                    const float a = AA[(long long)l * M + i];
                    v = make_float2(a * v.x, a * v.y);
                }
            }
            s_v[sl][bb] = v;
        }
        __syncthreads();
        for (int sl = 0; sl < cnt; ++sl) {
            const unsigned dx = (unsigned)(px - s_lo[sl][0]);
            const unsigned dy = (unsigned)(py - s_lo[sl][1]);
            const unsigned dz = D3 ? (unsigned)(pz - s_lo[sl][2]) : 0u;
            if (dx < (unsigned)NT && dy < (unsigned)NT && dz < (unsigned)(D3 ? NT : 1)) {
                float w = s_w[sl][0][dx] * s_w[sl][1][dy];
                if (D3) w = w * s_w[sl][2][dz];
#pragma unroll
                for (int bb = 0; bb < BC; ++bb) {
                    const float2 v = s_v[sl][bb];
                    acc[bb].x += w * v.x;
                    acc[bb].y += w * v.y;
                }
            }
        }
    }
    if (px < Gx && py < Gy && pz < (D3 ? Gz : 1)) {
        const long long p = (long long)pz * Gx * Gy + (long long)px * Gy + py;
#pragma unroll
        for (int bb = 0; bb < BC; ++bb)
            if (b0 + bb < B) grids[(long long)(b0 + bb) * G + p] = acc[bb];
    }
}

template <int NT, bool D3>
void spread_dispatch(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc, int L,
                     const float* AA, const float2* resid_y, const float* Wd, float2* grids,
                     hipStream_t s)
{
    const pg::Geometry& g = G.geo;
    const int B = nc * L;
    const int ntiles = G.ntx * G.nty * G.ntz;
    const float halfw = g.W * g.os / 2.0f;
    const float W2 = g.W * g.W;
    if (B == 1) {
        hipLaunchKernelGGL((k_spread<NT, D3, 1>), dim3(ntiles, 1), dim3(256), 0, s, G.kx, G.ky,
                           G.kz, G.M, g.Nx, g.Ny, g.Nz, g.Gx, g.Gy, g.Gz, g.os, halfw, W2, G.lut,
                           G.S, G.tile_off, G.tile_idx, G.ntx, G.nty, in, in_cstride, resid_y, Wd,
                           L, AA, B, grids, g.G());
    } else {
        constexpr int BC = 16;
        hipLaunchKernelGGL((k_spread<NT, D3, BC>), dim3(ntiles, (B + BC - 1) / BC), dim3(256), 0,
                           s, G.kx, G.ky, G.kz, G.M, g.Nx, g.Ny, g.Nz, g.Gx, g.Gy, g.Gz, g.os,
                           halfw, W2, G.lut, G.S, G.tile_off, G.tile_idx, G.ntx, G.nty, in,
                           in_cstride, resid_y, Wd, L, AA, B, grids, g.G());
    }
}

template <int NT, bool D3>
void interp_dispatch(const pg::Gnufft& G, const float2* grids, int nc, int L, const float* AA,
                     float2* out, long long out_cstride, hipStream_t s)
{
    const pg::Geometry& g = G.geo;
    const float halfw = g.W * g.os / 2.0f;
    hipLaunchKernelGGL((k_interp<NT, D3>), dim3((G.M + 255) / 256, nc), dim3(256), 0, s, G.kx,
                       G.ky, G.kz, G.M, g.Nx, g.Ny, g.Nz, g.Gx, g.Gy, g.os, halfw, g.W * g.W,
                       G.lut, G.S, grids, g.G(), L, AA, out, out_cstride);
}

}  // namespace

namespace pgk {

void launch_interp(const pg::Gnufft& G, const float2* grids, int nc, int L, const float* AA,
                   float2* out, long long out_cstride, hipStream_t s)
{
    const bool d3 = G.geo.dims == 3;
    switch (G.geo.ntaps) {
    case 8:
        d3 ? interp_dispatch<8, true>(G, grids, nc, L, AA, out, out_cstride, s)
           : interp_dispatch<8, false>(G, grids, nc, L, AA, out, out_cstride, s);
        break;
    case 6:
        d3 ? interp_dispatch<6, true>(G, grids, nc, L, AA, out, out_cstride, s)
           : interp_dispatch<6, false>(G, grids, nc, L, AA, out, out_cstride, s);
        break;
    default:
        throw pg::Error{PG_ERR_UNSUPPORTED, "kernel width * gridos must give 6 or 8 taps"};
    }
}

void launch_spread(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc, int L,
                   const float* AA, const float2* resid_y, const float* Wd, float2* grids,
                   hipStream_t s)
{
    const bool d3 = G.geo.dims == 3;
    switch (G.geo.ntaps) {
    case 8:
        d3 ? spread_dispatch<8, true>(G, in, in_cstride, nc, L, AA, resid_y, Wd, grids, s)
           : spread_dispatch<8, false>(G, in, in_cstride, nc, L, AA, resid_y, Wd, grids, s);
        break;
    case 6:
        d3 ? spread_dispatch<6, true>(G, in, in_cstride, nc, L, AA, resid_y, Wd, grids, s)
           : spread_dispatch<6, false>(G, in, in_cstride, nc, L, AA, resid_y, Wd, grids, s);
        break;
    default:
        throw pg::Error{PG_ERR_UNSUPPORTED, "kernel width * gridos must give 6 or 8 taps"};
    }
}

}  // namespace pgk
