// pg_grid.hip -- Kaiser-Bessel interpolation (forward gather) and spreading
// (adjoint scatter) on the oversampled grid, batched over coils x segments.
//
// Tap weights are precomputed per sample (pg_gnufft.h); taps outside the
// clamped window of gridding.cpp:67-73 carry weight 0 and their addresses
// are clamped into the grid, so no tap wraps around.
//
// interp  -- one thread per (sample, coil); the L segment grids of the coil
//            are read with the same 64 (2-D) / 512 (3-D) weights and summed
//            with the AA weights (TimeSegmentation.cpp:279-295).  Coils are
//            mapped to XCDs so that each XCD's L2 holds its coils' grids.
// spread  -- first the value of every (sample, grid) pair is formed once
//            (AA weight, CG residual), sample-major.
//            2-D: one wavefront per 4x4 grid patch, lane = grid
//            (coil x segment), accumulating in registers over the samples
//            whose window meets the patch; heavy patches are split into
//            chunks whose partial sums are added in a fixed order.  No
//            atomics, no zero-fill, deterministic.
//            3-D, >= 16 grids: one block per 8x8x4 tile, thread per point.
//            3-D, < 16 grids: one thread per sample, hardware float atomics
//            into a zeroed grid (the reference's OpenACC spread,
//            gridding.cpp:102-106).
#include "pg_gnufft.h"
#include "pg_internal.h"

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// ---------------------------------------------------------------- interp --
// Four lanes per sample: lane q of a sample reads the taps 2q, 2q+1 along y
// of every row, so one load instruction covers 16 samples x 64 contiguous
// bytes; the four partial sums are combined with two lane shuffles.
template <int NT, bool D3>
__global__ void __launch_bounds__(256) k_interp(
    const int* __restrict__ lo, const float* __restrict__ w, int M, int Gx, int Gy, int Gz,
    const float2* __restrict__ grids, long long G, int nc, int L, const float* __restrict__ AA,
    float2* __restrict__ out, long long out_cstride, int nsb, int xcd_map)
{
    int c, sb;
    const int p = blockIdx.x;
    if (xcd_map) {
        // consecutive block ids go to consecutive XCDs: keep every coil on
        // one XCD (nc is a multiple of 8 here)
        const int xcd = p & 7, r = p >> 3;
        const int cl = r / nsb;
        sb = r - cl * nsb;
        c = cl * 8 + xcd;
    } else {
        c = p / nsb;
        sb = p - c * nsb;
    }
    if (c >= nc) return;
    const int q = threadIdx.x & 3;
    const int i = sb * (int)(blockDim.x >> 2) + (int)(threadIdx.x >> 2);
    const bool valid = i < M;
    const int ii = valid ? i : M - 1;
    const int lx = lo[ii], ly = lo[M + ii];
    const int lz = D3 ? lo[2 * M + ii] : 0;
    const int j0 = 2 * q, j1 = 2 * q + 1;
    const float wy0 = (j0 < NT) ? w[(long long)(NT + j0) * M + ii] : 0.f;
    const float wy1 = (j1 < NT) ? w[(long long)(NT + j1) * M + ii] : 0.f;
    const int oy0 = clampi(ly + j0, 0, Gy - 1), oy1 = clampi(ly + j1, 0, Gy - 1);
    float wx[NT], wz[D3 ? NT : 1];
    int ox[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        wx[j] = w[(long long)j * M + ii];
        ox[j] = clampi(lx + j, 0, Gx - 1) * Gy;
    }
    long long oz[D3 ? NT : 1];
    if (D3) {
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            wz[j] = w[(long long)(2 * NT + j) * M + ii];
            oz[j] = (long long)clampi(lz + j, 0, Gz - 1) * Gx * Gy;
        }
    } else {
        wz[0] = 1.f;
        oz[0] = 0;
    }
    float tr = 0.f, ti = 0.f;
#pragma unroll 1
    for (int l = 0; l < L; ++l) {
        const float2* g = grids + ((long long)c * L + l) * G;
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int jz = 0; jz < (D3 ? NT : 1); ++jz) {
            const float2* gz = g + oz[jz];
#pragma unroll
            for (int jx = 0; jx < NT; ++jx) {
                const float2* row = gz + ox[jx];
                const float2 v0 = row[oy0], v1 = row[oy1];
                const float rr = wy0 * v0.x + wy1 * v1.x;
                const float ri = wy0 * v0.y + wy1 * v1.y;
                const float wxz = D3 ? wx[jx] * wz[jz] : wx[jx];
                ar += wxz * rr;
                ai += wxz * ri;
            }
        }
        if (AA) {
            const float a = AA[(long long)l * M + ii];
            ar *= a;
            ai *= a;
        }
        tr += ar;
        ti += ai;
    }
    tr += __shfl_xor(tr, 1);
    ti += __shfl_xor(ti, 1);
    tr += __shfl_xor(tr, 2);
    ti += __shfl_xor(ti, 2);
    if (q == 0 && valid) out[c * out_cstride + i] = make_float2(tr, ti);
}

// value of sample i in grid b: AA_l * (in_c[i] or W (y_c[i] - Ax_c[i])),
// the residual formed in FP64 from the FP64 Ax of the CG
__device__ __forceinline__ float2 spread_value(const float2* __restrict__ in, long long in_cstride,
                                               const float2* __restrict__ resid_y,
                                               const double2* __restrict__ ax64,
                                               const float* __restrict__ Wd,
                                               const float* __restrict__ AA, int L, int M, int b,
                                               int i)
{
    const int c = b / L, l = b - c * L;
    const long long di = c * in_cstride + i;
    float2 v;
    if (resid_y) {
        // W % (yi - Ax)   (solve_pwls_pcg.hpp:83)
        const float2 yv = resid_y[di];
        const double2 a = ax64[di];
        double rr = (double)yv.x - a.x, ri = (double)yv.y - a.y;
        if (Wd) {
            const double wv = Wd[di];
            rr *= wv;
            ri *= wv;
        }
        v = make_float2((float)rr, (float)ri);
    } else {
        v = in[di];
    }
    if (AA) {
        // conj(AA) % d with AA real  (TimeSegmentation.cpp:304-310)
        const float a = AA[(long long)l * M + i];
        v = make_float2(a * v.x, a * v.y);
    }
    return v;
}

// V[i*B + b] = value of sample i in grid b, formed once per spread so that the
// spread kernels read one contiguous row of B values per sample
__global__ void __launch_bounds__(256) k_spread_values(
    const float2* __restrict__ in, long long in_cstride, const float2* __restrict__ resid_y,
    const double2* __restrict__ ax64, const float* __restrict__ Wd, int L,
    const float* __restrict__ AA, int M, int B, float2* __restrict__ V)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)M * B) return;
    const int i = (int)(t / B), b = (int)(t - (long long)i * B);
    V[t] = spread_value(in, in_cstride, resid_y, ax64, Wd, AA, L, M, b, i);
}

// ------------------------------------------------- 2-D patch spread (B>=16)
__global__ void __launch_bounds__(64) k_spread_patch(
    const int4* __restrict__ items, const int* __restrict__ entries, const int* __restrict__ lo,
    const float* __restrict__ wpad, int M, int npy, const float2* __restrict__ V, int B,
    float2* __restrict__ grids, long long G, int Gy, float2* __restrict__ partial, int n_slots)
{
    const int4 it = items[blockIdx.x];
    const int b = blockIdx.y * 64 + threadIdx.x;
    const bool active = b < B;
    const int bb = active ? b : B - 1;
    const int px0 = (it.x / npy) * pg::kPatch, py0 = (it.x % npy) * pg::kPatch;
    float2 acc[pg::kPatch][pg::kPatch];
#pragma unroll
    for (int a = 0; a < pg::kPatch; ++a)
#pragma unroll
        for (int q = 0; q < pg::kPatch; ++q) acc[a][q] = make_float2(0.f, 0.f);

    // entries are processed in groups of EB: all loads of a group are issued
    // before its FMAs, so several memory round trips overlap
    constexpr int EB = 8;
    for (int e0 = it.y; e0 < it.z; e0 += EB) {
        const int n = min(EB, it.z - e0);
        float wx[EB][pg::kPatch], wy[EB][pg::kPatch];
        float2 v[EB];
#pragma unroll
        for (int k = 0; k < EB; ++k) {
            const int i = entries[e0 + min(k, n - 1)];
            const int dx = px0 - lo[i] + 3, dy = py0 - lo[M + i] + 3;  // in [0, 10]
            const float* wxp = wpad + (long long)i * 16 + dx;
            const float* wyp = wpad + ((long long)M + i) * 16 + dy;
#pragma unroll
            for (int a = 0; a < pg::kPatch; ++a) {
                wx[k][a] = (k < n) ? wxp[a] : 0.f;
                wy[k][a] = wyp[a];
            }
            v[k] = V[(long long)i * B + bb];
        }
#pragma unroll
        for (int k = 0; k < EB; ++k)
#pragma unroll
            for (int a = 0; a < pg::kPatch; ++a)
#pragma unroll
                for (int q = 0; q < pg::kPatch; ++q) {
                    const float ww = wx[k][a] * wy[k][q];
                    acc[a][q].x += ww * v[k].x;
                    acc[a][q].y += ww * v[k].y;
                }
    }
    if (!active) return;
    if (it.w < 0) {
        float2* g = grids + (long long)b * G + (long long)px0 * Gy + py0;
#pragma unroll
        for (int a = 0; a < pg::kPatch; ++a)
#pragma unroll
            for (int q = 0; q < pg::kPatch; ++q) g[(long long)a * Gy + q] = acc[a][q];
    } else {
        float2* pp = partial + ((long long)blockIdx.y * n_slots + it.w) * 16 * 64 + threadIdx.x;
#pragma unroll
        for (int a = 0; a < pg::kPatch; ++a)
#pragma unroll
            for (int q = 0; q < pg::kPatch; ++q) pp[(a * pg::kPatch + q) * 64] = acc[a][q];
    }
}

// sum the partial patches of split patches, in slot order; 4 wavefronts per
// patch, each summing 4 of its 16 points for 64 grids
__global__ void __launch_bounds__(256) k_spread_combine(const int4* __restrict__ comb, int npy,
                                                        const float2* __restrict__ partial,
                                                        int n_slots, int B,
                                                        float2* __restrict__ grids, long long G,
                                                        int Gy)
{
    const int4 cb = comb[blockIdx.x];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int b = blockIdx.y * 64 + lane;
    if (b >= B) return;
    const int px0 = (cb.x / npy) * pg::kPatch, py0 = (cb.x % npy) * pg::kPatch;
    const float2* pp = partial + ((long long)blockIdx.y * n_slots + cb.y) * 16 * 64 + lane;
    float2* g = grids + (long long)b * G + (long long)px0 * Gy + py0;
    float2 s[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) s[r] = pp[(wv * 4 + r) * 64];
    for (int k = 1; k < cb.z; ++k) {
        float2 v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = pp[((long long)k * 16 + wv * 4 + r) * 64];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            s[r].x += v[r].x;
            s[r].y += v[r].y;
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int p = wv * 4 + r;
        g[(long long)(p / pg::kPatch) * Gy + (p % pg::kPatch)] = s[r];
    }
}

// ------------------------------------------- 3-D tile spread (B >= 16) ----
constexpr int kChunk = 64;

template <int NT, int BC>
__global__ void __launch_bounds__(256) k_spread_tile3(
    const int* __restrict__ lo, const float* __restrict__ w, int M, int Gx, int Gy, int Gz,
    const int* __restrict__ tile_off, const int* __restrict__ tile_idx, int ntx, int nty,
    const float2* __restrict__ V, int B, float2* __restrict__ grids, long long G)
{
    __shared__ int s_lo[kChunk][3];
    __shared__ float s_w[kChunk][3][NT];
    __shared__ float2 s_v[kChunk][BC];

    const int tile = blockIdx.x;
    const int b0 = blockIdx.y * BC;
    const int tid = threadIdx.x;
    const int txy = ntx * nty;
    const int tz = tile / txy, r = tile - tz * txy;
    const int tx = r % ntx, ty = r / ntx;
    const int px = tx * 8 + ((tid >> 3) & 7);
    const int py = ty * 8 + (tid & 7);
    const int pz = tz * 4 + (tid >> 6);
    float2 acc[BC];
#pragma unroll
    for (int bb = 0; bb < BC; ++bb) acc[bb] = make_float2(0.f, 0.f);

    const int beg = tile_off[tile], end = tile_off[tile + 1];
    for (int base = beg; base < end; base += kChunk) {
        const int cnt = min(kChunk, end - base);
        __syncthreads();
        if (tid < cnt) {
            const int i = tile_idx[base + tid];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s_lo[tid][d] = lo[(long long)d * M + i];
#pragma unroll
                for (int j = 0; j < NT; ++j) s_w[tid][d][j] = w[(long long)(d * NT + j) * M + i];
            }
        }
        for (int t = tid; t < cnt * BC; t += blockDim.x) {
            const int sl = t / BC, bb = t - sl * BC;
            const int b = b0 + bb;
            s_v[sl][bb] = (b < B) ? V[(long long)tile_idx[base + sl] * B + b] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        for (int sl = 0; sl < cnt; ++sl) {
            const unsigned dx = (unsigned)(px - s_lo[sl][0]);
            const unsigned dy = (unsigned)(py - s_lo[sl][1]);
            const unsigned dz = (unsigned)(pz - s_lo[sl][2]);
            if (dx < (unsigned)NT && dy < (unsigned)NT && dz < (unsigned)NT) {
                const float ww = s_w[sl][0][dx] * s_w[sl][1][dy] * s_w[sl][2][dz];
#pragma unroll
                for (int bb = 0; bb < BC; ++bb) {
                    const float2 v = s_v[sl][bb];
                    acc[bb].x += ww * v.x;
                    acc[bb].y += ww * v.y;
                }
            }
        }
    }
    if (px < Gx && py < Gy && pz < Gz) {
        const long long p = (long long)pz * Gx * Gy + (long long)px * Gy + py;
#pragma unroll
        for (int bb = 0; bb < BC; ++bb)
            if (b0 + bb < B) grids[(long long)(b0 + bb) * G + p] = acc[bb];
    }
}

// --------------------------------------- atomic spread (one NUFFT at a time)
__global__ void k_zero(float2* __restrict__ p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float2(0.f, 0.f);
}

template <int NT, bool D3>
__global__ void __launch_bounds__(256) k_spread_atomic(
    const int* __restrict__ lo, const float* __restrict__ w, int M, int Gx, int Gy, int Gz,
    const float2* __restrict__ in, long long in_cstride, const float2* __restrict__ resid_y,
    const double2* __restrict__ ax64, const float* __restrict__ Wd, int L,
    const float* __restrict__ AA, int B, float2* __restrict__ grids, long long G)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= M) return;
    const float2 v = spread_value(in, in_cstride, resid_y, ax64, Wd, AA, L, M, b, i);
    const int lx = lo[i], ly = lo[M + i], lz = D3 ? lo[2 * M + i] : 0;
    float* g = reinterpret_cast<float*>(grids + (long long)b * G);
    for (int jz = 0; jz < (D3 ? NT : 1); ++jz) {
        const int nz = lz + jz;
        const float wz = D3 ? w[(long long)(2 * NT + jz) * M + i] : 1.f;
        if (D3 && (wz == 0.f || nz < 0 || nz >= Gz)) continue;
        for (int jx = 0; jx < NT; ++jx) {
            const int nx = lx + jx;
            const float wx = w[(long long)jx * M + i];
            if (wx == 0.f || nx < 0 || nx >= Gx) continue;
            for (int jy = 0; jy < NT; ++jy) {
                const int ny = ly + jy;
                const float wy = w[(long long)(NT + jy) * M + i];
                if (wy == 0.f || ny < 0 || ny >= Gy) continue;
                const float ww = D3 ? wx * wy * wz : wx * wy;
                const long long idx = (long long)nz * Gx * Gy + (long long)nx * Gy + ny;
                unsafeAtomicAdd(g + 2 * idx, ww * v.x);
                unsafeAtomicAdd(g + 2 * idx + 1, ww * v.y);
            }
        }
    }
}

template <int NT, bool D3>
void interp_dispatch(const pg::Gnufft& G, const float2* grids, int nc, int L, const float* AA,
                     float2* out, long long out_cstride, hipStream_t s)
{
    const pg::Geometry& g = G.geo;
    const int nsb = (G.M + 63) / 64;  // 64 samples (4 lanes each) per block
    const int xmap = (nc % 8 == 0) ? 1 : 0;
    hipLaunchKernelGGL((k_interp<NT, D3>), dim3(nsb * nc), dim3(256), 0, s, G.lo, G.w, G.M, g.Gx,
                       g.Gy, g.Gz, grids, g.G(), nc, L, AA, out, out_cstride, nsb, xmap);
}

template <int NT, bool D3>
void spread_dispatch(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc, int L,
                     const float* AA, const float2* resid_y, const double2* ax64, const float* Wd,
                     float2* grids, hipStream_t s)
{
    const pg::Geometry& g = G.geo;
    const int B = nc * L;
    if (D3 && B < 16) {
        const long long n = (long long)B * g.G();
        hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, grids, n);
        hipLaunchKernelGGL((k_spread_atomic<NT, D3>), dim3((G.M + 255) / 256, B), dim3(256), 0, s,
                           G.lo, G.w, G.M, g.Gx, g.Gy, g.Gz, in, in_cstride, resid_y, ax64, Wd, L,
                           AA, B, grids, g.G());
        return;
    }
    // per-sample values of every grid, once
    const long long nv = (long long)G.M * B;
    hipLaunchKernelGGL(k_spread_values, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, in,
                       in_cstride, resid_y, ax64, Wd, L, AA, G.M, B, G.vals);
    if (!D3) {
        const int nch = (B + 63) / 64;
        const int npy = g.Gy / pg::kPatch;
        hipLaunchKernelGGL(k_spread_patch, dim3(G.n_items, nch), dim3(64), 0, s, G.items, G.entries,
                           G.lo, G.wpad, G.M, npy, G.vals, B, grids, g.G(), g.Gy, G.partial,
                           G.n_slots);
        if (G.n_combine > 0)
            hipLaunchKernelGGL(k_spread_combine, dim3(G.n_combine, nch), dim3(256), 0, s, G.combine,
                               npy, G.partial, G.n_slots, B, grids, g.G(), g.Gy);
        return;
    }
    constexpr int BC = 16;
    const int ntiles = G.ntx * G.nty * G.ntz;
    hipLaunchKernelGGL((k_spread_tile3<NT, BC>), dim3(ntiles, (B + BC - 1) / BC), dim3(256), 0, s,
                       G.lo, G.w, G.M, g.Gx, g.Gy, g.Gz, G.tile_off, G.tile_idx, G.ntx, G.nty,
                       G.vals, B, grids, g.G());
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
                   const float* AA, const float2* resid_y, const double2* ax64, const float* Wd,
                   float2* grids, hipStream_t s)
{
    const bool d3 = G.geo.dims == 3;
    switch (G.geo.ntaps) {
    case 8:
        d3 ? spread_dispatch<8, true>(G, in, in_cstride, nc, L, AA, resid_y, ax64, Wd, grids, s)
           : spread_dispatch<8, false>(G, in, in_cstride, nc, L, AA, resid_y, ax64, Wd, grids, s);
        break;
    case 6:
        d3 ? spread_dispatch<6, true>(G, in, in_cstride, nc, L, AA, resid_y, ax64, Wd, grids, s)
           : spread_dispatch<6, false>(G, in, in_cstride, nc, L, AA, resid_y, ax64, Wd, grids, s);
        break;
    default:
        throw pg::Error{PG_ERR_UNSUPPORTED, "kernel width * gridos must give 6 or 8 taps"};
    }
}

}  // namespace pgk
