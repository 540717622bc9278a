// pg_grid.hip -- Kaiser-Bessel interpolation (forward gather) and spreading
// (adjoint scatter) on the oversampled grid, batched over coils x segments.
//
// Tap weights are precomputed per sample (pg_gnufft.h); taps outside the
// clamped window of gridding.cpp:67-73 carry weight 0 and their addresses
// are clamped into the grid, so no tap wraps around.
//
// interp  -- 2-D with 2..16 segments: the grid region under a tile of
//            samples is staged in LDS for all segments of a coil and each
//            sample is one wavefront (lane = segment x y-tap).  Otherwise
//            four lanes per (sample, coil) gather from global memory.  In
//            both, the L segment grids of a coil are read with the same
//            weights and summed with the AA weights (TimeSegmentation.cpp:
//            279-295), and coils are mapped to XCDs so that each XCD's L2
//            holds its coils' grids.
// spread  -- first the value of every (sample, grid) pair is formed once
//            (AA weight, CG residual), sample-major.
//            2-D: one wavefront per 4x4 grid patch, lane = grid
//            (coil x segment), accumulating in registers over the samples
//            whose window meets the patch (per-entry patch weights are
//            precomputed); heavy patches are split into chunks of <= 64
//            entries whose partial sums are added in a fixed order.  No
//            atomics, no zero-fill, deterministic.
//            3-D, >= 16 grids: one block per 8x8x4 tile, thread per point.
//            3-D, < 16 grids: one thread per sample, hardware float atomics
//            into a zeroed grid (the reference's OpenACC spread,
//            gridding.cpp:102-106).
#include "pg_gnufft.h"
#include "pg_internal.h"
#include "pg_fftreg.h"

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// ---------------------------------------------------------------- interp --
// Four lanes per sample: lane q of a sample reads the taps q, q+4, q+8 (as
// many as NT has) along y of every row, so one load instruction covers 16
// samples x 32 contiguous bytes; the four partial sums are combined with two
// lane shuffles.
template <int NT, bool D3>
__global__ void __launch_bounds__(256) k_interp(
    const int* __restrict__ lo, const float* __restrict__ w, int M, int Gx, int Gy, int Gz,
    const float2* __restrict__ grids, long long G, int nc, int L, const float* __restrict__ AA,
    const float* __restrict__ AAi, float2* __restrict__ out, long long out_cstride, int nsb,
    int xcd_map, const int* __restrict__ pl, long long pstride)
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
    // lane q takes the y taps q, q + 4, q + 8 (zero weight past NT)
    constexpr int JY = (NT + 3) / 4;
    float wy[JY];
    int oy[JY];
#pragma unroll
    for (int j = 0; j < JY; ++j) {
        const int jj = q + 4 * j;
        wy[j] = (jj < NT) ? w[(long long)(NT + jj) * M + ii] : 0.f;
        oy[j] = clampi(ly + jj, 0, Gy - 1);
    }
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
    // z-class planes (2-D gather on the plane of the sample's class)
    const float2* gp = (!D3 && pl) ? grids + (long long)pl[ii] * pstride : grids;
    float tr = 0.f, ti = 0.f;
#pragma unroll 1
    for (int l = 0; l < L; ++l) {
        const float2* g = gp + ((long long)c * L + l) * G;
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int jz = 0; jz < (D3 ? NT : 1); ++jz) {
            const float2* gz = g + oz[jz];
#pragma unroll
            for (int jx = 0; jx < NT; ++jx) {
                const float2* row = gz + ox[jx];
                float rr = 0.f, ri = 0.f;
#pragma unroll
                for (int j = 0; j < JY; ++j) {
                    const float2 v = row[oy[j]];
                    rr += wy[j] * v.x;
                    ri += wy[j] * v.y;
                }
                const float wxz = D3 ? wx[jx] * wz[jz] : wx[jx];
                ar += wxz * rr;
                ai += wxz * ri;
            }
        }
        if (AA) {
            const float a = AA[(long long)l * M + ii];
            if (AAi) {
                // complex weights, real and imaginary parts apart
                const float b = AAi[(long long)l * M + ii];
                const float nr = a * ar - b * ai;
                ai = a * ai + b * ar;
                ar = nr;
            } else {
                ar *= a;
                ai *= a;
            }
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

// ------------------------------------------------- 2-D interp through LDS --
// Sum over the 8 lanes of each aligned 8-lane group with DPP; the total ends
// in the group's lane 7 (row_half_mirror pairs lane k with lane 7 - k).
__device__ __forceinline__ float group8_sum_to_7(float v)
{
#define PG_DPP_ADD8(v, CTRL)                                                                 \
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), (CTRL), 0xF, 0xF, \
                                                    false))
    PG_DPP_ADD8(v, 0xB1);   // quad_perm [1,0,3,2]
    PG_DPP_ADD8(v, 0x4E);   // quad_perm [2,3,0,1]
    PG_DPP_ADD8(v, 0x141);  // row_half_mirror
#undef PG_DPP_ADD8
    return v;
}

// One block per (work item, coil): the kIR x kIR grid region under the
// item's tile is staged in LDS for all L segments of the coil ([x][l][y],
// y fastest, LP segment slots), together with the item's sample records and
// their AA weights.  Each wavefront then takes 8 samples at a time with
// lane = (sample s, segment l):
//   acc_l = sum_x wx[x] sum_y wy[y] g_l(ox + x, oy + y),  value = sum_l AA_l acc_l
// the segment sum being a 3-step DPP reduction inside each 8-lane group.
// Blocks are ordered coil-fastest so that coil c stays on XCD c % 8.
constexpr int kInterpWaves = 8;

template <int LP, bool CPX>
__global__ void __launch_bounds__(64 * kInterpWaves) k_interp_lds(
    const int4* __restrict__ items, const float4* __restrict__ rec, int ity, int padl,
    const float2* __restrict__ grids, long long G, int Gx, int Gy, int nc, int L,
    const float* __restrict__ AA, const float* __restrict__ AAi, int M, float2* __restrict__ out,
    long long out_cstride, long long pstride)
{
    constexpr int R = pg::kIR;
    constexpr int CH = pg::kIChunk;
    __shared__ float2 tile[R * LP * R];
    __shared__ float4 srec[CH * 5];
    __shared__ float saa[CH * LP];
    __shared__ float saai[CPX ? CH * LP : 1];
    const int c = blockIdx.x % nc;
    const int4 it = items[blockIdx.x / nc];
    const int n = it.z - it.y;  // 1..CH
    const int tx = it.x / ity, ty = it.x - (it.x / ity) * ity;
    const int rx0 = tx * pg::kIT - padl, ry0 = ty * pg::kIT - padl;
    // item plane it.w (z-class planes; 0 for a 2-D plan)
    const float2* gc = grids + (long long)it.w * pstride + (long long)c * L * G;
    // Staging: every global load of the block is issued before any LDS
    // store, then one barrier.  The region goes as rows (xr, l) of R values
    // along y, 32 lanes per row (R <= 32), row and segment by shifts, the
    // grid offset in 32 bits (L*G of one coil); only points inside the grid
    // are loaded.  The weights are read with the sample indices taken from
    // the records in global memory.
    static_assert(R <= 32 && (LP & (LP - 1)) == 0, "staging layout");
    constexpr int NTH = 64 * kInterpWaves;
    constexpr int PASS = NTH / 32;                  // rows per pass
    constexpr int NP = (R * LP + PASS - 1) / PASS;  // passes over the rows
    constexpr int NR = (CH * 5 + NTH - 1) / NTH;    // record loads per thread
    constexpr int NA = (CH * LP + NTH - 1) / NTH;   // weight loads per thread
    const int yr = threadIdx.x & 31;
    const int y = ry0 + yr;
    const bool yok = yr < R && y >= 0 && y < Gy;
    // loads from clamped, in-range addresses with the condition applied as a
    // select, so that none of them sits behind a branch and all are in flight
    // together
    const int yc = min(max(y, 0), Gy - 1);
    float2 tv[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int row = (int)(threadIdx.x >> 5) + k * PASS;
        const int l = row & (LP - 1), xr = row / LP;
        const int x = rx0 + xr;
        const bool ok = row < R * LP && yok && l < L && x >= 0 && x < Gx;
        const int lc = min(l, L - 1), xc = min(max(x, 0), Gx - 1);
        const float2 v = gc[lc * (int)G + xc * Gy + yc];
        tv[k] = ok ? v : make_float2(0.f, 0.f);
    }
    float4 rv[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int e = min((int)threadIdx.x + k * NTH, n * 5 - 1);
        rv[k] = rec[(long long)it.y * 5 + e];
    }
    float av[NA], aiv[CPX ? NA : 1];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int e = (int)threadIdx.x + k * NTH;
        const int sl = min(e / LP, n - 1), l = e & (LP - 1);
        const bool ok = e < n * LP && l < L;
        const int i = __float_as_int(rec[((long long)it.y + sl) * 5].x);
        const long long o = (long long)min(l, L - 1) * M + i;
        av[k] = ok ? AA[o] : 0.f;
        if (CPX) aiv[k] = ok ? AAi[o] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int row = (int)(threadIdx.x >> 5) + k * PASS;
        if (row < R * LP && yr < R) tile[row * R + yr] = tv[k];
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int e = (int)threadIdx.x + k * NTH;
        if (e < n * 5) srec[e] = rv[k];
    }
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int e = (int)threadIdx.x + k * NTH;
        if (e < n * LP) {
            saa[e] = av[k];
            if (CPX) saai[e] = aiv[k];
        }
    }
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // scalar sample loop
    const int lane = threadIdx.x & 63;
    const int s8 = lane >> 3, l8 = lane & 7;
    for (int base = wv * 8; base < n; base += kInterpWaves * 8) {
        const int sl = min(base + s8, n - 1);
        const float4* rp = srec + sl * 5;
        const float4 h = rp[0];
        const int ox = __float_as_int(h.y), oy = __float_as_int(h.z);
        const float4 xa = rp[1], xb = rp[2], ya = rp[3], yb = rp[4];
        const float wx[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
        const float wy[8] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y, yb.z, yb.w};
        float tr = 0.f, ti = 0.f;
#pragma unroll
        for (int g = 0; g < LP / 8; ++g) {
            const int l = g * 8 + l8;
            const float2* tp = tile + ((ox * LP + l) * R + oy);
            float ar = 0.f, ai = 0.f;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const float2* row = tp + x * LP * R;
                float rr = 0.f, ri = 0.f;
#pragma unroll
                for (int y = 0; y < 8; ++y) {
                    const float2 v = row[y];
                    rr += wy[y] * v.x;
                    ri += wy[y] * v.y;
                }
                ar += wx[x] * rr;
                ai += wx[x] * ri;
            }
            const float a = saa[sl * LP + l];
            if (CPX) {
                const float b = saai[sl * LP + l];
                tr += a * ar - b * ai;
                ti += a * ai + b * ar;
            } else {
                tr += a * ar;
                ti += a * ai;
            }
        }
        tr = group8_sum_to_7(tr);
        ti = group8_sum_to_7(ti);
        if (l8 == 7 && base + s8 < n)
            out[c * out_cstride + __float_as_int(h.x)] = make_float2(tr, ti);
    }
}

// The same interpolation with the segments staged four at a time: one block
// of 256 threads per (work item of <= 64 samples, coil); lane task = (sample
// s, segment slot l4) with s = tid / 4, so a block's 256 tasks cover its
// item's samples once per chunk of four segments.  Per chunk the 23 x 23
// region of the four segment grids is staged in LDS as [x][y][l] (17 KB: a
// point's four segments side by side, so a sample's four lanes read 32
// contiguous bytes at every tap): a thread loads its points (rows of 23
// along y across the threads: coalesced) from each of the four grids and
// stores each point as two 16-byte writes.  The stores follow the previous
// chunk's taps, and the loads of the next chunk are issued before this
// chunk's taps.  Each task forms
//   acc_l = sum_x wx[x] sum_y wy[y] g_l(ox + x, oy + y)
// and adds AA_l acc_l to its running sum; the four slots of a sample are
// added with two DPP steps at the end.  The sample records go straight to
// registers; the small LDS footprint (21 KB: the region and the AA weights)
// lets seven blocks share a CU, so one block's staging overlaps the others'
// taps.
constexpr int kI4Threads = 256;
constexpr int kI4Chunk = 64;
// the last word of a raw (stride 0) buffer resource on gfx950: 32-bit data
// format, no swizzle; loads at offsets at or past the resource's size return
// zero
constexpr int kBufferWord3 = 0x00020000;

template <bool CPX>
__global__ void __launch_bounds__(kI4Threads, CPX ? 6 : 7) k_interp_lds4(
    const int4* __restrict__ items, const float4* __restrict__ rec, int ity, int padl,
    const float2* __restrict__ grids, long long G, int Gx, int Gy, int nc, int L,
    const float* __restrict__ AA, const float* __restrict__ AAi, int M, float2* __restrict__ out,
    long long out_cstride, long long pstride)
{
    constexpr int R = pg::kIR;
    constexpr int LC = 4;
    constexpr int CH = kI4Chunk;
    constexpr int NTH = kI4Threads;
    constexpr int NQ = R * R;                  // points of the region
    constexpr int NP = (NQ + NTH - 1) / NTH;   // points per thread
    constexpr int PMAX = R + pg::kI4Pitches - 1;
    static_assert(CH * LC == NTH, "one task per thread");
    // the region's R rows at the item's pitch P (R .. PMAX points)
    __shared__ __align__(16) float2 tile[R * PMAX * LC];
    constexpr int NA = (CH * 16 + NTH - 1) / NTH;  // weight loads per thread
    __shared__ float saa[CH * 16];
    __shared__ float saai[CPX ? CH * 16 : 1];
    const int tid = threadIdx.x;
    // blocks (item, coil), coil fastest: the items of coil c run on XCD
    // c % 8, whose L2 serves neighbouring tiles' overlapping regions (item
    // fastest measured 40% slower)
    const int c = blockIdx.x % nc;
    const int4 it = items[blockIdx.x / nc];
    const int n = it.z - it.y;  // 1..CH
    const int tcode = it.x & 0xFFFFFF, P = R + (it.x >> 24);
    const int tx = tcode / ity, ty = tcode - (tcode / ity) * ity;
    const int rx0 = tx * pg::kIT - padl, ry0 = ty * pg::kIT - padl;
    // item plane it.w (z-class planes; 0 for a 2-D plan)
    const float2* gc = grids + (long long)it.w * pstride + (long long)c * L * G;
    const int s = tid >> 2, l4 = tid & 3;
    const bool live = s < n;
    int ox = 0, oy = 0, idx = 0;
    float wx[8], wy[8];
    float tr = 0.f, ti = 0.f;
    // Staging: point q = tid + k NTH of the region (xr = q / R, yr = q % R)
    // sits at byte offset voff[k] in each segment grid, the same for every
    // chunk.  The loads go through one buffer resource per segment slot of
    // the chunk, sized zero for a slot past L: a point outside the x-y grid
    // gets an offset of 2^31, and the hardware returns zero for it and for
    // every load of an empty slot, so the loads need neither clamps nor
    // selects and are all issued back to back.
    int voff[NP], soff[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int q = tid + k * NTH;
        const int xr = q / R, yr = q - (q / R) * R;
        const int x = rx0 + xr, y = ry0 + yr;
        const bool ok = q < NQ && x >= 0 && x < Gx && y >= 0 && y < Gy;
        voff[k] = ok ? (x * Gy + y) * 8 : (int)0x80000000u;
        soff[k] = (xr * P + yr) * LC;
    }
    float2 tv[NP][LC];
    auto load_chunk = [&](int g) {
#pragma unroll
        for (int lc = 0; lc < LC; ++lc) {
            const bool has = g * LC + lc < L;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(gc + (has ? (long long)(g * LC + lc) * G : 0LL)), (short)0,
                has ? (int)G * 8 : 0, kBufferWord3);
#pragma unroll
            for (int k = 0; k < NP; ++k)
                tv[k][lc] = __builtin_bit_cast(
                    float2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff[k], 0, 0));
        }
    };
    // The task's sample record (into registers: the four lanes of a sample
    // read the same 80 bytes), the sample indices of the (sample, segment
    // slot) pairs e = s * 16 + l, chunk 0 of the region and then the AA
    // weights (by those indices) are all requested before the first LDS
    // store; a record or weight of a sample past n or a slot past L reads
    // zero (offset past the resource's size, or 2^31).
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(rec + (long long)it.y * 5), (short)0, n * 5 * 16, kBufferWord3);
    float4 rv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        rv[k] = __builtin_bit_cast(
            float4, __builtin_amdgcn_raw_buffer_load_b128(rr, (s * 5 + k) * 16, 0, 0));
    int iv[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int e = tid + k * NTH;
        iv[k] = __float_as_int(rec[((long long)it.y + min(e >> 4, n - 1)) * 5].x);
    }
    load_chunk(0);
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc((void*)AA, (short)0, L * M * 4, kBufferWord3);
    const __amdgpu_buffer_rsrc_t rai =
        __builtin_amdgcn_make_buffer_rsrc((void*)(CPX ? AAi : AA), (short)0, L * M * 4,
                                          kBufferWord3);
    float av[NA], aiv[CPX ? NA : 1];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int e = tid + k * NTH, l = e & 15;
        const int o = (e >> 4) < n && l < L ? (l * M + iv[k]) * 4 : (int)0x80000000u;
        av[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ra, o, 0, 0));
        if (CPX)
            aiv[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rai, o, 0, 0));
    }
    auto store_chunk = [&]() {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const int q = tid + k * NTH;
            if (q < NQ) {
                float4* t4 = reinterpret_cast<float4*>(tile + soff[k]);
                t4[0] = make_float4(tv[k][0].x, tv[k][0].y, tv[k][1].x, tv[k][1].y);
                t4[1] = make_float4(tv[k][2].x, tv[k][2].y, tv[k][3].x, tv[k][3].y);
            }
        }
    };
    store_chunk();
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        saa[tid + k * NTH] = av[k];
        if (CPX) saai[tid + k * NTH] = aiv[k];
    }
    auto take_record = [&]() {
        const float4 h = rv[0];
        idx = __float_as_int(h.x);
        ox = __float_as_int(h.y);
        oy = __float_as_int(h.z);
        const float4 xa = rv[1], xb = rv[2], ya = rv[3], yb = rv[4];
        wx[0] = xa.x; wx[1] = xa.y; wx[2] = xa.z; wx[3] = xa.w;
        wx[4] = xb.x; wx[5] = xb.y; wx[6] = xb.z; wx[7] = xb.w;
        wy[0] = ya.x; wy[1] = ya.y; wy[2] = ya.z; wy[3] = ya.w;
        wy[4] = yb.x; wy[5] = yb.y; wy[6] = yb.z; wy[7] = yb.w;
    };
    // the taps of chunk g (in LDS) for this task's segment
    auto taps = [&](int g) {
        const int l = g * LC + l4;
        if (live && l < L) {
            const float2* tp = tile + ((ox * P + oy) * LC + l4);
            float ar = 0.f, ai = 0.f;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const float2* row = tp + x * P * LC;
                float rr = 0.f, ri = 0.f;
#pragma unroll
                for (int y = 0; y < 8; ++y) {
                    const float2 v = row[y * LC];
                    rr += wy[y] * v.x;
                    ri += wy[y] * v.y;
                }
                ar += wx[x] * rr;
                ai += wx[x] * ri;
            }
            const float a = saa[s * 16 + l];
            if (CPX) {
                const float b = saai[s * 16 + l];
                tr += a * ar - b * ai;
                ti += a * ai + b * ar;
            } else {
                tr += a * ar;
                ti += a * ai;
            }
        }
    };
    for (int g = 0;; ++g) {
        __syncthreads();  // chunk g is in LDS
        if (g == 0) take_record();
        // the next chunk's loads are in flight during this chunk's taps
        const bool more = (g + 1) * LC < L;
        if (more) load_chunk(g + 1);
        taps(g);
        if (!more) break;
        __syncthreads();  // chunk g's taps are done
        store_chunk();
    }
    // the four segment slots of a sample (a quad of lanes)
#define PG_DPP_ADDQ(v, CTRL)                                                                 \
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), (CTRL), 0xF, 0xF, \
                                                    false))
    PG_DPP_ADDQ(tr, 0xB1);  // quad_perm [1,0,3,2]
    PG_DPP_ADDQ(ti, 0xB1);
    PG_DPP_ADDQ(tr, 0x4E);  // quad_perm [2,3,0,1]
    PG_DPP_ADDQ(ti, 0x4E);
#undef PG_DPP_ADDQ
    if (l4 == 0 && live) out[c * out_cstride + idx] = make_float2(tr, ti);
}

// value of sample i in grid b: AA_l * (in_c[i] or W (y_c[i] - Ax_c[i])),
// the residual formed in FP64 from the FP64 Ax of the CG
__device__ __forceinline__ float2 spread_value(const float2* __restrict__ in, long long in_cstride,
                                               const float2* __restrict__ resid_y,
                                               const double2* __restrict__ ax64,
                                               const float* __restrict__ Wd,
                                               const float* __restrict__ AA,
                                               const float* __restrict__ AAi, int L, int M, int b,
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
        // conj(AA) % d  (TimeSegmentation.cpp:304-310); AAi, when set, holds
        // the imaginary parts of complex weights
        const float a = AA[(long long)l * M + i];
        if (AAi) {
            const float b = AAi[(long long)l * M + i];
            v = make_float2(a * v.x + b * v.y, a * v.y - b * v.x);
        } else {
            v = make_float2(a * v.x, a * v.y);
        }
    }
    return v;
}

// V[i*B + b] = value of sample i in grid b, formed once per spread so that the
// spread kernels read one contiguous row of B values per sample.  Block:
// 4 samples x 64 lanes over b.  (Inside the PCG the update kernel forms the
// next iteration's values itself, with the same arithmetic.)
__global__ void __launch_bounds__(256) k_spread_values(
    const float2* __restrict__ in, long long in_cstride, const float2* __restrict__ resid_y,
    const double2* __restrict__ ax64, const float* __restrict__ Wd, int L,
    const float* __restrict__ AA, const float* __restrict__ AAi, int M, int B,
    float2* __restrict__ V)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= M) return;
    for (int b = threadIdx.x & 63; b < B; b += 64)
        V[(long long)i * B + b] =
            spread_value(in, in_cstride, resid_y, ax64, Wd, AA, AAi, L, M, b, i);
}

// the same for an even grid count, two grids per thread (16-byte stores;
// consecutive threads fill consecutive pairs of a sample's row)
__global__ void __launch_bounds__(256) k_spread_values2(
    const float2* __restrict__ in, long long in_cstride, const float2* __restrict__ resid_y,
    const double2* __restrict__ ax64, const float* __restrict__ Wd, int L,
    const float* __restrict__ AA, const float* __restrict__ AAi, int M, int B,
    float2* __restrict__ V)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int hb = B >> 1;
    const long long i = p / hb;
    if (i >= M) return;
    const int b = 2 * (int)(p - i * hb);
    const float2 v0 = spread_value(in, in_cstride, resid_y, ax64, Wd, AA, AAi, L, M, b, (int)i);
    const float2 v1 = spread_value(in, in_cstride, resid_y, ax64, Wd, AA, AAi, L, M, b + 1, (int)i);
    *reinterpret_cast<float4*>(V + i * B + b) = make_float4(v0.x, v0.y, v1.x, v1.y);
}

void launch_values(const float2* in, long long in_cstride, const float2* resid_y,
                   const double2* ax64, const float* Wd, int L, const float* AA, const float* AAi,
                   int M, int B, float2* V, hipStream_t s)
{
    if (B % 2 == 0) {
        const long long n = (long long)M * (B / 2);
        hipLaunchKernelGGL(k_spread_values2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in,
                           in_cstride, resid_y, ax64, Wd, L, AA, AAi, M, B, V);
    } else {
        hipLaunchKernelGGL(k_spread_values, dim3((unsigned)((M + 3) / 4)), dim3(256), 0, s, in,
                           in_cstride, resid_y, ax64, Wd, L, AA, AAi, M, B, V);
    }
}

// Compact sample rows for the spread of a time-segmented batch: row i holds
// the nc coil values r_c(i) (the input, or W (y - Ax) rounded to FP32, as
// spread_value forms them) as float2, then the L segment weights AA_l(i) (and,
// for complex weights, their imaginary parts), padded to S floats (a multiple
// of 4).  The spread forms each grid's value AA_l r_c from the row, with the
// products of spread_value, so both spreads sum the same values; the rows
// are S floats per sample instead of 2 nc L.  Block: 64 samples, staged in
// LDS and written out as contiguous rows.
__global__ void __launch_bounds__(256) k_spread_rows(
    const float2* __restrict__ in, long long in_cstride, const float2* __restrict__ resid_y,
    const double2* __restrict__ ax64, const float* __restrict__ Wd, int L,
    const float* __restrict__ AA, const float* __restrict__ AAi, int M, int nc, int S,
    float* __restrict__ rows)
{
    extern __shared__ float buf[];  // 64 x S
    const int i0 = blockIdx.x * 64;
    const int tid = threadIdx.x;
    const int n = min(64, M - i0);
    for (int e = tid; e < nc * 64; e += 256) {
        const int c = e >> 6, k = e & 63;
        float2 v = make_float2(0.f, 0.f);
        if (k < n) v = spread_value(in, in_cstride, resid_y, ax64, Wd, nullptr, nullptr, 1, M, c,
                                    i0 + k);
        buf[k * S + 2 * c] = v.x;
        buf[k * S + 2 * c + 1] = v.y;
    }
    const int na = AAi ? 2 * L : L;
    for (int e = tid; e < na * 64; e += 256) {
        const int q = e >> 6, k = e & 63;
        const int i = min(i0 + k, M - 1);
        buf[k * S + 2 * nc + q] = q < L ? AA[(long long)q * M + i] : AAi[(long long)(q - L) * M + i];
    }
    for (int e = tid; e < (S - 2 * nc - na) * 64; e += 256) {
        const int q = e >> 6, k = e & 63;
        buf[k * S + 2 * nc + na + q] = 0.f;
    }
    __syncthreads();
    float* dst = rows + (long long)i0 * S;
    for (int e = tid; e < n * S; e += 256) dst[e] = buf[e];
}

void launch_rows(const float2* in, long long in_cstride, const float2* resid_y, const double2* ax64,
                 const float* Wd, int L, const float* AA, const float* AAi, int M, int nc, int S,
                 float2* rows, hipStream_t s)
{
    hipLaunchKernelGGL(k_spread_rows, dim3((unsigned)((M + 63) / 64)), dim3(256),
                       sizeof(float) * 64 * S, s, in, in_cstride, resid_y, ax64, Wd, L, AA, AAi, M,
                       nc, S, reinterpret_cast<float*>(rows));
}

// conj(a + i ai) v for the compact-row spreads, with the FMA contraction
// written out so that every kernel forms the same bits
__device__ __forceinline__ float2 row_conj_weight(float a, float ai, float2 v)
{
    return make_float2(__builtin_fmaf(a, v.x, ai * v.y), __builtin_fmaf(a, v.y, -(ai * v.x)));
}

// ------------------------------------------------- 2-D patch spread --------
// A block of W = kSpreadWaves wavefronts; lane = grid (coil x segment).
// Wavefront w runs the work item witems[W block + w] = {patch (-1: none),
// entry begin, entry end, slot}: it accumulates the 4x4 patch in registers
// over its entries.  Entry k carries its sample index and the 16 weights wx*wy of the
// patch's points (zero outside the sample's clamped window), precomputed
// with the trajectory.  Per stage of up to 64 entries a wavefront stages the
// weights in LDS (read back as broadcasts) and the indices one per lane
// (readlane); the only global load per entry is the sample's contiguous row
// of grid values, and the next 8 rows are in flight while the current 8 are
// summed.  A patch of 65..64 W entries is split over gs = 2..W consecutive
// wavefronts of a block (bgs[block] = gs, the same for all groups of a
// block), whose sums are added through LDS in a fixed tree order; the
// group's first wavefront writes the patch, or its partial slot when the
// patch is one of the parts of a heavier patch (added by
// k_spread_heavy_combine).  No atomics, no zero-fill, deterministic.
using pg::kSpreadWaves;

// ROWS = 1 (2: complex weights): the entries' values are formed from the
// compact rows of k_spread_rows (stride S floats, nc coils, L segments):
// lane b = c L + l reads r_c at row + 2c and AA_l at row + 2 nc + l, and
// forms AA_l r_c with spread_value's products (conj(AA) r for complex AA).
template <int ROWS>
__global__ void __launch_bounds__(64 * kSpreadWaves) k_spread_patch(
    const int4* __restrict__ witems, const int* __restrict__ bgs, const int* __restrict__ entries,
    const float4* __restrict__ ew, int npy, const float2* __restrict__ V, int B,
    float2* __restrict__ grids, long long G, int Gy, float2* __restrict__ partial, int n_slots,
    int npp, long long pstride, int L, int S)
{
    __shared__ float2 red[kSpreadWaves / 2][16][64];
    __shared__ float4 sw[kSpreadWaves][64][4];
    // wave-uniform (scalar) work item: its loop bounds live in SGPRs
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int4 it = witems[blockIdx.x * kSpreadWaves + wv];
    const int gs = bgs[blockIdx.x];
    const int b = blockIdx.y * 64 + lane;
    const bool active = b < B;
    const int bb = active ? b : B - 1;

    float2 acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = make_float2(0.f, 0.f);
    const float2* Vb = V + bb;
    // compact rows: this lane's coil and segment offsets within a row
    const int lc = ROWS ? bb / L : 0, ll = ROWS ? bb - lc * L : 0;
    const int nc = ROWS ? B / L : 0;
    const float* Rf = reinterpret_cast<const float*>(V);
    const int roff = 2 * lc, aoff = 2 * nc + ll;
    constexpr int U = 8;  // entries per group
    float4(*mw)[4] = sw[wv];
    for (int s0 = it.y; s0 < it.z; s0 += 64) {
        const int n = min(64, it.z - s0);
        // stage weights and indices; the weights of rows past n are zero (the
        // sums below read every row of a group of 8 and mask only the values,
        // so a stale row would otherwise enter as 0 * stale)
        const bool ok = lane < n;
        const int e = ok ? s0 + lane : s0;
        const int ei = entries[e];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float4 q = ew[4 * e + r];
            mw[lane][r] = ok ? q : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        pgf::wave_lds_sync();
        // rows of group k0; entries past n load the last entry's row, which
        // is zeroed with a bit mask where it is used (no branch around the
        // loads, so that the next group's loads stay in flight while this
        // group is summed)
        auto load_group = [&](int k0, float2 (&v)[U], float (&a)[U], float (&ai)[U]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = __builtin_amdgcn_readlane(ei, min(k0 + u, n - 1));
                if (ROWS) {
                    const float* row = Rf + (long long)i * S;
                    v[u] = *reinterpret_cast<const float2*>(row + roff);
                    a[u] = row[aoff];
                    if (ROWS == 2) ai[u] = row[aoff + L];
                } else {
                    v[u] = Vb[(long long)i * B];
                }
            }
        };
        // sum group k0 (rows in v) into the patch
        auto sum_group = [&](int k0, float2 (&v)[U], float (&a)[U], float (&ai)[U]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int m = (k0 + u < n) ? -1 : 0;
                float2 val = v[u];
                if (ROWS == 1) val = make_float2(a[u] * val.x, a[u] * val.y);
                if (ROWS == 2)
                    val = row_conj_weight(a[u], ai[u], val);
                const float2 vu = make_float2(__int_as_float(__float_as_int(val.x) & m),
                                              __int_as_float(__float_as_int(val.y) & m));
                const int k = min(k0 + u, 63);
                float w[16];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float4 q4 = mw[k][r];
                    w[4 * r] = q4.x;
                    w[4 * r + 1] = q4.y;
                    w[4 * r + 2] = q4.z;
                    w[4 * r + 3] = q4.w;
                }
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    acc[q].x += w[q] * vu.x;
                    acc[q].y += w[q] * vu.y;
                }
            }
        };
        // two row buffers in turn (the scheduling barriers keep each group's
        // loads ahead of the sums; the second group is summed
        // unconditionally, masked past n, so that its loads are not sunk next
        // to their use)
        float2 va[U], vb[U];
        float aa[U], ab[U], aia[U], aib[U];
        load_group(0, va, aa, aia);
        for (int k0 = 0; k0 < n; k0 += 2 * U) {
            load_group(k0 + U, vb, ab, aib);
            __builtin_amdgcn_sched_barrier(0);
            sum_group(k0, va, aa, aia);
            load_group(k0 + 2 * U, va, aa, aia);
            __builtin_amdgcn_sched_barrier(0);
            sum_group(k0 + U, vb, ab, aib);
        }
        pgf::wave_lds_sync();  // the weights of the next stage replace these
    }
    // groups of gs wavefronts: tree sum, wavefront w + r into w
    for (int r = 1; r < gs; r *= 2) {
        __syncthreads();
        if ((wv & (2 * r - 1)) == r)
#pragma unroll
            for (int q = 0; q < 16; ++q) red[wv >> 1][q][lane] = acc[q];
        __syncthreads();
        if ((wv & (2 * r - 1)) == 0)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float2 o = red[(wv + r) >> 1][q][lane];
                acc[q].x += o.x;
                acc[q].y += o.y;
            }
    }
    if (!active || it.x < 0 || (wv & (gs - 1)) != 0) return;
    if (it.w >= 0) {
        float2* pp = partial + ((long long)blockIdx.y * n_slots + it.w) * 16 * 64 + lane;
#pragma unroll
        for (int q = 0; q < 16; ++q) pp[q * 64] = acc[q];
        return;
    }
    // patch it.x = plane * npp + q (plane 0 for a 2-D plan)
    const int pln = it.x / npp, q0 = it.x - pln * npp;
    const int px0 = (q0 / npy) * pg::kPatch, py0 = (q0 % npy) * pg::kPatch;
    float2* g = grids + pln * pstride + (long long)b * G + (long long)px0 * Gy + py0;
#pragma unroll
    for (int a = 0; a < pg::kPatch; ++a)
#pragma unroll
        for (int q = 0; q < pg::kPatch; ++q) g[(long long)a * Gy + q] = acc[a * 4 + q];
}

// The same patch spread from compact rows with each stage's row slices
// staged in LDS.  A stage takes up to kStEntries entries; the wavefront's
// 64 grids b0 .. b0+63 use coils c_lo .. c_hi and all L segments, i.e. the
// float runs [2 c_lo, 2 c_hi + 2) and [2 nc, 2 nc + F2) of each entry's row
// (F2 = L, or 2L for complex weights, rounded up to even).  The lanes load
// those F = F1 + F2 floats of every entry of the stage as float2 units, all
// in flight together, and store them to LDS; the sums then read each
// entry's coil value and segment weight from LDS (a broadcast for the lanes
// of one coil) instead of from global memory, so a stage waits for memory
// once instead of once per group of entries.  Same values, products and
// summation order as k_spread_patch.  FM (32 or 48) bounds the floats staged
// per entry; the 32-float form keeps 48 KB of LDS and 80 VGPRs per block of
// 8 wavefronts, so six wavefronts per SIMD.  Requires F <= FM for every
// 64-grid chunk (checked by the launch).
constexpr int kStEntries = 32;
constexpr int kStFloats = 48;  // the largest row slice staged per entry

template <int ROWS, int FM>
__global__ void __launch_bounds__(64 * kSpreadWaves, FM <= 32 ? 6 : 4) k_spread_patch_st(
    const int4* __restrict__ witems, const int* __restrict__ bgs, const int* __restrict__ entries,
    const float4* __restrict__ ew, int npy, const float2* __restrict__ V, int B,
    float2* __restrict__ grids, long long G, int Gy, float2* __restrict__ partial, int n_slots,
    int npp, long long pstride, int L, int S)
{
    static_assert(ROWS == 1 || ROWS == 2, "compact rows only");
    constexpr int NE = kStEntries;
    constexpr int NU = (NE * FM / 2 + 63) / 64;  // float2 row loads per lane per stage
    // the stage rows of the block's wavefronts, reused for the tree sum after the
    // last stage (red: [kSpreadWaves / 2][16][64] float2)
    constexpr int POOL = kSpreadWaves * NE * FM > kSpreadWaves / 2 * 16 * 64 * 2
                             ? kSpreadWaves * NE * FM
                             : kSpreadWaves / 2 * 16 * 64 * 2;
    __shared__ __align__(16) float pool[POOL];
    __shared__ float4 sw[kSpreadWaves][NE][4];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // blocks (patch block, 64-grid chunk), patch blocks fastest: the blocks
    // that run together write different patches, spread over the memory
    // channels (the chunks of one patch block write 64-grid-apart rows of
    // the same addresses, which run together measured 40% slower)
    const int chunk = blockIdx.y;
    const int4 it = witems[blockIdx.x * kSpreadWaves + wv];
    const int gs = bgs[blockIdx.x];
    const int b0 = chunk * 64;
    const int b = b0 + lane;
    const bool active = b < B;
    const int bb = active ? b : B - 1;
    const int nc = B / L;
    const int lc = bb / L, ll = bb - lc * L;
    const int c_lo = b0 / L, c_hi = (min(b0 + 64, B) - 1) / L;
    const int F1 = 2 * (c_hi - c_lo + 1);
    const int F2 = ((ROWS == 2 ? 2 * L : L) + 1) & ~1;
    const int U2 = (F1 + F2) / 2;        // float2 units per entry
    const int F = 2 * U2;                // LDS floats per entry
    const int qd = 64 / U2, qr = 64 - qd * U2;
    float* rows = pool + wv * NE * FM;
    float4(*mw)[4] = sw[wv];
    const float* Rf = reinterpret_cast<const float*>(V);
    const int voff = 2 * (lc - c_lo), aoff = F1 + ll;

    float2 acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = make_float2(0.f, 0.f);
    for (int s0 = it.y; s0 < it.z; s0 += NE) {
        const int n = min(NE, it.z - s0);
        const bool ok = lane < n;
        const int e = ok ? s0 + lane : s0;
        const int ei = entries[e];
        // the stage's weights: lane k < n loads entry k's 64 bytes through a
        // buffer resource of n entries (other lanes read zeros), all four
        // loads issued before the rows' (no branch, no wait between them)
        float4 wq[4];
        {
            const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(ew + 4 * (long long)s0), (short)0, n * 64, kBufferWord3);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                wq[r] = __builtin_bit_cast(
                    float4, __builtin_amdgcn_raw_buffer_load_b128(rw, lane * 64 + r * 16, 0, 0));
        }
        // unit t = lane + 64 j of the stage: entry k = t / U2, unit u = t % U2
        float2 rv[NU];
        {
            // the rows through one buffer resource (32-bit byte offsets: the
            // launch keeps M S 4 below 2^31)
            const __amdgpu_buffer_rsrc_t rrow = __builtin_amdgcn_make_buffer_rsrc(
                (void*)Rf, (short)0, 0x7fffffff, kBufferWord3);
            int k = lane / U2, u = lane - (lane / U2) * U2;
            // formed anew per stage (kept out of registers across stages)
            asm volatile("" : "+v"(k), "+v"(u));
#pragma unroll
            for (int j = 0; j < NU; ++j) {
                const int src = __shfl(ei, min(k, n - 1), 64);
                const int off = u < U2 - F2 / 2 ? 2 * c_lo + 2 * u : 2 * nc + 2 * (u - (U2 - F2 / 2));
                rv[j] = __builtin_bit_cast(
                    float2, __builtin_amdgcn_raw_buffer_load_b64(rrow, (src * S + off) * 4, 0, 0));
                k += qd;
                u += qr;
                if (u >= U2) {
                    u -= U2;
                    ++k;
                }
            }
        }
        if (lane < NE) {
#pragma unroll
            for (int r = 0; r < 4; ++r) mw[lane][r] = wq[r];
        }
        // every unit t < 64 NU <= NE FM / 2 lies in the wavefront's rows;
        // those past n U2 are never read
        static_assert(64 * NU <= NE * FM / 2, "stage rows");
#pragma unroll
        for (int j = 0; j < NU; ++j)
            *reinterpret_cast<float2*>(rows + 2 * (lane + 64 * j)) = rv[j];
        pgf::wave_lds_sync();
#pragma unroll 1
        for (int k = 0; k < n; ++k) {
            const float* row = rows + k * F;
            float2 val = *reinterpret_cast<const float2*>(row + voff);
            const float a = row[aoff];
            if (ROWS == 1) val = make_float2(a * val.x, a * val.y);
            if (ROWS == 2) {
                const float ai = row[aoff + L];
                val = row_conj_weight(a, ai, val);
            }
            float w[16];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float4 q4 = mw[k][r];
                w[4 * r] = q4.x;
                w[4 * r + 1] = q4.y;
                w[4 * r + 2] = q4.z;
                w[4 * r + 3] = q4.w;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                acc[q].x += w[q] * val.x;
                acc[q].y += w[q] * val.y;
            }
        }
        pgf::wave_lds_sync();  // the next stage replaces these rows and weights
    }
    float2(*red)[16][64] = reinterpret_cast<float2(*)[16][64]>(pool);
    for (int r = 1; r < gs; r *= 2) {
        __syncthreads();
        if ((wv & (2 * r - 1)) == r)
#pragma unroll
            for (int q = 0; q < 16; ++q) red[wv >> 1][q][lane] = acc[q];
        __syncthreads();
        if ((wv & (2 * r - 1)) == 0)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float2 o = red[(wv + r) >> 1][q][lane];
                acc[q].x += o.x;
                acc[q].y += o.y;
            }
    }
    if (!active || it.x < 0 || (wv & (gs - 1)) != 0) return;
    if (it.w >= 0) {
        float2* pp = partial + ((long long)chunk * n_slots + it.w) * 16 * 64 + lane;
#pragma unroll
        for (int q = 0; q < 16; ++q) pp[q * 64] = acc[q];
        return;
    }
    const int pln = it.x / npp, q0 = it.x - pln * npp;
    const int px0 = (q0 / npy) * pg::kPatch, py0 = (q0 % npy) * pg::kPatch;
    // each patch row is 32 contiguous bytes at a 32-byte boundary (py0, Gy
    // and the plane and grid strides are multiples of 4 points): two
    // 16-byte stores per row
    float2* g = grids + pln * pstride + (long long)b * G + (long long)px0 * Gy + py0;
#pragma unroll
    for (int a = 0; a < pg::kPatch; ++a) {
        float4* g4 = reinterpret_cast<float4*>(g + (long long)a * Gy);
        g4[0] = make_float4(acc[a * 4].x, acc[a * 4].y, acc[a * 4 + 1].x, acc[a * 4 + 1].y);
        g4[1] = make_float4(acc[a * 4 + 2].x, acc[a * 4 + 2].y, acc[a * 4 + 3].x, acc[a * 4 + 3].y);
    }
}

// the largest staged floats per entry of a launch's 64-grid chunks
int spread_st_floats(int B, int L, bool cpx)
{
    int fm = 0;
    for (int b0 = 0; b0 < B; b0 += 64) {
        const int f1 = 2 * ((std::min(b0 + 64, B) - 1) / L - b0 / L + 1);
        fm = std::max(fm, f1 + (((cpx ? 2 * L : L) + 1) & ~1));
    }
    return fm;
}

// The same work list on the matrix cores (B a multiple of 4).  For a chunk of
// 4 entries the patch update is a product
//   acc[16 points][128 floats] += W[16 points][4 entries] * V[4 entries][128 floats]
// of the entries' weights and their rows of grid values (re and im as
// separate columns), done as 8 v_mfma_f32_16x16x4f32, one per 16-column
// tile.  Tile t, column j holds row float 8j + t, so that each lane's share
// of a row is 8 consecutive floats (two 16-byte loads).  Lane layouts of the
// 16x16x4 f32 product (one block): A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, D[i][j] in lane j + 16(i/4), register i % 4.
typedef float pg_f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(64 * kSpreadWaves) k_spread_patch_mfma(
    const int4* __restrict__ witems, const int* __restrict__ bgs, const int* __restrict__ entries,
    const float4* __restrict__ ew, int npy, const float2* __restrict__ V, int B,
    float2* __restrict__ grids, long long G, int Gy, float2* __restrict__ partial, int n_slots,
    int npp, long long pstride)
{
    __shared__ float red[kSpreadWaves / 2][32][64];
    __shared__ float sw[kSpreadWaves][64][16];
    __shared__ int si[kSpreadWaves][64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int4 it = witems[blockIdx.x * kSpreadWaves + wv];
    const int gs = bgs[blockIdx.x];
    const int cw = min(64, B - 64 * (int)blockIdx.y);  // grids in this chunk (multiple of 4)
    const int kq = lane >> 4, jq = lane & 15;          // MFMA k (entry) and j (column) of this lane
    const bool colok = 4 * jq < cw;                     // this lane's 4 grids exist
    // this lane's 8 floats of a row: grids 4jq .. 4jq+3 of the chunk
    const float* Vc = reinterpret_cast<const float*>(V) + 2 * 64 * (long long)blockIdx.y +
                      8 * (colok ? jq : 0);
    const long long rowf = 2LL * B;  // floats per sample row

    pg_f4 acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = pg_f4{0.f, 0.f, 0.f, 0.f};
    float(*mw)[16] = sw[wv];
    int* mi = si[wv];
    for (int s0 = it.y; s0 < it.z; s0 += 64) {
        const int n = min(64, it.z - s0);
        {
            // stage weights and indices; entries past n get zero weights
            const bool ok = lane < n;
            const int e = ok ? s0 + lane : s0;
            mi[lane] = entries[e];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float4 q = ew[4 * e + r];
                mw[lane][4 * r] = ok ? q.x : 0.f;
                mw[lane][4 * r + 1] = ok ? q.y : 0.f;
                mw[lane][4 * r + 2] = ok ? q.z : 0.f;
                mw[lane][4 * r + 3] = ok ? q.w : 0.f;
            }
        }
        pgf::wave_lds_sync();
        // rows of chunk k0 for this lane (entry k0 + kq; past n: zeroed)
        auto load_chunk = [&](int k0, float4 (&v)[2]) {
            const int k = min(k0 + kq, 63);
            const float4* src = reinterpret_cast<const float4*>(Vc + rowf * mi[k]);
            v[0] = src[0];
            v[1] = src[1];
        };
        auto mma_chunk = [&](int k0, float4 (&v)[2]) {
            const int m = (k0 + kq < n && colok) ? -1 : 0;
            const float a = mw[min(k0 + kq, 63)][jq];  // A[point jq][entry kq]
            float b[8] = {v[0].x, v[0].y, v[0].z, v[0].w, v[1].x, v[1].y, v[1].z, v[1].w};
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const float bt = __int_as_float(__float_as_int(b[t]) & m);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bt, acc[t], 0, 0, 0);
            }
        };
        float4 va[2], vb[2];
        load_chunk(0, va);
        for (int k0 = 0; k0 < n; k0 += 8) {
            load_chunk(k0 + 4, vb);
            __builtin_amdgcn_sched_barrier(0);
            mma_chunk(k0, va);
            load_chunk(k0 + 8, va);
            __builtin_amdgcn_sched_barrier(0);
            mma_chunk(k0 + 4, vb);
        }
        pgf::wave_lds_sync();  // the next stage's weights replace these
    }
    // groups of gs wavefronts: tree sum, wavefront w + r into w
    for (int r = 1; r < gs; r *= 2) {
        __syncthreads();
        if ((wv & (2 * r - 1)) == r)
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) red[wv >> 1][4 * t + q][lane] = acc[t][q];
        __syncthreads();
        if ((wv & (2 * r - 1)) == 0)
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[t][q] += red[(wv + r) >> 1][4 * t + q][lane];
    }
    if (it.x < 0 || (wv & (gs - 1)) != 0 || !colok) return;
    // lane holds points p = 4 kq + q and grids b = 4 jq + u (re: tile 2u, im: tile 2u+1)
    const int b0 = 64 * blockIdx.y + 4 * jq;
    if (it.w >= 0) {
        float2* pp = partial + ((long long)blockIdx.y * n_slots + it.w) * 16 * 64;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                pp[(4 * kq + q) * 64 + 4 * jq + u] = make_float2(acc[2 * u][q], acc[2 * u + 1][q]);
        return;
    }
    const int pln = it.x / npp, q0 = it.x - pln * npp;
    const int px0 = (q0 / npy) * pg::kPatch, py0 = (q0 % npy) * pg::kPatch;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int p = 4 * kq + q;
        const long long off = pln * pstride + (long long)(px0 + p / pg::kPatch) * Gy + py0 +
                              p % pg::kPatch;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            grids[(long long)(b0 + u) * G + off] = make_float2(acc[2 * u][q], acc[2 * u + 1][q]);
    }
}

// The matrix-core patch spread from compact rows: k_spread_patch_mfma's
// products (8 v_mfma_f32_16x16x4f32 per chunk of 4 entries, the same lane
// layouts) with the B operand formed from the row slices that
// k_spread_patch_st stages in LDS.  Lane (kq, jq) supplies, for entry
// k0 + kq, the re and im parts of grids b0 + 4 jq + u (u < 4): the value
// AA_l r_c (conj(AA_l) r_c for complex weights) of coil c = b / L, segment
// l = b % L, the products of k_spread_values, so the operands, and with the
// same chunks of 4 entries the sums, are those of the value-row kernel.
// B a multiple of 4; F <= kStFloats for every 64-grid chunk.
template <int ROWS>
__global__ void __launch_bounds__(64 * kSpreadWaves) k_spread_patch_mfma_st(
    const int4* __restrict__ witems, const int* __restrict__ bgs, const int* __restrict__ entries,
    const float4* __restrict__ ew, int npy, const float2* __restrict__ V, int B,
    float2* __restrict__ grids, long long G, int Gy, float2* __restrict__ partial, int n_slots,
    int npp, long long pstride, int L, int S)
{
    static_assert(ROWS == 1 || ROWS == 2, "compact rows only");
    constexpr int NE = kStEntries, FM = kStFloats;
    constexpr int NU = (NE * FM / 2 + 63) / 64;
    constexpr int POOL = kSpreadWaves * NE * FM > kSpreadWaves / 2 * 32 * 64
                             ? kSpreadWaves * NE * FM
                             : kSpreadWaves / 2 * 32 * 64;
    __shared__ __align__(16) float pool[POOL];
    __shared__ float sw[kSpreadWaves][NE][16];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int4 it = witems[blockIdx.x * kSpreadWaves + wv];
    const int gs = bgs[blockIdx.x];
    const int b0 = blockIdx.y * 64;
    const int cw = min(64, B - b0);                     // grids in this chunk (multiple of 4)
    const int kq = lane >> 4, jq = lane & 15;          // MFMA k (entry) and j (column) of this lane
    const bool colok = 4 * jq < cw;                     // this lane's 4 grids exist
    const int nc = B / L;
    const int c_lo = b0 / L, c_hi = (b0 + cw - 1) / L;
    const int F1 = 2 * (c_hi - c_lo + 1);
    const int F2 = ((ROWS == 2 ? 2 * L : L) + 1) & ~1;
    const int U2 = (F1 + F2) / 2;
    const int F = 2 * U2;
    const int qd = 64 / U2, qr = 64 - qd * U2;
    float* rows = pool + wv * NE * FM;
    float(*mw)[16] = sw[wv];
    const float* Rf = reinterpret_cast<const float*>(V);
    // this lane's four grids: coil value and segment weight offsets in a
    // staged row
    int vo[4], ao[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int bq = min(b0 + 4 * jq + u, B - 1);
        const int cq = bq / L, lq = bq - cq * L;
        vo[u] = 2 * (cq - c_lo);
        ao[u] = F1 + lq;
    }

    pg_f4 acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = pg_f4{0.f, 0.f, 0.f, 0.f};
    for (int s0 = it.y; s0 < it.z; s0 += NE) {
        const int n = min(NE, it.z - s0);
        const bool ok = lane < n;
        const int e = ok ? s0 + lane : s0;
        const int ei = entries[e];
        float2 rv[NU];
        {
            int k = lane / U2, u = lane - (lane / U2) * U2;
            asm volatile("" : "+v"(k), "+v"(u));
#pragma unroll
            for (int j = 0; j < NU; ++j) {
                const int src = __shfl(ei, min(k, n - 1), 64);
                const int off = u < U2 - F2 / 2 ? 2 * c_lo + 2 * u : 2 * nc + 2 * (u - (U2 - F2 / 2));
                rv[j] = *reinterpret_cast<const float2*>(Rf + (long long)src * S + off);
                k += qd;
                u += qr;
                if (u >= U2) {
                    u -= U2;
                    ++k;
                }
            }
        }
        if (lane < NE) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float4 q = ew[4 * e + r];
                mw[lane][4 * r] = ok ? q.x : 0.f;
                mw[lane][4 * r + 1] = ok ? q.y : 0.f;
                mw[lane][4 * r + 2] = ok ? q.z : 0.f;
                mw[lane][4 * r + 3] = ok ? q.w : 0.f;
            }
        }
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const int t = lane + 64 * j;
            if (t < n * U2) *reinterpret_cast<float2*>(rows + 2 * t) = rv[j];
        }
        pgf::wave_lds_sync();
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int ke = k0 + kq;
            const int m = (ke < n && colok) ? -1 : 0;
            const float a = mw[min(ke, NE - 1)][jq];  // A[point jq][entry kq]
            const float* row = rows + min(ke, n - 1) * F;
            float b[8];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float2 val = *reinterpret_cast<const float2*>(row + vo[u]);
                const float w = row[ao[u]];
                if (ROWS == 1) val = make_float2(w * val.x, w * val.y);
                if (ROWS == 2) val = row_conj_weight(w, row[ao[u] + L], val);
                b[2 * u] = __int_as_float(__float_as_int(val.x) & m);
                b[2 * u + 1] = __int_as_float(__float_as_int(val.y) & m);
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[t], acc[t], 0, 0, 0);
        }
        pgf::wave_lds_sync();  // the next stage replaces these rows and weights
    }
    float(*red)[32][64] = reinterpret_cast<float(*)[32][64]>(pool);
    for (int r = 1; r < gs; r *= 2) {
        __syncthreads();
        if ((wv & (2 * r - 1)) == r)
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) red[wv >> 1][4 * t + q][lane] = acc[t][q];
        __syncthreads();
        if ((wv & (2 * r - 1)) == 0)
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[t][q] += red[(wv + r) >> 1][4 * t + q][lane];
    }
    if (it.x < 0 || (wv & (gs - 1)) != 0 || !colok) return;
    const int bl = b0 + 4 * jq;
    if (it.w >= 0) {
        float2* pp = partial + ((long long)blockIdx.y * n_slots + it.w) * 16 * 64;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                pp[(4 * kq + q) * 64 + 4 * jq + u] = make_float2(acc[2 * u][q], acc[2 * u + 1][q]);
        return;
    }
    const int pln = it.x / npp, q0 = it.x - pln * npp;
    const int px0 = (q0 / npy) * pg::kPatch, py0 = (q0 % npy) * pg::kPatch;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int p = 4 * kq + q;
        const long long off = pln * pstride + (long long)(px0 + p / pg::kPatch) * Gy + py0 +
                              p % pg::kPatch;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            grids[(long long)(bl + u) * G + off] = make_float2(acc[2 * u][q], acc[2 * u + 1][q]);
    }
}

// patches summed in parts: add the parts' slots in slot order; block
// (patch, grid chunk, point)
__global__ void __launch_bounds__(64) k_spread_heavy_combine(const int4* __restrict__ comb, int npy,
                                                             const float2* __restrict__ partial,
                                                             int n_slots, int B,
                                                             float2* __restrict__ grids,
                                                             long long G, int Gy, int npp,
                                                             long long pstride)
{
    const int4 cb = comb[blockIdx.x];
    const int q = blockIdx.z;
    const int b = blockIdx.y * 64 + threadIdx.x;
    if (b >= B) return;
    const float2* pp = partial + ((long long)blockIdx.y * n_slots + cb.y) * 16 * 64 + q * 64 +
                       threadIdx.x;
    float2 t = pp[0];
    for (int k = 1; k < cb.z; ++k) {
        const float2 v = pp[(long long)k * 16 * 64];
        t.x += v.x;
        t.y += v.y;
    }
    const int pln = cb.x / npp, q0 = cb.x - pln * npp;
    const int px0 = (q0 / npy) * pg::kPatch, py0 = (q0 % npy) * pg::kPatch;
    grids[pln * pstride + (long long)b * G + (long long)(px0 + q / pg::kPatch) * Gy + py0 +
          q % pg::kPatch] = t;
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
    const float* __restrict__ AA, const float* __restrict__ AAi, int B, float2* __restrict__ grids,
    long long G)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= M) return;
    const float2 v = spread_value(in, in_cstride, resid_y, ax64, Wd, AA, AAi, L, M, b, i);
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
                     const float* AAi, float2* out, long long out_cstride, hipStream_t s,
                     bool planes)
{
    const pg::Geometry& g = G.geo;
    const pg::Gnufft::Tuned* tn = G.tuned_for(nc * L);
    // z-class planes: a 2-D interpolation on the plane of each sample's class
    const long long gsz = planes ? (long long)g.Gx * g.Gy : g.G();
    const long long pstride = planes ? (long long)nc * L * gsz : 0;
    if ((!D3 || planes) && NT <= 8 && pgk::interp_lds_possible(G, L, AA != nullptr) &&
        (!tn || tn->lds)) {
        // four segments per stage when the autotuner chose it, or else from
        // four segments up (measured faster there: the smaller blocks and
        // LDS footprint keep more of them per CU)
        const int v4 = pgk::g_interp4.load(std::memory_order_relaxed);
        const bool four = tn ? tn->lds == 2 : (v4 == 2 || (v4 == 1 && L >= 4));
        // (its staging and weight offsets are 32-bit byte offsets)
        if (G.n_iitems4 > 0 && four && gsz * 4 * 8 < (1LL << 31) &&
            (long long)L * G.M * 4 < (1LL << 31)) {
            const dim3 grid((unsigned)G.n_iitems4 * nc), block(kI4Threads);
            auto launch = [&](auto kern) {
                hipLaunchKernelGGL(kern, grid, block, 0, s, G.iitems4, G.irec4, G.ity, G.ipadl,
                                   grids, gsz, g.Gx, g.Gy, nc, L, AA, AAi, G.M, out, out_cstride,
                                   pstride);
            };
            AAi ? launch(k_interp_lds4<true>) : launch(k_interp_lds4<false>);
            return;
        }
        const dim3 grid((unsigned)G.n_iitems * nc), block(64 * kInterpWaves);
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, block, 0, s, G.iitems, G.irec, G.ity, G.ipadl, grids,
                               gsz, g.Gx, g.Gy, nc, L, AA, AAi, G.M, out, out_cstride, pstride);
        };
        if (L <= 8)
            AAi ? launch(k_interp_lds<8, true>) : launch(k_interp_lds<8, false>);
        else
            AAi ? launch(k_interp_lds<16, true>) : launch(k_interp_lds<16, false>);
        return;
    }
    const int nsb = (G.M + 63) / 64;  // 64 samples (4 lanes each) per block
    const int xmap = (nc % 8 == 0) ? 1 : 0;
    if (planes) {
        hipLaunchKernelGGL((k_interp<NT, false>), dim3(nsb * nc), dim3(256), 0, s, G.lo, G.w, G.M,
                           g.Gx, g.Gy, 1, grids, gsz, nc, L, AA, AAi, out, out_cstride, nsb, xmap,
                           (const int*)G.zcls, pstride);
        return;
    }
    hipLaunchKernelGGL((k_interp<NT, D3>), dim3(nsb * nc), dim3(256), 0, s, G.lo, G.w, G.M, g.Gx,
                       g.Gy, g.Gz, grids, g.G(), nc, L, AA, AAi, out, out_cstride, nsb, xmap,
                       (const int*)nullptr, 0LL);
}

template <int NT, bool D3>
void spread_dispatch(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc, int L,
                     const float* AA, const float* AAi, const float2* resid_y, const double2* ax64,
                     const float* Wd, float2* grids, bool vals_ready, hipStream_t s, bool planes,
                     float2* vals, float2* partial)
{
    const pg::Geometry& g = G.geo;
    const int B = nc * L;
    if (D3 && !planes && B < 16) {
        const long long n = (long long)B * g.G();
        hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, grids, n);
        hipLaunchKernelGGL((k_spread_atomic<NT, D3>), dim3((G.M + 255) / 256, B), dim3(256), 0, s,
                           G.lo, G.w, G.M, g.Gx, g.Gy, g.Gz, in, in_cstride, resid_y, ax64, Wd, L,
                           AA, AAi, B, grids, g.G());
        return;
    }
    // per-sample values of every grid (or the compact rows they are formed
    // from), once
    const int S = pgk::spread_row_floats(G, nc, L, AA, AAi);
    if (!vals_ready) {
        if (S)
            launch_rows(in, in_cstride, resid_y, ax64, Wd, L, AA, AAi, G.M, nc, S, vals, s);
        else
            launch_values(in, in_cstride, resid_y, ax64, Wd, L, AA, AAi, G.M, B, vals, s);
    }
    if (!D3 || planes) {
        // 2-D patches (of every z-class plane); plane p of grid b at
        // (p*B + b)*Gx*Gy
        const long long gsz = planes ? (long long)g.Gx * g.Gy : g.G();
        const long long pstride = planes ? (long long)B * gsz : 0;
        const int nch = (B + 63) / 64;
        const int npy = g.Gy / pg::kPatch;
        const pg::Gnufft::Tuned* tn = G.tuned_for(B);
        const bool mfma =
            B % 4 == 0 &&
            (tn ? tn->mfma != 0 : pgk::g_spread_mfma.load(std::memory_order_relaxed) != 0);
        if (mfma && S) {
            // compact rows reach the matrix cores only through the staged form
            auto launch = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(G.n_sblocks, nch), dim3(64 * kSpreadWaves), 0, s,
                                   G.witems, G.bgs, G.entries, (const float4*)G.ent_w, npy, vals,
                                   B, grids, gsz, g.Gy, partial, G.n_slots, G.npp, pstride, L, S);
            };
            AAi ? launch(k_spread_patch_mfma_st<2>) : launch(k_spread_patch_mfma_st<1>);
        } else if (mfma)
            hipLaunchKernelGGL(k_spread_patch_mfma, dim3(G.n_sblocks, nch), dim3(64 * kSpreadWaves),
                               0, s, G.witems, G.bgs, G.entries, (const float4*)G.ent_w, npy, vals,
                               B, grids, gsz, g.Gy, partial, G.n_slots, G.npp, pstride);
        else {
            auto launch = [&](auto kern, int L_, int S_) {
                hipLaunchKernelGGL(kern, dim3(G.n_sblocks, nch), dim3(64 * kSpreadWaves), 0, s,
                                   G.witems, G.bgs, G.entries, (const float4*)G.ent_w, npy, vals,
                                   B, grids, gsz, g.Gy, partial, G.n_slots, G.npp, pstride, L_, S_);
            };
            const bool st = S && pgk::g_spread_stage.load(std::memory_order_relaxed) != 0 &&
                            spread_st_floats(B, L, AAi != nullptr) <= kStFloats &&
                            (long long)G.M * S * 4 < (1LL << 31);
            auto launch_st = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(G.n_sblocks, nch), dim3(64 * kSpreadWaves), 0, s,
                                   G.witems, G.bgs, G.entries, (const float4*)G.ent_w, npy, vals,
                                   B, grids, gsz, g.Gy, partial, G.n_slots, G.npp, pstride, L, S);
            };
            const bool small = spread_st_floats(B, L, AAi != nullptr) <= 32;
            if (st && AAi)
                small ? launch_st(k_spread_patch_st<2, 32>) : launch_st(k_spread_patch_st<2, 48>);
            else if (st)
                small ? launch_st(k_spread_patch_st<1, 32>) : launch_st(k_spread_patch_st<1, 48>);
            else if (S && AAi)
                launch(k_spread_patch<2>, L, S);
            else if (S)
                launch(k_spread_patch<1>, L, S);
            else
                launch(k_spread_patch<0>, L, 0);
        }
        if (G.n_comb > 0)
            hipLaunchKernelGGL(k_spread_heavy_combine, dim3(G.n_comb, nch, 16), dim3(64), 0, s, G.comb,
                               npy, partial, G.n_slots, B, grids, gsz, g.Gy, G.npp, pstride);
        return;
    }
    PG_REQUIRE(G.tile_off && G.tile_idx, PG_ERR_STATE,
               "plan built for z-class planes: no 3-D tile lists for a batch of this size");
    constexpr int BC = 16;
    const int ntiles = G.ntx * G.nty * G.ntz;
    hipLaunchKernelGGL((k_spread_tile3<NT, BC>), dim3(ntiles, (B + BC - 1) / BC), dim3(256), 0, s,
                       G.lo, G.w, G.M, g.Gx, g.Gy, g.Gz, G.tile_off, G.tile_idx, G.ntx, G.nty,
                       vals, B, grids, g.G());
}

}  // namespace

namespace pgk {

std::atomic<int> g_spread_mfma{0};
std::atomic<int> g_zclass{1};
std::atomic<int> g_interp4{1};
std::atomic<int> g_spread_rows{1};
std::atomic<int> g_zfuse{1};
std::atomic<int> g_zmfma{1};
std::atomic<int> g_zamfma{1};
std::atomic<int> g_zfwd{1};
std::atomic<int> g_spread_stage{1};

int spread_row_floats(const pg::Gnufft& G, int nc, int L, const float* AA, const float* AAi)
{
    const int B = nc * L;
    if (!g_spread_rows.load(std::memory_order_relaxed) || !AA || L < 2) return 0;
    if (G.geo.dims == 3 && !zclass_active(G, B)) return 0;
    const pg::Gnufft::Tuned* tn = G.tuned_for(B);
    // the matrix-core spread reads compact rows only in its staged form
    if (B % 4 == 0 && (tn ? tn->mfma != 0 : g_spread_mfma.load(std::memory_order_relaxed) != 0) &&
        !(g_spread_stage.load(std::memory_order_relaxed) &&
          spread_st_floats(B, L, AAi != nullptr) <= kStFloats))
        return 0;
    const int S = (2 * nc + (AAi ? 2 : 1) * L + 3) & ~3;
    // the rows live in the plan's value buffer (M x B float2) and a block of
    // 64 rows in LDS
    return (S <= 2 * B && S <= 128) ? S : 0;
}

bool zclass_active(const pg::Gnufft& G, int B)
{
    return G.nzc > 0 && B >= 16 && (long long)B * G.nzc <= 65535;
}

bool interp_lds_possible(const pg::Gnufft& G, int L, bool aa)
{
    const pg::Geometry& g = G.geo;
    return (g.dims == 2 || G.nzc > 0) && g.ntaps <= 8 && aa && L >= 2 && L <= 16 &&
           G.n_iitems > 0 && (long long)L * g.Gx * g.Gy < (1LL << 31);
}

void launch_interp(const pg::Gnufft& G, const float2* grids, int nc, int L, const float* AA,
                   const float* AAi, float2* out, long long out_cstride, hipStream_t s, bool planes)
{
    PG_REQUIRE(!planes || zclass_active(G, nc * L), PG_ERR_STATE, "no z-class planes for this batch");
    const bool d3 = G.geo.dims == 3;
    switch (G.geo.ntaps) {
    case 8:
        d3 ? interp_dispatch<8, true>(G, grids, nc, L, AA, AAi, out, out_cstride, s, planes)
           : interp_dispatch<8, false>(G, grids, nc, L, AA, AAi, out, out_cstride, s, planes);
        break;
    case 6:
        d3 ? interp_dispatch<6, true>(G, grids, nc, L, AA, AAi, out, out_cstride, s, planes)
           : interp_dispatch<6, false>(G, grids, nc, L, AA, AAi, out, out_cstride, s, planes);
        break;
    case 12:
        d3 ? interp_dispatch<12, true>(G, grids, nc, L, AA, AAi, out, out_cstride, s, planes)
           : interp_dispatch<12, false>(G, grids, nc, L, AA, AAi, out, out_cstride, s, planes);
        break;
    default:
        throw pg::Error{PG_ERR_UNSUPPORTED, "kernel width * gridos must give 6, 8 or 12 taps"};
    }
}

bool spread_uses_values(const pg::Gnufft& G, int B) { return !(G.geo.dims == 3 && B < 16); }

void launch_spread_values(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc,
                          int L, const float* AA, const float* AAi, const float2* resid_y,
                          const double2* ax64, const float* Wd, hipStream_t s, float2* vals)
{
    const int S = spread_row_floats(G, nc, L, AA, AAi);
    if (S)
        launch_rows(in, in_cstride, resid_y, ax64, Wd, L, AA, AAi, G.M, nc, S, vals ? vals : G.vals,
                    s);
    else
        launch_values(in, in_cstride, resid_y, ax64, Wd, L, AA, AAi, G.M, nc * L,
                      vals ? vals : G.vals, s);
}

void launch_spread(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc, int L,
                   const float* AA, const float* AAi, const float2* resid_y, const double2* ax64,
                   const float* Wd, float2* grids, bool vals_ready, hipStream_t s, bool planes,
                   float2* vals, float2* partial)
{
    PG_REQUIRE(!planes || zclass_active(G, nc * L), PG_ERR_STATE, "no z-class planes for this batch");
    if (!vals) vals = G.vals;
    if (!partial) partial = G.partial;
    const bool d3 = G.geo.dims == 3;
    switch (G.geo.ntaps) {
    case 8:
        d3 ? spread_dispatch<8, true>(G, in, in_cstride, nc, L, AA, AAi, resid_y, ax64, Wd,
                                      grids, vals_ready, s, planes, vals, partial)
           : spread_dispatch<8, false>(G, in, in_cstride, nc, L, AA, AAi, resid_y, ax64, Wd,
                                       grids, vals_ready, s, planes, vals, partial);
        break;
    case 6:
        d3 ? spread_dispatch<6, true>(G, in, in_cstride, nc, L, AA, AAi, resid_y, ax64, Wd,
                                      grids, vals_ready, s, planes, vals, partial)
           : spread_dispatch<6, false>(G, in, in_cstride, nc, L, AA, AAi, resid_y, ax64, Wd,
                                       grids, vals_ready, s, planes, vals, partial);
        break;
    case 12:
        d3 ? spread_dispatch<12, true>(G, in, in_cstride, nc, L, AA, AAi, resid_y, ax64, Wd,
                                       grids, vals_ready, s, planes, vals, partial)
           : spread_dispatch<12, false>(G, in, in_cstride, nc, L, AA, AAi, resid_y, ax64, Wd,
                                        grids, vals_ready, s, planes, vals, partial);
        break;
    default:
        throw pg::Error{PG_ERR_UNSUPPORTED, "kernel width * gridos must give 6, 8 or 12 taps"};
    }
}

}  // namespace pgk
