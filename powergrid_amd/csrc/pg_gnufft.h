// pg_gnufft.h -- device-resident state of one Gnufft (a trajectory on one
// geometry): the setup of Gnufft.cpp:34-122 done once, on the device.
//
// The trajectory is fixed for the life of a plan, so every Kaiser-Bessel tap
// weight is computed once, on the host, in double precision (window start,
// distance and LUT index of gridding.cpp:64-97 evaluated in FP64) and stored
// as FP32.  The kernels only load weights; all gridding arithmetic and
// accumulation stays FP32.
#pragma once

#include "pg_internal.h"

namespace pg {

constexpr int kPatch = 4;  // 2-D spread patch edge (grid points)

struct Gnufft {
    Geometry geo;
    int M = 0;
    // window starts (unclamped) per dimension: lo[d*M + i]
    int* lo = nullptr;
    // tap weights, structure-of-arrays: w[(d*NT + j)*M + i] (0 for taps that
    // fall outside the clamped window)
    float* w = nullptr;
    // 2-D spread: weights padded to 16 entries per sample and dimension,
    // wpad[(d*M + i)*16 + 3 + j] = weight of tap j (zeros around), so that a
    // 4-point patch reads 4 consecutive entries at offset (patch - lo + 3)
    float* wpad = nullptr;
    // per-axis KB transforms psi(u) (griddingSupport.cpp:246-260); the
    // deapodization divides by dfx[x]*dfy[y](*dfz[z])
    float* dfx = nullptr;
    float* dfy = nullptr;
    float* dfz = nullptr;
    float inv_osd = 0.25f;  // 1/os^d
    // FFT plans for the three grid lengths
    FftPlan1D px, py, pz;
    // 2-D batched spread: work items over 4x4 patches
    //   items[k] = {patch, entry begin, entry end, partial slot (-1: direct)}
    int4* items = nullptr;
    int n_items = 0;
    int* entries = nullptr;  // sample indices, grouped by patch
    int4* combine = nullptr; // {patch, first slot, slot count, 0} of split patches
    int n_combine = 0;
    int n_slots = 0;
    // 3-D spread (and B < 16): point-stationary tiles
    int ntx = 0, nty = 0, ntz = 0;
    int* tile_off = nullptr;  // ntiles + 1
    int* tile_idx = nullptr;  // concatenated sample indices
    long long tile_entries = 0;
    // workspace for batched transforms (grown on demand)
    int cap_batch = 0;
    float2* grids = nullptr;    // cap_batch * G
    float2* tbuf = nullptr;     // cap_batch * Nx*Nz*Gy
    float2* tbuf2 = nullptr;    // 3-D only: cap_batch * Nz*Gx*Gy
    float2* partial = nullptr;  // ceil(cap_batch/64) * n_slots * 16 * 64
    float2* vals = nullptr;     // M * cap_batch per-sample spread values
    float2* img_tmp = nullptr;
    float2* dat_tmp = nullptr;

    void ensure_batch(int B);
    void release();
};

}  // namespace pg
