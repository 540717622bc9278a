// pg_gnufft.h -- device-resident state of one Gnufft (a trajectory on one
// geometry): the setup of Gnufft.cpp:34-122 done once, on the device.
#pragma once

#include "pg_internal.h"

namespace pg {

struct Gnufft {
    Geometry geo;
    int M = 0;
    // trajectory (device)
    float* kx = nullptr;
    float* ky = nullptr;
    float* kz = nullptr;
    // KB look-up table (griddingSupport.cpp:52-76), size S = 10000*W
    float* lut = nullptr;
    int S = 0;
    // per-axis KB transforms psi(u) (griddingSupport.cpp:246-260); the
    // deapodization factor is dfx[x]*dfy[y](*dfz[z]), divided out as the
    // reference does
    float* dfx = nullptr;
    float* dfy = nullptr;
    float* dfz = nullptr;
    float inv_osd = 0.25f;  // 1/os^d
    // FFT plans for the three grid lengths
    FftPlan1D px, py, pz;
    // spread tiles: samples whose (clamped) window intersects each tile
    int ntx = 0, nty = 0, ntz = 0;
    int* tile_off = nullptr;  // ntiles + 1
    int* tile_idx = nullptr;  // concatenated sample indices
    long long tile_entries = 0;
    int max_tile = 0;
    // workspace for batched transforms (grown on demand)
    int cap_batch = 0;
    float2* grids = nullptr;  // cap_batch * G
    float2* tbuf = nullptr;   // cap_batch * Nx*Nz*Gy (and Nx*Gy*Nz rows)
    float2* tbuf2 = nullptr;  // 3-D only: cap_batch * Nz*Gx*Gy
    // scratch for the one-NUFFT-at-a-time (plain) path and for composites
    float2* img_tmp = nullptr;  // N
    float2* img_tmp2 = nullptr; // N
    float2* dat_tmp = nullptr;  // M
    float2* dat_tmp2 = nullptr; // M

    void ensure_batch(int B);
    void release();
};

}  // namespace pg
