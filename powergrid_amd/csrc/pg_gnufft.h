// pg_gnufft.h -- device-resident state of one Gnufft (a trajectory on one
// geometry): the setup of Gnufft.cpp:34-122 done once, on the device.
//
// The trajectory is fixed for the life of a plan, so every Kaiser-Bessel tap
// weight is computed once, on the host, in double precision (window start,
// distance and LUT index of gridding.cpp:64-97 evaluated in FP64) and stored
// as FP32.  The kernels only load weights; all gridding arithmetic and
// accumulation stays FP32.
#pragma once

#include <vector>

#include "pg_internal.h"

namespace pg {

constexpr int kPatch = 4;  // 2-D spread patch edge (grid points)
// wavefronts (work items) per block of the 2-D patch spread, a power of two:
// a patch of up to 64 kSpreadWaves entries is summed by up to that many
// wavefronts of one block, a heavier one in parts through partial slots
constexpr int kSpreadWaves = 8;
constexpr int kIT = 16;    // 2-D interp: tile edge in window starts
constexpr int kIR = kIT + 7;  // staged grid region edge (8 taps)
constexpr int kIChunk = 128;  // samples per interp work item
constexpr int kIChunk4 = 64;  // the same for the four-segment staging
// row pitches (in points) that an item of the four-segment staging may take
// for its kIR x kIR region: kIR .. kIR + kI4Pitches - 1
constexpr int kI4Pitches = 3;

struct Gnufft {
    Geometry geo;
    int M = 0;
    // window starts (unclamped) per dimension: lo[d*M + i]
    int* lo = nullptr;
    // tap weights, structure-of-arrays: w[(d*NT + j)*M + i] (0 for taps that
    // fall outside the clamped window)
    float* w = nullptr;
    // per-axis KB transforms psi(u) (griddingSupport.cpp:246-260); the
    // deapodization divides by dfx[x]*dfy[y](*dfz[z])
    float* dfx = nullptr;
    float* dfy = nullptr;
    float* dfz = nullptr;
    float inv_osd = 0.25f;  // 1/os^d
    // full deapodization factor per image voxel, computed in FP64:
    // dap[n] = 1/(os^d psi_x psi_y psi_z), 0 where psi is NaN
    float* dap = nullptr;
    // FFT plans for the three grid lengths; fft_ok is false when a grid
    // length has a prime factor above 13 (only the spatial interpolation is available)
    FftPlan1D px, py, pz;
    bool fft_ok = false;
    // 2-D batched spread over 4x4 patches: blocks of W = kSpreadWaves
    // wavefront items
    //   witems[W*k + w] = {patch or -1, entry begin, entry end, slot or -1}
    // and the per-block group size bgs[k] (wavefronts sharing a patch);
    // patches with more than 64 W entries are summed in parts into partial
    // slots and combined by comb[] = {patch, first slot, parts, 0}
    int4* witems = nullptr;
    int* bgs = nullptr;
    int n_sblocks = 0;
    int4* comb = nullptr;
    int n_comb = 0;
    int n_slots = 0;
    float2* partial = nullptr;  // ceil(cap_batch/64) * n_slots * 16 * 64
    int* entries = nullptr;  // sample indices, grouped by patch
    // per entry k: the 16 weights wx*wy of the patch's points for that
    // entry's sample (ent_w[16k + 4a + q], point (a, q)), 0 outside its
    // clamped window
    float* ent_w = nullptr;
    // 2-D interp through LDS: samples binned by the kIT x kIT tile of their
    // (re-based) window start; work items {tile, rec begin, rec end, 0} of at
    // most kIChunk samples; per binned sample a record of 5 float4:
    //   {i, ox, oy, 0 (ints)}, wx[0..7], wy[0..7]
    // with (ox, oy) the window start inside the tile's staged region
    // (origin tile*kIT - ipadl) and zero weights past the taps
    int4* iitems = nullptr;
    int n_iitems = 0;
    // the same tiles cut into items of <= kIChunk4, x = tile | (pitch - kIR) << 24
    int4* iitems4 = nullptr;
    int n_iitems4 = 0;
    float4* irec = nullptr;
    // the same records with each item's samples reordered for the
    // four-segment kernel: within a tile the records of irec are ordered so
    // that every wavefront's samples of k_interp_lds fall on different LDS
    // banks, and those of irec4 for k_interp_lds4 (whose wavefronts take 16
    // samples instead of 8, with another tile layout)
    float4* irec4 = nullptr;
    int itx = 0, ity = 0, ipadl = 0;
    // 3-D trajectories whose samples take few distinct z windows (a stack of
    // 2-D trajectories, kz on a Cartesian raster): class k's z taps, with the
    // zero padding, shifts, FFT and deapodization along z, form one complex
    // matrix zmat[k*Nz + z] applied to the image's z columns, and the rest is
    // the 2-D pipeline (y FFT, x FFT, x-y gridding) on nzc planes per grid.
    // The x-y work lists above are built over the patches and tiles of all
    // planes (plane p, patch q at p*npp + q; interp items carry their plane in
    // .w); the planes of B grids are stored plane-major, (k*B + b)*Gx*Gy.
    // dapxy: the x-y part of the deapodization per image column (y fastest);
    // zcls[i]: sample i's class.  nzc = 0: the full 3-D grid is used.
    int nzc = 0;
    int npp = 0;  // 4x4 patches per plane
    float2* zmat = nullptr;
    float* dapxy = nullptr;
    int* zcls = nullptr;
    // 3-D spread (and B < 16): point-stationary tiles
    int ntx = 0, nty = 0, ntz = 0;
    int* tile_off = nullptr;  // ntiles + 1
    int* tile_idx = nullptr;  // concatenated sample indices
    long long tile_entries = 0;
    // kernel variants chosen by the autotuner for a batch of B grids: the
    // interpolation as a gather (0), LDS-staged with all segments at once over
    // bins of <= kIChunk samples (1) or four segments at a time over bins of
    // <= kIChunk4 samples (2); the matrix-core (1) or vector-unit (0) spread;
    // no entry: the defaults
    struct Tuned {
        int B, lds, mfma;
    };
    std::vector<Tuned> tuned;
    const Tuned* tuned_for(int B) const
    {
        for (const Tuned& t : tuned)
            if (t.B == B) return &t;
        return nullptr;
    }
    // workspace for batched transforms (grown on demand)
    int cap_batch = 0;
    float2* grids = nullptr;    // cap_batch * G
    float2* tbuf = nullptr;     // cap_batch * Nx*Nz*Gy
    float2* tbuf2 = nullptr;    // 3-D only: cap_batch * Nz*Gx*Gy
    float2* vals = nullptr;     // M * cap_batch per-sample spread values
    float2* img_tmp = nullptr;
    float2* dat_tmp = nullptr;

    void ensure_batch(int B);
    void release();
};

}  // namespace pg
