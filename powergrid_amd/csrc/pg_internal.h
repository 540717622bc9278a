// pg_internal.h -- shared definitions for the gfx950 implementation of the
// field-corrected SENSE path (Gnufft / TimeSegmentation / SENSE / Gdft and
// the PWLS-PCG loop around them).
//
// Layout conventions (identical to the reference, PowerGrid/):
//   image  : y-fastest, idx = y + x*Ny + z*Nx*Ny        (reconSolve.cpp:40-53)
//   grid   : ny + nx*Gy + nz*Gx*Gy                       (gridding.cpp:100,202)
//   k-space: coil-major blocks, sample i of coil c at c*M + i (SENSE.cpp:70)
//   complex: interleaved float2
// Batched grids are stored grid-major: grid b = c*L + l occupies
// [b*G, (b+1)*G).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/powergrid_amd.h"

namespace pg {

// ------------------------------------------------------------ errors ----
struct Error {
    int code;
    std::string msg;
};

#define PG_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            throw ::pg::Error{PG_ERR_HIP, std::string(#call) + ": " +       \
                                              hipGetErrorString(e_)};       \
    } while (0)

#define PG_REQUIRE(cond, code, text)                                        \
    do {                                                                    \
        if (!(cond)) throw ::pg::Error{(code), (text)};                     \
    } while (0)

void set_last_error(const std::string& s);

// ------------------------------------------------------- FFT plans ----
// A 1-D length is factored into radix-4 / radix-2 / radix-3 stages.
constexpr int kMaxStages = 16;
struct FftPlan1D {
    int n = 1;
    int nst = 0;
    int radix[kMaxStages];
    float2* tw = nullptr;  // device: exp(-2 pi i k / n), k < n
};

bool fft_length_supported(int n);
FftPlan1D make_fft_plan(int n);
void free_fft_plan(FftPlan1D& p);

// Kernel-side copy of a 1-D plan (passed by value).
struct FftArgs {
    int n;
    int nst;
    int radix[kMaxStages];
    const float2* tw;
};
inline FftArgs fft_args(const FftPlan1D& p)
{
    FftArgs a;
    a.n = p.n;
    a.nst = p.nst;
    for (int i = 0; i < kMaxStages; ++i) a.radix[i] = i < p.nst ? p.radix[i] : 1;
    a.tw = p.tw;
    return a;
}

// ------------------------------------------------------- geometry ----
struct Geometry {
    int Nx, Ny, Nz;     // image
    int Gx, Gy, Gz;     // oversampled grid
    int dims;           // 2 or 3
    float os;           // grid oversampling (gridOS)
    float W;            // kernel width in original-grid units (reference: 4.0)
    float beta;         // KB shape (Gnufft.cpp:85-88)
    int ntaps;          // taps per dimension, ceil(W*os)
    int offX, offY, offZ;  // zero_pad offsets
    long long N() const { return (long long)Nx * Ny * Nz; }
    long long G() const { return (long long)Gx * Gy * Gz; }
};

// Spread binning: for each grid tile, the samples whose window intersects it.
constexpr int kSpreadTX = 16;  // tile extent along x (grid points)
constexpr int kSpreadTY = 16;  // tile extent along y
constexpr int kSpreadTZ = 4;   // tile extent along z (3-D)

struct Gnufft;

}  // namespace pg

// --------------------------------------------------- kernel launchers ----
// (defined in the .hip files; all asynchronous on `stream`)
namespace pgk {

// 2-D batched spread on the matrix cores (k_spread_patch_mfma) when set and
// the grid count is a multiple of 4; the vector-unit kernel otherwise
extern std::atomic<int> g_spread_mfma;
// plans of 3-D trajectories with few distinct z windows are built for the
// z-class planes (pg_gnufft.h) when set (the default); 0 builds them for the
// full 3-D grid
extern std::atomic<int> g_zclass;
// the LDS-staged interpolation stages four segments at a time in 256-thread
// blocks over items of <= 64 samples (k_interp_lds4): 2 always, 1 (the
// default) from 4 segments up, 0 never (all segments at once in
// 512-thread blocks over items of <= 128 samples)
extern std::atomic<int> g_interp4;

// the 2-D (and z-class plane) spread of a time-segmented batch forms each
// grid's value from a compact sample row (the nc coil values and the L
// segment weights) instead of reading a row of nc*L values, when set (the
// default)
extern std::atomic<int> g_spread_rows;

// on z-class planes with gridos 2, the adjoint y IFFT and z stage run as one
// kernel (the cropped class columns staged in LDS) when set (the default)
extern std::atomic<int> g_zfuse;
// the forward z stage of z-class planes runs on the matrix cores (Nz and
// the class count at most 32) when set (default), else on the vector units
extern std::atomic<int> g_zmfma;
// the fused adjoint y IFFT + z stage of z-class planes runs its IFFTs on
// four wavefronts and its class sums on the matrix cores (Ny 64 or 128, Nz
// and the class count at most 32) when set (default)
extern std::atomic<int> g_zamfma;
// the forward z stage of z-class planes on the matrix cores runs fused with
// the planes' y FFT (Ny 64 or 128 at gridos 2) when set (default)
extern std::atomic<int> g_zfwd;

// the compact-row patch spread stages each stage's row slices in LDS
// (k_spread_patch_st) when set (default), else reads them per entry
extern std::atomic<int> g_spread_stage;

// whether a batched launch of B grids of plan G runs on its z-class planes
bool zclass_active(const pg::Gnufft& G, int B);
// floats per compact sample row of a spread of nc coils x L segments with
// weights AA (and AAi), or 0 when that spread reads value rows
int spread_row_floats(const pg::Gnufft& G, int nc, int L, const float* AA, const float* AAi);

struct WeightArgs {
    // image-side weights applied before the forward NUFFT and after the
    // adjoint one.  Any pointer may be null (= 1).
    const float2* S;    // coil maps, nc x N (SENSE.cpp:59-63 / :89-92)
    const float2* Wo;   // segment phases exp(-i fm (l tau + Tmin)), L x N
    const float2* WoH;  // conj(Wo), L x N
    // data-side weights: AA (L x M, real) -- TimeSegmentation.cpp:291-293
    const float* AA;
};

// forward: image(s) -> batched grids  (deapod, pad, fftshift, FFT, ifftshift;
// with `planes`, the z-class planes of the grids: the z taps applied in the
// z pass)
void launch_fwd_prep_fft(const pg::Gnufft& G, const float2* img, long long img_cstride,
                         WeightArgs w, int nc, int L, float2* tbuf, float2* tbuf2, float2* grids,
                         hipStream_t s, bool planes = false);
// forward: batched grids (or z-class planes) -> samples (KB interpolation,
// sum over segments)
void launch_interp(const pg::Gnufft& G, const float2* grids, int nc, int L, const float* AA,
                   const float* AAi, float2* out, long long out_cstride, hipStream_t s,
                   bool planes = false);
// adjoint: samples -> batched grids (KB spread, deterministic tiles)
// (with resid_y set, the samples are W (resid_y - ax64), ax64 the FP64 Ax of
// the CG, and `in` is unused; with vals_ready the per-sample values
// G.vals have already been formed, by the PCG update kernel)
// (vals, partial: the per-sample value rows and the partial-sum slots of
// heavy patches, when not the plan's own workspaces)
void launch_spread(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc, int L,
                   const float* AA, const float* AAi, const float2* resid_y, const double2* ax64,
                   const float* Wd, float2* grids, bool vals_ready, hipStream_t s,
                   bool planes = false, float2* vals = nullptr, float2* partial = nullptr);
// whether launch_spread reads per-sample values (all but the 3-D atomic path)
bool spread_uses_values(const pg::Gnufft& G, int B);
// whether the LDS-staged interpolation serves L segments of plan G (the
// gather kernel does otherwise)
bool interp_lds_possible(const pg::Gnufft& G, int L, bool aa);
// the per-sample values G.vals of a spread (see launch_spread)
void launch_spread_values(const pg::Gnufft& G, const float2* in, long long in_cstride, int nc,
                          int L, const float* AA, const float* AAi, const float2* resid_y,
                          const double2* ax64, const float* Wd, hipStream_t s,
                          float2* vals = nullptr);
// adjoint: batched grids (or z-class planes, expanded along z in the z
// pass) -> image(s) (ifftshift, IFFT, fftshift, crop, deapod, x WoH, sum
// over segments, x conj(S))
void launch_adj_fft_post(const pg::Gnufft& G, float2* grids, float2* tbuf, float2* tbuf2,
                         WeightArgs w, int nc, int L, float2* out, long long out_cstride,
                         hipStream_t s, bool planes = false);

}  // namespace pgk
