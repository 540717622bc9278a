/*
 * powergrid_amd.h -- C ABI of the MI355X (gfx950) implementation of
 * PowerGrid's field-corrected non-Cartesian SENSE path.
 *
 * The reference exposes this path as C++ operator objects with
 * `Col<complex<T>> operator*(const Col<complex<T>>&) const` (forward) and
 * `operator/` (adjoint), consumed by solve_pwls_pcg / reconSolve
 * (SENSE.cpp:51-97, TimeSegmentation.cpp:252-324, Gnufft.cpp:153-211,
 * Gdft.cpp:52-136, solve_pwls_pcg.hpp:48-167, reconSolve.h:63-82).  Below the
 * operators the reference crosses a raw-pointer seam, computeFd_CPU_Grid /
 * computeFH_CPU_Grid (gridding.h:84-103).  This header replaces both: the
 * operator objects become opaque handles with the reference's constructor
 * arguments, and operator* / operator/ become pg_op_forward / pg_op_adjoint
 * on device buffers (or pg_op_*_host on host buffers).  The C++ classes in
 * include/powergrid/ rebuild the reference's class API on top of it.
 *
 * Conventions: complex data is interleaved float (re, im); the image vector
 * is y-fastest (idx = y + x*Ny + z*Nx*Ny, reconSolve.cpp:30-54); multi-coil
 * k-space is coil-major (sample i of coil c at c*M + i, SENSE.cpp:70); the
 * field map is in rad/s and times in s (TimeSegmentation.h:56).  All calls
 * return a pg_status; pg_last_error() describes the most recent failure.
 * Nothing here exits the process (the reference exit(1)s on malloc failure,
 * gridding.cpp:454-457).  Launches are asynchronous on the given stream
 * (NULL = the default stream); a handle must not be used concurrently from
 * two streams.
 */
#ifndef POWERGRID_AMD_H
#define POWERGRID_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PG_OK = 0,
    PG_ERR_ARG = 1,          /* invalid argument */
    PG_ERR_UNSUPPORTED = 2,  /* valid for the reference, not supported here */
    PG_ERR_HIP = 3,          /* HIP runtime error */
    PG_ERR_OOM = 4,          /* device allocation failed */
    PG_ERR_STATE = 5,        /* call not valid for this handle */
    PG_ERR_NUMERIC = 6       /* NaN/Inf detected where the reference warns */
} pg_status;

typedef struct pg_op_s* pg_op;           /* Gnufft, Gdft, TimeSegmentation or SENSE */
typedef struct pg_penalty_s* pg_penalty; /* QuadPenalty or TVPenalty */

const char* pg_status_string(int status);
const char* pg_last_error(void);
/* library version, e.g. "0.1.0" */
const char* pg_version(void);

/* ---- device plumbing (so a C or C++ caller needs no other runtime) ---- */
int pg_set_device(int device);
int pg_malloc(void** ptr, size_t bytes);
int pg_free(void* ptr);
/* free and total device memory in bytes (hipMemGetInfo), for callers that
 * admit work by free HBM (no counterpart in the reference) */
int pg_mem_info(size_t* free_bytes, size_t* total_bytes);
/* blocking copies; with stream NULL they wait for all work on the device
 * (the library's streams are non-blocking), pg_memcpy_dtod only enqueues */
int pg_memcpy_htod(void* dst_dev, const void* src_host, size_t bytes, void* stream);
int pg_memcpy_dtoh(void* dst_host, const void* src_dev, size_t bytes, void* stream);
int pg_memcpy_dtod(void* dst_dev, const void* src_dev, size_t bytes, void* stream);
int pg_stream_create(void** stream);
int pg_stream_destroy(void* stream);
int pg_stream_synchronize(void* stream);

/* ---- operators ---------------------------------------------------------
 * pg_gnufft_create replaces Gnufft<float>::Gnufft (Gnufft.cpp:34-122,
 * Gnufft.h:51-53): same arguments plus kernelWidth (original-grid units; the
 * reference hard-codes 4.0 at Gnufft.cpp:84; W*gridos must be 6, 8 or 12).
 * Requirements: gridos*Nx integer and a multiple of 4, Nx == Ny (the
 * reference's zero_pad2d/circshift2 are x-fastest and only consistent for
 * square images, griddingSupport.cpp:488,565).  Grid lengths with a prime
 * factor above 13 leave only the spatial interpolation available (operator* and
 * operator/ then return PG_ERR_UNSUPPORTED).  ix/iy/iz are accepted for
 * signature parity and may be NULL (Gnufft does not use them). */
int pg_gnufft_create(pg_op* out, size_t M, float gridos, size_t Nx, size_t Ny, size_t Nz,
                     const float* kx, const float* ky, const float* kz, const float* ix,
                     const float* iy, const float* iz, float kernelWidth);

/* replaces Gdft<float>::Gdft (Gdft.cpp:32-49, Gdft.h:47-49).  precision 0:
 * the reference's arithmetic class (FP32 phase, FP32 sin/cos,
 * ftCpu.cpp:120-130); precision 1: FP64 phase and sin/cos with FP64
 * accumulation (the exact-NUDFT oracle form). */
int pg_gdft_create(pg_op* out, size_t M, size_t N, const float* kx, const float* ky,
                   const float* kz, const float* ix, const float* iy, const float* iz,
                   const float* fm, const float* t, int precision);

/* replaces GdftR2<float>::GdftR2 (GdftR2.cpp:32-57, GdftR2.h:45-48): Gdft
 * (precision 0) whose every (sample, voxel) pair is weighted by
 * sinc(kx/Nx + Gx t) sinc(ky/Ny + Gy t) sinc(kz/Nz + Gz t)
 * (ftCpuWithGrads.cpp:98-212), G = the field map's differences along the
 * image's offsets 1, Ny and Nx*Ny (GdftR2.cpp:67-107).  In 2-D the
 * reference's z map reads outside the field map; it is zero here.
 * N must equal Nx*Ny*Nz. */
int pg_gdftr2_create(pg_op* out, size_t M, size_t N, const float* kx, const float* ky,
                     const float* kz, const float* ix, const float* iy, const float* iz,
                     const float* fm, const float* t, size_t Nx, size_t Ny, size_t Nz);

/* replaces TimeSegmentation<float,G>::TimeSegmentation
 * (TimeSegmentation.cpp:36-221, TimeSegmentation.h:70-71).  interptype 1:
 * Hanning (:80-93); 2: the exact least-squares min-max interpolator
 * (:94-150), its complex weights formed in FP64 (a Jacobi eigen-solve of the
 * L x L Gram matrix, then one device pass over samples x voxels), applied
 * one segment at a time; 3: the histogram approximation of 2 (:157-207),
 * the field map binned to 256 frequencies, formed the same way over the
 * bins (L <= 16; a field map of zero range returns PG_ERR_UNSUPPORTED, where
 * the reference's FFT length would be infinite).  G is not owned and must
 * outlive the new handle (TimeSegmentation.h:55). */
int pg_ts_create(pg_op* out, pg_op G, const float* fieldmap, const float* tvec, size_t M,
                 size_t N, size_t L, size_t interptype, size_t shots);

/* replaces SENSE<float,G>::SENSE (SENSE.cpp:32-44, SENSE.h:64); sens is N*nc
 * interleaved complex, coil-major (SENSE.cpp:40).  G is not owned. */
int pg_sense_create(pg_op* out, pg_op G, const float* sens, size_t M, size_t N, size_t nc);

/* replaces pcSenseTimeSeg<float>::pcSenseTimeSeg (pcSenseTimeSeg.cpp:40-124,
 * pcSenseTimeSeg.h:48-50): shot-phase-corrected, field-corrected SENSE.
 * kx, ky, kz (kz may be null for 2-D) and tvec hold n_shots shots of Nd
 * samples, shot-major; sens is Nx*Ny*Nz*nc interleaved complex, coil-major;
 * fieldmap is Nx*Ny*Nz (rad/s); shotphase is Nx*Ny*Nz*n_shots real phases,
 * shot-major.  The reference derives n_shots and Nd from the array lengths;
 * here they are passed.  kernelWidth: 4.0 is the reference's (Gnufft.cpp:84).
 * Forward data is Nd x (n_shots*nc), column s + c*n_shots
 * (pcSenseTimeSeg.cpp:129-148); the adjoint sums over coils and shots
 * (:152-171).  The per-shot operators are owned by the handle. */
int pg_pcsense_create(pg_op* out, const float* kx, const float* ky, const float* kz, size_t Nx,
                      size_t Ny, size_t Nz, size_t nc, const float* tvec, size_t L,
                      size_t interptype, const float* sens, const float* fieldmap,
                      const float* shotphase, size_t n_shots, size_t Nd, float kernelWidth);

/* replaces one MPI rank's share of mpipcSENSETimeSeg<float>
 * (MPI/mpipcSENSETimeSeg.cpp:39-150, mpipcSENSETimeSeg.h): the tasks
 * t = 0..n_shots*nc-1 (shot t / nc, coil t % nc, :106-113) are dealt round
 * robin, t to rank t mod nranks (:115-124); this handle covers rank `rank`'s
 * tasks, on one device.  Arguments otherwise as pg_pcsense_create.  As in the
 * reference, each task's TimeSegmentation is built with shots = n_shots
 * (:146-148), so sample i of a shot takes the interpolator weights of sample
 * i mod (Nd/n_shots) (Nd must be a multiple of n_shots).  Forward: the
 * rank's Nd x ntask columns in task order (tempOutData, :163-171); adjoint:
 * those columns -- column jj is the full data's column shot + coil*n_shots
 * (:221-223) -- to the rank's partial image (:218-228), summed in task order.
 * The gather to rank 0, the cross-rank sum and the broadcasts (:177-212,
 * :230-260) are the caller's collectives. */
int pg_pcsense_rank_create(pg_op* out, const float* kx, const float* ky, const float* kz,
                           size_t Nx, size_t Ny, size_t Nz, size_t nc, const float* tvec, size_t L,
                           size_t interptype, const float* sens, const float* fieldmap,
                           const float* shotphase, size_t n_shots, size_t Nd, float kernelWidth,
                           size_t rank, size_t nranks);

/* the tasks of a pg_pcsense_create / pg_pcsense_rank_create handle in data
 * column order (the reference's taskList, shotList, coilList,
 * MPI/mpipcSENSETimeSeg.cpp:104-124): *ntask, and per task (each array may be
 * null, else holds ntask entries) its shot, coil and column in the handle's
 * forward output (shot + coil*n_shots for pg_pcsense_create) */
int pg_pcsense_tasks(pg_op h, size_t* ntask, uint32_t* shot, uint32_t* coil, uint32_t* column);

/* replaces pcSENSE<float>::pcSENSE (pcSENSE.cpp:39-102, pcSENSE.h:47-49):
 * shot-phase-corrected SENSE over an exact DFT per shot (Gdft, FP32 trig).
 * Arguments as pg_pcsense_create, without L/interptype/kernelWidth.  As in the
 * reference, the forward applies S_c exp(+i P_s) (pcSENSE.cpp:117) and the
 * adjoint conj(S_c) exp(+i P_s) (:142): the pair is not adjoint unless the
 * shot phases are zero. */
int pg_pcsense_gdft_create(pg_op* out, const float* kx, const float* ky, const float* kz,
                           size_t Nx, size_t Ny, size_t Nz, size_t nc, const float* tvec,
                           const float* sens, const float* fieldmap, const float* shotphase,
                           size_t n_shots, size_t Nd);

int pg_op_destroy(pg_op op);
/* n_in = image length (N), n_out = data length (M, or M*nc for SENSE) */
int pg_op_shape(pg_op op, size_t* n_in, size_t* n_out);

/* PG_MODE_BATCHED (default): one launch sequence over all coils x segments;
 * PG_MODE_PLAIN: one NUFFT at a time in the reference's loop order
 * (SENSE.cpp:65-67, TimeSegmentation.cpp:272-288). */
#define PG_MODE_BATCHED 0
#define PG_MODE_PLAIN 1
int pg_op_set_mode(pg_op op, int mode);

/* operator* (forward) and operator/ (adjoint) on device buffers:
 * replaces Gnufft::operator* / operator/ (Gnufft.cpp:153-211, i.e.
 * computeFd_CPU_Grid / computeFH_CPU_Grid, gridding.h:84-103),
 * TimeSegmentation::operator* / operator/ (TimeSegmentation.cpp:252-324),
 * SENSE::operator* / operator/ (SENSE.cpp:51-97) and Gdft::operator* /
 * operator/ (Gdft.cpp:52-136).  in and out must not alias. */
int pg_op_forward(pg_op op, const void* in_dev, void* out_dev, void* stream);
int pg_op_adjoint(pg_op op, const void* in_dev, void* out_dev, void* stream);
/* the same on host buffers (copies in, runs, copies out, synchronizes) */
int pg_op_forward_host(pg_op op, const void* in_host, void* out_host);
int pg_op_adjoint_host(pg_op op, const void* in_host, void* out_host);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the forward
 * (adjoint = 0) or adjoint (adjoint != 0) of a Gdft / GdftR2 handle, or of a
 * SENSE handle over one in PG_MODE_BATCHED, with the partial sums in a static
 * 64 MiB device array of the library instead of the handle's workspace.  It
 * calls no hipMalloc or hipFree; results are bit-identical to pg_op_forward /
 * pg_op_adjoint.  An operator needing more than 8M complex partial sums
 * (forward: slices x coils x M; adjoint: slices x N) is rejected with
 * PG_ERR_ARG.  Every call on a device shares that device's array: calls that
 * overlap on one device (other streams or threads) race. */
int pg_dft_apply_static(pg_op op, int adjoint, const void* in_dev, void* out_dev, void* stream);


/* ---- gridding alone and direct reconstruction ---------------------------
 * The oversampled grid is Gx x Gy (x Gz) complex, y fastest, then x, then z
 * (gridding.cpp:100, idx = ny + nx*Gy + nz*Gx*Gy), with no FFT shift:
 * Gx = ceil(gridos*Nx) rounded up to even, Gz = ceil(gridos*Nz) in 3-D and
 * 1 in 2-D (Gnufft.cpp:228-234).  pg_gnufft_grid_size returns them. */
int pg_gnufft_grid_size(pg_op G, size_t* Gx, size_t* Gy, size_t* Gz);
/* Gnufft<float>::forwardSpatialInterp (Gnufft.cpp:214-263): the grid
 * interpolated at the trajectory with the Kaiser-Bessel taps
 * (gridding_forward_2D/3D, gridding.cpp:217-415); samples: M complex. */
int pg_gnufft_forward_spatial(pg_op G, const void* grid_dev, void* samples_dev, void* stream);
/* Gnufft<float>::adjointSpatialInterp (Gnufft.cpp:266-337): the samples
 * spread onto the zeroed grid (gridding_adjoint_2D/3D, gridding.cpp:31-214,
 * sdc 1); grid: Gx*Gy*Gz complex, fully overwritten. */
int pg_gnufft_adjoint_spatial(pg_op G, const void* samples_dev, void* grid_dev, void* stream);
/* calc3DDensityCompensation (directRecon.cpp:28-60): the Pipe-Menon
 * iteration dcf <- dcf / real(interp(spread(dcf))), twice, from
 * dcf = 1/M^2, with a Gnufft on a 3x oversampled grid (kernel width 4), then
 * |dcf| / ||dcf||_2.  kx, ky, kz: host arrays of M (kz ignored when
 * Nz == 1); dcf_dev: M floats on the device.  3*Ninplane must be a multiple
 * of 4 (and 3*Nz in 3-D). */
int pg_density_compensation(const float* kx, const float* ky, const float* kz, size_t M,
                            size_t Ninplane, size_t Nz, float* dcf_dev, void* stream);
/* the coil loop of gridCoilImages (directRecon.cpp:118-127): image c =
 * G / (dcf % data_c) for c < ncoils, all coils in one batched adjoint.
 * G: a Gnufft handle; dcf: M floats; data: ncoils x M complex (coil-major);
 * images: ncoils x N complex (coil-major). */
int pg_grid_coil_images(pg_op G, const float* dcf_dev, const void* data_dev, size_t ncoils,
                        void* images_dev, void* stream);
/* calcSumOfSquaresImage (directRecon.cpp:63-73): out[v] =
 * sqrt(sum_c |x[v + c*nvox]|) over nvox voxels x ncoils coils.  As in the
 * reference the coil magnitudes are summed, not their squares. */
int pg_sum_of_squares(const void* coil_images_dev, size_t nvox, size_t ncoils, float* out_dev,
                      void* stream);

/* ---- image output -------------------------------------------------------
 * writeNiftiRealImage (processNIFTI.hpp:41-101): <filename>.nii, a single-
 * file NIfTI-1 image of float32 values (host buffer), with the reference's
 * header: dim = {3, Nx, Ny, Nz, Nslab, N_TR, 1, 1}, pixdim[0..3] = 1,
 * vox_offset 352, scl_slope 100, mm units, magic "n+1"; the fields the
 * reference leaves uninitialised are zero. */
int pg_write_nifti_real(const char* filename, const float* data_host, size_t Nx, size_t Ny,
                        size_t Nz, size_t Nslab, size_t N_TR);
/* writeNiftiMagPhsImage (processNIFTI.hpp:104-119): |x| and arg(x) of a
 * complex device image, formed on the device, to <filename>_mag.nii and
 * <filename>_phs.nii. */
int pg_write_nifti_mag_phase(const char* filename, const void* image_dev, size_t Nx, size_t Ny,
                             size_t Nz, size_t Nslab, size_t N_TR, void* stream);

/* ---- penalties ----------------------------------------------------------
 * kind 0: QuadPenalty (QuadPenalty.cpp:32-64), kind 1: TVPenalty
 * (TVPenalty.cpp:32-68, delta used); the Robject methods Cd / Ctd /
 * Gradient / Denom / Penalty follow Robject.cpp:49-193. */
#define PG_PENALTY_QUAD 0
#define PG_PENALTY_TV 1
int pg_penalty_create(pg_penalty* out, int kind, size_t Nx, size_t Ny, size_t Nz, float beta,
                      float delta, size_t dims2penalize);
int pg_penalty_destroy(pg_penalty R);
int pg_penalty_gradient(pg_penalty R, const void* x_dev, void* g_dev, void* stream);
/* Denom(ddir, x) (Robject.cpp:167-193), returned as (re, im) in out2 (host) */
int pg_penalty_denom(pg_penalty R, const void* ddir_dev, const void* x_dev, double* out2,
                     void* stream);
int pg_penalty_value(pg_penalty R, const void* x_dev, double* out, void* stream);

/* ---- solver --------------------------------------------------------------
 * replaces solve_pwls_pcg<float,Tobj,Robj> (solve_pwls_pcg.hpp:48-167).
 * The whole loop runs on the device: vectors FP32, inner products, step
 * sizes and norms FP64.  W_dev may be NULL (= ones, as reconSolve.h:74-75).
 * mode PG_CG_GRAPH: one iteration captured as a HIP graph and replayed,
 * all branches decided on the device; PG_CG_HOST: host-driven loop that
 * reads the scalars back every iteration (the reference's control flow).
 * resid (host, niter entries, may be NULL) receives ||y - A x|| after each
 * completed iteration (solve_pwls_pcg.hpp:162). */
#define PG_CG_GRAPH 0
#define PG_CG_HOST 1
#define PG_EXIT_NITER 0       /* ran all iterations */
#define PG_EXIT_ZERO_GRAD 1   /* solve_pwls_pcg.hpp:88-91 */
#define PG_EXIT_NOT_DESCENT 2 /* :113-116 */
#define PG_EXIT_BAD_DENOM 3   /* :137-145 */
#define PG_EXIT_TOL 4         /* the tolerance of pg_solve_pwls_pcg_tol */
typedef struct {
    int iterations;   /* completed updates */
    int exit_reason;  /* PG_EXIT_* */
    double final_resid;
} pg_cg_info;
int pg_solve_pwls_pcg(pg_op A, pg_penalty R, const void* x0_dev, const float* W_dev,
                      const void* y_dev, size_t niter, int mode, void* x_out_dev, double* resid,
                      pg_cg_info* info, void* stream);
/* replaces reconSolve (reconSolve.h:63-82): W = ones, x0 = 0 */
int pg_recon_solve(pg_op A, pg_penalty R, const void* y_dev, size_t niter, int mode,
                   void* x_out_dev, double* resid, pg_cg_info* info, void* stream);
/* (no counterpart in the reference, which always runs niter iterations) the
 * same two with a stopping tolerance, for callers that run a reconstruction
 * "to a relative residual or niter iterations": the loop also ends, with
 * PG_EXIT_TOL, before the update of the first iteration k >= 1 whose
 * gradient ngrad = A' W (y - A x) - R'(x) (solve_pwls_pcg.hpp:80-86)
 * satisfies ||ngrad_k|| <= tol ||ngrad_0||.  tol <= 0 gives the plain
 * functions' results. */
int pg_solve_pwls_pcg_tol(pg_op A, pg_penalty R, const void* x0_dev, const float* W_dev,
                          const void* y_dev, size_t niter, double tol, int mode, void* x_out_dev,
                          double* resid, pg_cg_info* info, void* stream);
int pg_recon_solve_tol(pg_op A, pg_penalty R, const void* y_dev, size_t niter, double tol,
                       int mode, void* x_out_dev, double* resid, pg_cg_info* info, void* stream);
/* (no counterpart in the reference, which reconstructs one slice at a time)
 * K independent reconstructions -- reconSolve on A[k], R[k] (may be null),
 * y[k], each of its own size -- ordered after the work already on `stream`
 * (required), which waits for all of them: every solve is one captured
 * graph with its own CG scalars, and the graphs run concurrently on up to
 * four of the library's streams (pg_set_tuning "batch_streams"), the most
 * expensive first, each onto the stream with the least work so far; with
 * "batch_streams" 0 they are the children of one parent graph launch.
 * Stops as pg_recon_solve_tol; infos (may be null) receives K results.  The
 * K operators must not share operator objects (including inner
 * Gnufft/TimeSegmentation plans): their workspaces are used at once. */
int pg_recon_solve_batch(size_t K, const pg_op* A, const pg_penalty* R, const void* const* y_dev,
                         size_t niter, double tol, void* const* x_out_dev, pg_cg_info* infos,
                         void* stream);

/* ---- tuning (no counterpart in the reference) ----------------------------
 * "spread_mfma" 1: the 2-D batched spread runs on the matrix cores when the
 * grid count is a multiple of 4; 0 (default): on the vector units.  Both
 * give the same sums up to FP32 rounding order.
 * "autotune" 1: when an operator is first prepared for a batch of B grids,
 * its Gnufft plan times each variant of the 2-D (or z-class plane) batched
 * interpolation (gather, or LDS-staged tiles with sample bins of 128 or 64)
 * and spread (vector units or matrix cores) on its own workspaces and keeps
 * the fastest for that B, overriding "spread_mfma" and "interp_lds4";
 * 0 (default): the fixed choice.  The variants differ in
 * FP32 summation order only, but with autotuning on, which one runs -- and
 * so the last bits of a result -- depends on the timing.
 * "zclass" 1 (default): a 3-D Gnufft plan whose samples take at most Gz
 * distinct z windows (a stack of 2-D trajectories: kz on a Cartesian
 * raster) is built for z-class planes: batches of 16 grids or more apply
 * the separable kernel's z taps inside the z FFT pass and grid in x-y on
 * one plane per class; 0: plans made afterwards grid on the full 3-D grid.
 * The two differ in FP32 summation order only.
 * "interp_lds4" 1 (default): the LDS-staged interpolation stages four
 * segments at a time (256-thread blocks, items of <= 64 samples) when a
 * batch has 4 segments or more; 2: always; 0: never (all segments at once,
 * 512-thread blocks, items of <= 128 samples).  Same sums, FP32 order aside.
 * "batch_streams" n (0..8, default 4): streams of pg_recon_solve_batch.
 * "spread_rows" 1 (default): the 2-D (and z-class plane) spread of a
 * time-segmented batch forms each grid's value from a compact sample row
 * (the nc coil values and L segment weights) instead of reading a row of
 * nc*L values; 0: value rows.  Same products, same order.
 * "trace" 1: tracing ranges (pg_trace_ranges); 0 (default): none.
 * "zfuse" 1 (default): on z-class planes at gridos 2 the adjoint's y IFFT
 * and z stage run as one kernel; 0: two kernels with the class planes in
 * between.  Same sums, same order.
 * "zstage_mfma" 1 (default): the forward z stage of z-class planes (Nz and
 * the class count at most 32) forms its class sums on the matrix cores;
 * 0: on the vector units.  Same products, z sums in another order.
 * "zadj_mfma" 1 (default): the fused adjoint y IFFT + z stage of z-class
 * planes (Ny 64 or 128, Nz and the class count at most 32) runs its IFFTs on
 * all four wavefronts and its class sums on the matrix cores; 0: two
 * wavefronts' IFFTs and vector-unit sums.  Same products, sums in another
 * order.
 * "zfwd_fuse" 1 (default): the forward z stage on the matrix cores (Ny 64 or
 * 128 at gridos 2) and the class planes' y FFT run as one kernel, the class
 * planes kept in LDS; 0: two kernels.  Same products, same order.
 * "spread_stage" 1 (default): the compact-row spread loads the row slices of
 * each group of up to 32 patch entries into LDS with all loads in flight,
 * then sums from LDS (when a 64-grid chunk needs at most 48 floats of a
 * row); 0: one row load per entry and lane.  Same products, same order.
 * Settings take effect for launches (and solve graphs) made after the call;
 * "zclass" for plans created after it. */
int pg_set_tuning(const char* key, int value);
/* the number of z-class planes of the Gnufft plan under h (a Gnufft, or a
 * TimeSegmentation or SENSE over one); 0 when it grids on the full grid */
int pg_gnufft_zclasses(pg_op h, size_t* nzc);
/* the variants the autotuner chose for the Gnufft plan under h (a Gnufft,
 * or a TimeSegmentation or SENSE over one) at B grids: interp_lds 1 =
 * LDS-staged tiles with every segment at once over bins of <= 128 samples,
 * 2 = four segments at a time over bins of <= 64 samples, 0 = gather;
 * spread_mfma 1 = matrix cores, 0 = vector units; both -1 when the plan was
 * not tuned for B */
int pg_op_tuning(pg_op h, int B, int* interp_lds, int* spread_mfma);

/* ---- measurement --------------------------------------------------------
 * When enabled, every kernel launch of the library is bracketed by HIP
 * events on its stream and its duration accumulated per kernel name.  Not
 * for use under graph capture. */
int pg_profile_enable(int on);
int pg_profile_reset(void);
/* name: kernel family, e.g. "interp", "spread", "fwd_prep_fft", "adj_fft_post",
 * "gdft_fwd"; returns the launch count and total milliseconds */
int pg_profile_query(const char* name, int* count, double* total_ms);
/* tracing ranges, the reference's RANGE markers (Tracer.hpp:19-34, NVTX
 * there; roctx here, recorded by rocprofv3 --marker-trace): with
 * pg_set_tuning("trace", 1), operator applications (named as the
 * reference's operators, e.g. "SENSE::operator*"), their launch stages,
 * the PCG solve and its host-driven iterations, operator set-up and penalty
 * calls push ranges.  *available: whether libroctx64 could be opened;
 * *pushed: ranges pushed since the process started. */
int pg_trace_ranges(int* available, unsigned long long* pushed);

#ifdef __cplusplus
}
#endif
#endif /* POWERGRID_AMD_H */
