/*
 * pg_oracle.h -- CPU restatement of PowerGrid's field-corrected SENSE path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in powergrid_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and never as the thing measured
 * or shipped.  The product path never calls into it.
 *
 * Every function is written from the reference's semantics (file:line
 * citations are into PowerGrid/ of mrfil/PowerGrid) and exists in a float
 * (_f) and a double (_d) variant, like the reference's float/double template
 * instantiations.  Complex data is interleaved (re, im), the image vector is
 * y-fastest (idx = y + x*Ny + z*Nx*Ny) and the oversampled grid is indexed
 * ny + nx*Gy + nz*Gx*Gy, as in the reference.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   LUT / bessi0        scipy.special.i0 at the floored LUT argument
 *   FFT                 numpy.fft (unnormalised, both signs)
 *   NUFFT pipeline      exact DFT at fm=0 (rel-L2 bound) + adjointness
 *   TimeSegmentation    L=1 == G; fm=0 closed form; dense-matrix evaluation
 *   SENSE               nc=1,S=1 == inner op; adjointness
 *   Gdft (float trig)   numpy float64 DFT
 *   exact NUDFT (f64)   numpy float64 DFT
 *   penalties           hand-computed small vectors (wrap + Ctd tail)
 *   PCG                 tiny dense least-squares problem vs numpy solve
 */
#ifndef PG_ORACLE_H
#define PG_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* thread count used by the parallel loops (1 = the reference's serial GCC
 * build; >1 mirrors the OPENACC_MP multicore build). */
void orc_set_threads(int n);
int orc_get_threads(void);

#define ORC_DECL(R, S)                                                        \
  /* griddingSupport.cpp:33-76 */                                             \
  size_t orc_lut_size_##S(R width);                                           \
  void orc_lut_##S(R beta, R width, R* lut);                                  \
  R orc_kb_beta_##S(R gridos, R width);                                       \
  /* fftCPU.cpp:48-106: in-place unnormalised C2C, sign -1 fwd, +1 bwd */     \
  void orc_fft_##S(R* data, int nx, int ny, int nz, int sign);                \
  void orc_fft1d_##S(R* data, int n, int sign);                               \
  /* griddingSupport.cpp:209-368 (Nz==1 -> 2-D) */                            \
  void orc_deapod_##S(R* dst, const R* src, int Nx, int Ny, int Nz, R W,      \
                      R beta, R os);                                          \
  /* griddingSupport.cpp:456-547 / 371-453 / 551-627 */                       \
  void orc_zero_pad_##S(R* dst, const R* src, int Nx, int Ny, int Nz, R os);  \
  void orc_crop_##S(R* dst, const R* src, int Nx, int Ny, int Nz, int Gx,     \
                    int Gy, int Gz);                                          \
  void orc_fftshift_##S(R* dst, const R* src, int Gx, int Gy, int Gz);        \
  void orc_ifftshift_##S(R* dst, const R* src, int Gx, int Gy, int Gz);       \
  /* gridding.cpp:31-415: raw interp / spread on an oversampled grid */       \
  void orc_grid_forward_##S(int M, int Nx, int Ny, int Nz, R os, R W,         \
                            const R* kx, const R* ky, const R* kz,            \
                            const R* lut, size_t S_lut, const R* grid,        \
                            R* samples);                                      \
  void orc_grid_adjoint_##S(int M, int Nx, int Ny, int Nz, R os, R W,         \
                            const R* kx, const R* ky, const R* kz,            \
                            const R* lut, size_t S_lut, const R* samples,     \
                            R* grid);                                         \
  /* ftCpu.cpp:94-215 (float sin/cos, as the reference) */                   \
  void orc_ft_##S(R* out, const R* in, const R* kx, const R* ky,              \
                  const R* kz, const R* ix, const R* iy, const R* iz,         \
                  const R* fm, const R* t, int M, int N);                     \
  void orc_ift_##S(R* out, const R* in, const R* kx, const R* ky,             \
                   const R* kz, const R* ix, const R* iy, const R* iz,        \
                   const R* fm, const R* t, int M, int N);                    \
  /* operators: Gnufft.cpp:34-211, TimeSegmentation.cpp:36-324,              \
     SENSE.cpp:32-97, Gdft.cpp:32-136 */                                      \
  void* orc_gnufft_create_##S(int M, R gridos, int Nx, int Ny, int Nz,        \
                              const R* kx, const R* ky, const R* kz,          \
                              R kernelWidth);                                 \
  void* orc_gdft_create_##S(int M, int N, const R* kx, const R* ky,           \
                            const R* kz, const R* ix, const R* iy,            \
                            const R* iz, const R* fm, const R* t);            \
  /* GdftR2.cpp:32-107, ftCpuWithGrads.cpp:98-212 */                         \
  void* orc_gdftr2_create_##S(int M, int N, const R* kx, const R* ky,         \
                              const R* kz, const R* ix, const R* iy,          \
                              const R* iz, const R* fm, const R* t, int nx,   \
                              int ny, int nz);                                \
  void orc_gdftr2_grad_##S(const R* fm, int nx, int ny, int nz, int dim,      \
                           R* out);                                           \
  void* orc_ts_create_##S(void* G, const R* fm, const R* t, int M, int N,     \
                          int L, int interptype, int shots);                  \
  void* orc_sense_create_##S(void* G, const R* sens, int M, int N, int nc);   \
  void* orc_dense_create_##S(const R* A, int rows, int cols);                 \
  /* pcSenseTimeSeg.cpp:40-171 */                                             \
  void* orc_pcsense_create_##S(int Nd, int Ns, const R* kx, const R* ky,      \
                               const R* kz, int Nx, int Ny, int Nz, int Nc,   \
                               const R* t, int L, int interptype,             \
                               const R* sens, const R* fm, const R* phase,    \
                               R kernelWidth);                                \
  /* MPI/mpipcSENSETimeSeg.cpp:39-262, one rank's share */                     \
  void* orc_pcsense_rank_create_##S(int Nd, int Ns, const R* kx,              \
                                    const R* ky, const R* kz, int Nx, int Ny, \
                                    int Nz, int Nc, const R* t, int L,        \
                                    int interptype, const R* sens,            \
                                    const R* fm, const R* phase,              \
                                    R kernelWidth, int rank, int nranks);     \
  int orc_pcsense_rank_tasks_##S(void* op, int* shot, int* coil);             \
  /* pcSENSE.cpp:39-153 (an exact DFT per shot) */                            \
  void* orc_pcsense_dft_create_##S(int Nd, int Ns, const R* kx, const R* ky,  \
                                   const R* kz, int Nx, int Ny, int Nz,       \
                                   int Nc, const R* t, const R* sens,         \
                                   const R* fm, const R* phase);              \
  void orc_op_forward_##S(void* op, const R* in, R* out);                     \
  void orc_op_adjoint_##S(void* op, const R* in, R* out);                     \
  int orc_op_n_in_##S(void* op);                                              \
  int orc_op_n_out_##S(void* op);                                             \
  void orc_op_destroy_##S(void* op);                                          \
  /* TS interpolator / phase tables (TimeSegmentation.cpp:64-93,213-218) */   \
  int orc_ts_tables_##S(void* ts, R* AA, R* Wo, R* tau_tmin);                 \
  int orc_ts_set_aa_complex_##S(void* ts, const R* AAc);                      \
  /* Robject.cpp:49-193, QuadPenalty.cpp:45-64, TVPenalty.cpp:43-68 */        \
  void* orc_penalty_create_##S(int kind, int Nx, int Ny, int Nz, R beta,      \
                               R delta, int dims2penalize);                   \
  void orc_penalty_destroy_##S(void* R_);                                     \
  void orc_penalty_cd_##S(void* R_, const R* d, int dim, R* out);             \
  void orc_penalty_ctd_##S(void* R_, const R* d, int dim, R* out);            \
  void orc_penalty_gradient_##S(void* R_, const R* x, R* g);                  \
  void orc_penalty_denom_##S(void* R_, const R* ddir, const R* x, R* out2);   \
  R orc_penalty_value_##S(void* R_, const R* x);                              \
  /* solve_pwls_pcg.hpp:48-167; W may be NULL (= ones, reconSolve.h:63-82) */ \
  int orc_pcg_##S(void* op, void* R_, const R* x0, const R* W, const R* y,    \
                  int niter, R* x_out, R* resid_norms);

ORC_DECL(float, f)
ORC_DECL(double, d)

/* exact field-corrected NUDFT in double precision with double sin/cos
 * (the FP64 oracle; not a restatement -- Gdft computes its trig in float) */
void orc_nudft_exact_fwd(double* out, const double* in, const double* kx,
                         const double* ky, const double* kz, const double* ix,
                         const double* iy, const double* iz, const double* fm,
                         const double* t, int M, int N);
void orc_nudft_exact_adj(double* out, const double* in, const double* kx,
                         const double* ky, const double* kz, const double* ix,
                         const double* iy, const double* iz, const double* fm,
                         const double* t, int M, int N);

#ifdef __cplusplus
}
#endif
#endif
