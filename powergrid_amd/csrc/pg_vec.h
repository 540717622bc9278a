// pg_vec.h -- elementwise, penalty and PWLS-PCG launchers (pg_vec.hip).
#pragma once

#include "pg_internal.h"

// device-resident PCG scalars (solve_pwls_pcg.hpp:60-77, kept in FP64)
struct CgScal {
    double ywy;        // yi' W yi
    double oldinprod;  // previous real(cdot(ngrad, ngrad))
    double newinprod;  // this iteration's (moved to oldinprod by the adir kernel)
    double gamma;
    double dAWAd, dAWr;
    double2 steps[4];  // line-search steps 0..3 (steps[0] = 0)
    double2 step;      // the final complex step, as the reference's CxT1 step
    int ii;            // loop index
    int iter;          // completed updates
    int stop;          // 0 or PG_EXIT_*
    int first;         // ddir = ngrad on the first iteration
    int pending;       // an update's residual norm still to be recorded
    double tol;        // relative gradient tolerance (0: none)
    double g0;         // real(cdot(ngrad, ngrad)) of the first iteration
};

namespace pgv {

// partial sums per reduction (blocks of a reducing kernel): up to
// kMaxParts (four 256-thread blocks per CU) over vectors of kBigParts
// elements or more, so that the bandwidth-bound CG passes over the data
// vectors (16.7M samples at the 3-D workload: 3-D workload 64.2 -> 65.2 CG
// it/s) keep enough loads in flight; up to kSmallParts below (at the 2-D
// workload's sizes the extra partials cost more than they gain)
constexpr int kMaxParts = 1024;
constexpr int kSmallParts = 256;
constexpr long long kBigParts = 1LL << 22;

struct PenSpec {
    int kind;  // 0 quad, 1 TV
    int nd;    // 2 or 3 penalised dimensions (Robject.cpp:128-133)
    long long n;
    long long off[3];  // 1, Ny, Nx*Ny (Robject.cpp:76)
    float beta, delta;
};

void cmul(float2* out, const float2* a, const float2* b, long long n, hipStream_t s);
void cmul_acc(float2* out, const float2* a, const float2* b, long long n, bool first, bool conj_b,
              hipStream_t s);
void rmul_acc(float2* out, const float* r, const float2* a, long long n, bool first, hipStream_t s);
void coil_sum(float2* out, const float2* P, int nc, long long N, hipStream_t s);
void fill(float2* out, long long n, float2 v, hipStream_t s);
void resid(float2* r, const float2* y, const double2* Ax, const float* W, long long n, hipStream_t s);
void f2d(double2* out, const float2* in, long long n, hipStream_t s);

void pen_gradient(const PenSpec& p, const float2* x, float2* g, hipStream_t s);
void pen_denom_pdot(const PenSpec& p, const float2* ddir, const float2* x, const double2* step,
                    bool want_pdot, double* part, double* out2, const int* stop, hipStream_t s);
void pen_value(const PenSpec& p, const float2* x, double* part, double* out, hipStream_t s);

// partial-sum buffers of one PCG iteration (per-block FP64 partials)
struct CgParts {
    double* g;     // grad: 2 x ng
    double* d;     // dir: nd
    double* a;     // adir: 2 x na
    double* l[3];  // penalty {denom, pdot} at line-search steps 0..2: 2 x nl each
    double* u;     // update: nu
    int ng, nd, na, nl, nu;
};
size_t cg_parts_doubles(long long N, long long M);
CgParts cg_parts(double* base, long long N, long long M);

void cg_init(const float2* y, const float* W, long long n, double* part, CgScal* sc, double tol,
             hipStream_t s);
void cg_grad(const float2* P, int nc, long long N, const PenSpec* p, const float2* x,
             float2* ngrad, const CgParts& cp, CgScal* sc, double* resid, hipStream_t s);
void cg_dir(const float2* ngrad, float2* ddir, long long N, const CgParts& cp, CgScal* sc,
            hipStream_t s);
void cg_adir(const float2* Adir, const float2* y, const double2* Ax, const float* W, long long n,
             const PenSpec* p, const float2* ddir, const float2* x, const CgParts& cp,
             CgScal* sc, hipStream_t s);
void cg_linesearch(const PenSpec* p, const float2* ddir, const float2* x, const CgParts& cp,
                   CgScal* sc, hipStream_t s);
void cg_update(double2* Ax, const float2* Adir, const float2* y, const float* W, long long M,
               double2* x, float2* x32, const float2* ddir, long long N, const PenSpec* p,
               const CgParts& cp, float2* rout, CgScal* sc, hipStream_t s);
void cg_finish(const CgParts& cp, CgScal* sc, double* resid, hipStream_t s);

}  // namespace pgv
