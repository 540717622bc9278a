// pg_vec.hip -- elementwise, reduction, penalty and PWLS-PCG kernels.
//
// Vectors are FP32 complex; every inner product, norm, step size and the
// line-search scalars are FP64 and live on the device (struct CgScal), so a
// whole iteration can be captured in a HIP graph.  Reductions are two-stage
// (per-block partials, then summed in a fixed order by the kernel that
// needs them), so results do not depend on scheduling.
#include "pg_vec.h"

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

constexpr int kThreads = 256;

// block-wide sum of NV doubles; result valid in thread 0
template <int NV>
__device__ void block_sum(double (&v)[NV])
{
    __shared__ double sh[NV][kThreads / 64];
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][wv] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0;
            for (int w = 0; w < kThreads / 64; ++w) s += sh[k][w];
            v[k] = s;
        }
}

// ---------------------------------------------------------- elementwise --

__global__ void k_cmul(float2* out, const float2* a, const float2* b, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cmul(a[i], b[i]);
}

// out = (first ? 0 : out) + a .* b  (b conjugated when conj_b)
__global__ void k_cmul_acc(float2* out, const float2* a, const float2* b, long long n, int first,
                           int conj_b)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 bb = b[i];
    if (conj_b) bb.y = -bb.y;
    const float2 p = cmul(a[i], bb);
    if (first) out[i] = p;
    else out[i] = make_float2(out[i].x + p.x, out[i].y + p.y);
}

// out = (first ? 0 : out) + r .* a  with r real
__global__ void k_rmul_acc(float2* out, const float* r, const float2* a, long long n, int first)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 p = make_float2(r[i] * a[i].x, r[i] * a[i].y);
    if (first) out[i] = p;
    else out[i] = make_float2(out[i].x + p.x, out[i].y + p.y);
}

__global__ void k_coil_sum(float2* out, const float2* P, int nc, long long N)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float2 s = P[i];
    for (int c = 1; c < nc; ++c) {
        const float2 p = P[(long long)c * N + i];
        s.x += p.x;
        s.y += p.y;
    }
    out[i] = s;
}

__global__ void k_fill(float2* out, long long n, float2 v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

// r = W .* (y - Ax), formed in FP64 from the FP64 Ax
__global__ void k_resid(float2* r, const float2* y, const double2* Ax, const float* W, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double vr = (double)y[i].x - Ax[i].x, vi = (double)y[i].y - Ax[i].y;
    if (W) {
        vr *= W[i];
        vi *= W[i];
    }
    r[i] = make_float2((float)vr, (float)vi);
}

__global__ void k_f2d(double2* out, const float2* in, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double2(in[i].x, in[i].y);
}

// ------------------------------------------------------------ penalties --
// Robject.cpp:49-120 on the y-fastest image, as per-element stencils.

struct Pen {
    int kind;  // 0 quad, 1 TV
    int nd;
    long long n;
    long long off[3];
    float beta, delta;
};

__device__ __forceinline__ float2 cd_at(const float2* d, long long i, long long off)
{
    // Cd: out[i] = d[i] - d[i-off] for i >= off, else 0 (Robject.cpp:76-79)
    if (i < off) return make_float2(0.f, 0.f);
    const float2 a = d[i], b = d[i - off];
    return make_float2(a.x - b.x, a.y - b.y);
}

// xs = x + step * dir on the fly (dir may be null)
__device__ __forceinline__ float2 xs_at(const float2* x, const float2* dir, double2 step,
                                        long long i)
{
    float2 v = x[i];
    if (dir) {
        const float2 d = dir[i];
        const float sr = (float)step.x, si = (float)step.y;
        v.x += sr * d.x - si * d.y;
        v.y += sr * d.y + si * d.x;
    }
    return v;
}

__device__ __forceinline__ float2 cd_xs(const float2* x, const float2* dir, double2 step,
                                        long long i, long long off)
{
    if (i < off) return make_float2(0.f, 0.f);
    const float2 a = xs_at(x, dir, step, i), b = xs_at(x, dir, step, i - off);
    return make_float2(a.x - b.x, a.y - b.y);
}

__device__ __forceinline__ float wpot(const Pen& p, float2 d)
{
    if (p.kind == 0) return 1.0f;  // QuadPenalty.cpp:45-50
    const float ar = d.x / p.delta, ai = d.y / p.delta;  // TVPenalty.cpp:43-49
    const float t2 = ar * ar + ai * ai;
    return 1.0f / sqrtf(1.0f + t2);
}

__device__ __forceinline__ float2 dpot(const Pen& p, float2 d)
{
    if (p.kind == 0) return d;  // QuadPenalty.cpp:52-56
    const float ar = d.x / p.delta, ai = d.y / p.delta;  // TVPenalty.cpp:51-57
    const float s = sqrtf(1.0f + ar * ar + ai * ai);
    return make_float2(d.x / s, d.y / s);
}

// Gradient(xs)[j] = beta * sum_dims Ctd(dpot(Cd xs))[j]   (Robject.cpp:143-165)
__device__ float2 pen_grad_at(const Pen& p, const float2* x, const float2* dir, double2 step,
                              long long j)
{
    float2 g = make_float2(0.f, 0.f);
    for (int k = 0; k < p.nd; ++k) {
        const long long off = p.off[k], n = p.n;
        float2 o;
        // Ctd (Robject.cpp:111-117): j < n-1-off: d[j]-d[j+off];
        // j == n-1-off: d[j]; j == off-1 (when not covered above): -d[off]
        if (j < n - 1 - off) {
            const float2 a = dpot(p, cd_xs(x, dir, step, j, off));
            const float2 b = dpot(p, cd_xs(x, dir, step, j + off, off));
            o = make_float2(a.x - b.x, a.y - b.y);
        } else if (j == n - 1 - off) {
            o = dpot(p, cd_xs(x, dir, step, j, off));
        } else if (j == off - 1) {
            const float2 b = dpot(p, cd_xs(x, dir, step, off, off));
            o = make_float2(-b.x, -b.y);
        } else {
            o = make_float2(0.f, 0.f);
        }
        g.x += o.x;
        g.y += o.y;
    }
    return make_float2(p.beta * g.x, p.beta * g.y);
}

__global__ void k_pen_gradient(Pen p, const float2* x, float2* g)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < p.n) g[j] = pen_grad_at(p, x, nullptr, make_double2(0, 0), j);
}

// partials of Denom(ddir, xs) (re) and of pdot = Re<ddir, Gradient(xs)>
__global__ void k_pen_denom_pdot(Pen p, const float2* ddir, const float2* x, const double2* stepp,
                                 int want_pdot, double* part, const int* stop)
{
    if (stop && *stop) return;
    const double2 step = stepp ? *stepp : make_double2(0, 0);
    const float2* dir = stepp ? ddir : nullptr;
    double v[2] = {0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (long long)gridDim.x * blockDim.x) {
        for (int k = 0; k < p.nd; ++k) {
            const float2 cdir = cd_at(ddir, i, p.off[k]);
            const float w = wpot(p, cd_xs(x, dir, step, i, p.off[k]));
            v[0] += (double)w * ((double)cdir.x * cdir.x + (double)cdir.y * cdir.y);
        }
        if (want_pdot) {
            const float2 g = pen_grad_at(p, x, dir, step, i);
            const float2 d = ddir[i];
            v[1] += (double)d.x * g.x + (double)d.y * g.y;
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0] * p.beta;
        part[2 * blockIdx.x + 1] = v[1];
    }
}

__global__ void k_pen_value(Pen p, const float2* x, double* part)
{
    double v[1] = {0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (long long)gridDim.x * blockDim.x) {
        for (int k = 0; k < p.nd; ++k) {
            const float2 d = cd_at(x, i, p.off[k]);
            const double a2 = (double)d.x * d.x + (double)d.y * d.y;
            if (p.kind == 0) v[0] += a2 / 2.0;
            else v[0] += (double)p.delta * p.delta * (sqrt(1.0 + a2 / ((double)p.delta * p.delta)) - 1.0);
        }
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// sum partials (nparts x NV) in a fixed order; one block
template <int NV>
__global__ void k_reduce_parts(const double* part, int nparts, double* out)
{
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += part[(long long)b * NV + k];
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = v[k];
}

// ------------------------------------------------------------------ CG ---
// One PCG iteration (solve_pwls_pcg.hpp:83-162) is, after the adjoint:
//   grad   ngrad = sum_c P_c - R.Gradient(x)           partials |g_data|^2, |ngrad|^2
//   dir    zero-gradient exit, gamma; ddir             partial  <ddir, ngrad>
//   (forward A ddir)
//   adir   descent check; dAWAd, dAWr partials; penalty partials at step 0
//   ls1    Newton step 1; penalty partials at step 1
//   ls2    Newton step 2; penalty partials at step 2
//   update Newton step 3; x += step ddir, Ax += step Adir; partial |y - Ax|^2
// Each kernel first sums the previous kernel's partials itself, in every
// block and in the same fixed order, so that every block takes the same
// (deterministic) decision and no single-block control kernel is needed.
// Block 0 records the scalars; a kernel never writes a field it reads.

// this thread's share of the fixed-order sum of nparts x NV partials:
// entries b = threadIdx.x, + blockDim.x, ...; the first is loaded without
// waiting for anything (the usual case: nparts <= blockDim.x)
template <int NV>
__device__ __forceinline__ void load_parts(const double* __restrict__ part, int nparts, double* v)
{
    const int b = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0;
    if (b < nparts)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += part[(long long)b * NV + k];
    for (int bb = b + blockDim.x; bb < nparts; bb += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += part[(long long)bb * NV + k];
}

// block-sum the per-thread shares and give the sums to every thread (each
// value is reduced on its own, so sums fused into one call are the same)
template <int NV>
__device__ void bcast_sum(double (&v)[NV])
{
    __shared__ double bc[NV];
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) bc[k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = bc[k];
}

// sum of nparts x NV partials in a fixed order, result in every thread
template <int NV>
__device__ void all_sum_parts(const double* __restrict__ part, int nparts, double (&out)[NV])
{
    load_parts<NV>(part, nparts, out);
    bcast_sum<NV>(out);
}

// One Newton step of the line search (solve_pwls_pcg.hpp:129-152) from the
// penalty sums {pdenom, pdot} at the current step; returns false and sets
// stop on a bad denominator
__device__ __forceinline__ bool newton_step(double dAWAd, double dAWr, double pdenom, double pdot,
                                            double2& step, int& stop)
{
    const double dr = dAWAd + pdenom;  // Denom is real-valued here
    if (!(fabs(dr) >= 1e-20) || fabs(dr) > 1e25) {
        stop = PG_EXIT_BAD_DENOM;
        return false;
    }
    const double2 s = step;
    const double numr = -dAWr + s.x * dAWAd + pdot;
    const double numi = s.y * dAWAd;
    step = make_double2(s.x - numr / dr, s.y - numi / dr);
    return true;
}

// penalty partials {beta * Denom, Re<ddir, Gradient(x + step ddir)>} of the
// elements i = i0, i0 + stride, ... (block-summed into part[2 * slot])
__device__ void pen_partials(const Pen& p, const float2* ddir, const float2* x, double2 step,
                             long long i0, long long stride, double* part, int slot)
{
    double v[2] = {0, 0};
    for (long long i = i0; i < p.n; i += stride) {
        for (int k = 0; k < p.nd; ++k) {
            const float2 cdir = cd_at(ddir, i, p.off[k]);
            const float w = wpot(p, cd_xs(x, ddir, step, i, p.off[k]));
            v[0] += (double)w * ((double)cdir.x * cdir.x + (double)cdir.y * cdir.y);
        }
        const float2 g = pen_grad_at(p, x, ddir, step, i);
        const float2 d = ddir[i];
        v[1] += (double)d.x * g.x + (double)d.y * g.y;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * slot] = v[0] * p.beta;
        part[2 * slot + 1] = v[1];
    }
}

// quadratic penalty: partials of q = Re<ddir, Gradient(ddir)>, the slope of
// pdot(s) = Re<ddir, Gradient(x + s ddir)> = pdot(0) + s q (the gradient is
// linear; Ctd is not the exact transpose of Cd, so q differs from Denom)
__device__ void pen_slope_partials(const Pen& p, const float2* ddir, long long i0,
                                   long long stride, double* part, int slot)
{
    double v[2] = {0, 0};
    for (long long i = i0; i < p.n; i += stride) {
        const float2 g = pen_grad_at(p, ddir, nullptr, make_double2(0, 0), i);
        const float2 d = ddir[i];
        v[0] += (double)d.x * g.x + (double)d.y * g.y;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * slot] = v[0];
        part[2 * slot + 1] = 0.0;
    }
}

// pen_partials at step 0 and pen_slope_partials in one pass over the
// elements: the same sums, each in the same order, with one block reduction
__device__ void pen_partials_and_slope(const Pen& p, const float2* ddir, const float2* x,
                                       long long i0, long long stride, double* part0,
                                       double* part1, int slot)
{
    double v[4] = {0, 0, 0, 0};
    for (long long i = i0; i < p.n; i += stride) {
        for (int k = 0; k < p.nd; ++k) {
            const float2 cdir = cd_at(ddir, i, p.off[k]);
            const float w = wpot(p, cd_xs(x, ddir, make_double2(0, 0), i, p.off[k]));
            v[0] += (double)w * ((double)cdir.x * cdir.x + (double)cdir.y * cdir.y);
        }
        const float2 g = pen_grad_at(p, x, ddir, make_double2(0, 0), i);
        const float2 gs = pen_grad_at(p, ddir, nullptr, make_double2(0, 0), i);
        const float2 d = ddir[i];
        v[1] += (double)d.x * g.x + (double)d.y * g.y;
        v[2] += (double)d.x * gs.x + (double)d.y * gs.y;
    }
    block_sum<4>(v);
    if (threadIdx.x == 0) {
        part0[2 * slot] = v[0] * p.beta;
        part0[2 * slot + 1] = v[1];
        part1[2 * slot] = v[2];
        part1[2 * slot + 1] = 0.0;
    }
}

// grad: also closes the previous iteration's bookkeeping (residual norm,
// iteration count) from its update partials
__global__ void k_cg_grad(const float2* P, int nc, long long N, Pen p, int have_pen,
                          const float2* x, float2* ngrad, pgv::CgParts cp, CgScal* sc,
                          double* resid)
{
    if (blockIdx.x == 0 && sc->pending) {
        double u[1];
        all_sum_parts<1>(cp.u, cp.nu, u);
        if (threadIdx.x == 0) {
            resid[sc->ii] = sqrt(u[0]);
            sc->iter += 1;
            sc->ii += 1;
            sc->pending = 0;
        }
    }
    if (sc->stop) return;
    double v[2] = {0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x) {
        // the penalty stencil's loads and each run of 8 coils' loads are
        // issued together; the coils are still added in coil order
        const float2 r = have_pen ? pen_grad_at(p, x, nullptr, make_double2(0, 0), i)
                                  : make_float2(0.f, 0.f);
        float2 g = P[i];
        for (int c0 = 1; c0 < nc; c0 += 8) {
            float2 q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                q[k] = (c0 + k < nc) ? P[(long long)(c0 + k) * N + i] : make_float2(0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (c0 + k < nc) {
                    g.x += q[k].x;
                    g.y += q[k].y;
                }
        }
        v[0] += (double)g.x * g.x + (double)g.y * g.y;
        if (have_pen) {
            g.x -= r.x;
            g.y -= r.y;
        }
        ngrad[i] = g;
        v[1] += (double)g.x * g.x + (double)g.y * g.y;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        cp.g[2 * blockIdx.x] = v[0];
        cp.g[2 * blockIdx.x + 1] = v[1];
    }
}

// dir: zero-gradient exit and gamma (solve_pwls_pcg.hpp:88-107), then
// ddir = ngrad + gamma ddir (ngrad on the first iteration); partial
// real(cdot(ddir, ngrad))
__global__ void k_cg_dir(const float2* ngrad, float2* ddir, long long N, pgv::CgParts cp,
                         CgScal* sc)
{
    // the first element's inputs, the partials and the stop flag are loaded
    // before any of them is used
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float2 g = make_float2(0.f, 0.f), o = g;
    if (i < N) {
        g = ngrad[i];
        o = ddir[i];
    }
    double gs[2];
    load_parts<2>(cp.g, cp.ng, gs);
    if (sc->stop) return;
    bcast_sum<2>(gs);
    const double gnorm = sqrt(gs[0]);
    if (gnorm / sc->ywy < 1e-10) {  // norm_grad (:41-46); NaN compares false
        if (blockIdx.x == 0 && threadIdx.x == 0) sc->stop = PG_EXIT_ZERO_GRAD;
        return;
    }
    const double newinprod = gs[1];
    const int first = (sc->ii == 0);
    // the optional tolerance exit (no counterpart in the reference): the
    // gradient, penalty term included, down to tol of the first iteration's
    if (!first && sc->tol > 0.0 && newinprod <= sc->tol * sc->tol * sc->g0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) sc->stop = PG_EXIT_TOL;
        return;
    }
    const double gam = first ? 0.0 : ((fabs(sc->oldinprod) < 1e-10) ? 0.0 : newinprod / sc->oldinprod);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->newinprod = newinprod;
        sc->gamma = gam;
        sc->first = first;
        if (first) sc->g0 = newinprod;
    }
    const float gamf = (float)gam;
    double v[1] = {0};
    while (i < N) {
        const long long j = i + stride;
        float2 g2 = make_float2(0.f, 0.f), o2 = g2;
        if (j < N) {
            g2 = ngrad[j];
            o2 = ddir[j];
        }
        float2 d = g;
        if (!first) d = make_float2(g.x + gamf * o.x, g.y + gamf * o.y);
        ddir[i] = d;
        v[0] += (double)d.x * g.x + (double)d.y * g.y;
        i = j;
        g = g2;
        o = o2;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) cp.d[blockIdx.x] = v[0];
}

// adir: descent check (:113-116); blocks [0, np) form the penalty partials
// at step 0 (first, so that their longer stencil work starts first), blocks
// [np, np + na) the partials of dAWAd = real(cdot(Adir, W Adir)) and
// dAWr = real(Adir' W (y - Ax)) (:123-125)
__global__ void k_cg_adir(const float2* Adir, const float2* y, const double2* Ax, const float* W,
                          long long n, int na, Pen p, int have_pen, const float2* ddir,
                          const float2* x, pgv::CgParts cp, CgScal* sc)
{
    // the first sample's inputs, the partials and the stop flag are loaded
    // before any of them is used; each later sample's before the previous
    // one is summed
    const int np = (int)gridDim.x - na;
    const int db = (int)blockIdx.x - np;  // data block index
    const bool data_block = db >= 0;
    const long long stride = (long long)na * blockDim.x;
    long long i = (long long)db * blockDim.x + threadIdx.x;
    float2 a = make_float2(0.f, 0.f), yv = a;
    double2 ax = make_double2(0, 0);
    float w = 1.f;
    if (data_block && i < n) {
        a = Adir[i];
        yv = y[i];
        ax = Ax[i];
        if (W) w = W[i];
    }
    double dd[1];
    load_parts<1>(cp.d, cp.nd, dd);
    if (sc->stop) return;
    bcast_sum<1>(dd);
    if (dd[0] < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) sc->stop = PG_EXIT_NOT_DESCENT;
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->oldinprod = sc->newinprod;
    if (!data_block) {
        if (have_pen) {
            const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
            const long long pstride = (long long)np * blockDim.x;
            // the quadratic penalty's line search is closed-form in the update
            // kernel; the step-1 slots are free for the slope partials
            if (p.kind == 0)
                pen_partials_and_slope(p, ddir, x, i0, pstride, cp.l[0], cp.l[1], blockIdx.x);
            else
                pen_partials(p, ddir, x, make_double2(0, 0), i0, pstride, cp.l[0], blockIdx.x);
        }
        return;
    }
    double v[2] = {0, 0};
    while (i < n) {
        const long long j = i + stride;
        float2 a2 = make_float2(0.f, 0.f), y2 = a2;
        double2 ax2 = make_double2(0, 0);
        float w2 = 1.f;
        if (j < n) {
            a2 = Adir[j];
            y2 = y[j];
            ax2 = Ax[j];
            if (W) w2 = W[j];
        }
        const double wd = W ? (double)w : 1.0;
        const double rr = (double)yv.x - ax.x, ri = (double)yv.y - ax.y;
        v[0] += wd * ((double)a.x * a.x + (double)a.y * a.y);
        v[1] += wd * ((double)a.x * rr + (double)a.y * ri);
        i = j;
        a = a2;
        yv = y2;
        ax = ax2;
        w = w2;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        cp.a[2 * db] = v[0];
        cp.a[2 * db + 1] = v[1];
    }
}

// ls j (j = 1, 2): Newton step j from the penalty partials at steps[j-1],
// then the penalty partials at steps[j]
__global__ void k_cg_ls(int j, Pen p, const float2* ddir, const float2* x, pgv::CgParts cp,
                        CgScal* sc)
{
    if (sc->stop) return;
    double dAWAd, dAWr;
    if (j == 1) {
        double a[2];
        all_sum_parts<2>(cp.a, cp.na, a);
        dAWAd = a[0];
        dAWr = a[1];
    } else {
        dAWAd = sc->dAWAd;
        dAWr = sc->dAWr;
    }
    double pl[2];
    all_sum_parts<2>(cp.l[j - 1], cp.nl, pl);
    double2 step = (j == 1) ? make_double2(0, 0) : sc->steps[j - 1];
    int stop = 0;
    const bool ok = newton_step(dAWAd, dAWr, pl[0], pl[1], step, stop);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (j == 1) {
            sc->dAWAd = dAWAd;
            sc->dAWr = dAWr;
        }
        if (ok) sc->steps[j] = step;
        else sc->stop = stop;
    }
    if (!ok) return;
    pen_partials(p, ddir, x, step, (long long)blockIdx.x * blockDim.x + threadIdx.x,
                 (long long)gridDim.x * blockDim.x, cp.l[j], blockIdx.x);
}

// update: the last Newton step, then Ax += step Adir, x += step ddir in FP64
// (x32: the FP32 copy for the penalty stencils); partial ||y - Ax||^2
// (:155-162).  pen_mode 0 (no penalty) and 2 (quadratic penalty) take all
// three Newton steps here: for a quadratic penalty Denom(ddir, x + s ddir)
// is the same for every s and Re<ddir, Gradient(x + s ddir)> = pdot(0) +
// s q (the step stays real), from the step-0 penalty and slope partials.
// pen_mode 1 (TV) takes the third step from the partials at steps[2].  With
// rout set, also the next adjoint's input W (y - Ax) from the new Ax, in FP32
// (the rounding of the spread's own value pass).
__global__ void k_cg_update(double2* Ax, const float2* Adir, const float2* y, const float* W,
                            long long M, double2* x, float2* x32, const float2* ddir, long long N,
                            int pen_mode, pgv::CgParts cp, float2* rout, CgScal* sc)
{
    // the first element's inputs, the partials and the stop flag are all
    // loaded before any of them is used; each later element's inputs are
    // loaded before the previous element is stored
    const long long n = M > N ? M : N;
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float2 a = make_float2(0.f, 0.f), yv = a, d = a;
    double2 ax = make_double2(0, 0), xo = ax;
    float w = 1.f;
    if (i < M) {
        a = Adir[i];
        ax = Ax[i];
        yv = y[i];
        if (W) w = W[i];
    }
    if (i < N) {
        d = ddir[i];
        xo = x[i];
    }
    const int have_pen = (pen_mode == 1);
    double q[6] = {0, 0, 0, 0, 0, 0};
    if (pen_mode == 2) {
        load_parts<2>(cp.a, cp.na, q);
        load_parts<2>(cp.l[0], cp.nl, q + 2);
        load_parts<2>(cp.l[1], cp.nl, q + 4);
    } else if (have_pen) {
        load_parts<2>(cp.l[2], cp.nl, q);
    } else {
        load_parts<2>(cp.a, cp.na, q);
    }
    if (sc->stop) return;
    bcast_sum<6>(q);
    double2 step;
    int stop = 0;
    bool ok = true;
    double dAWAd = 0, dAWr = 0;
    if (pen_mode == 2) {
        // q: {dAWAd, dAWr}, penalty {denom, pdot} at step 0, slope {q, -}
        dAWAd = q[0];
        dAWr = q[1];
        step = make_double2(0, 0);
        for (int j = 1; j <= 3 && ok; ++j)
            ok = newton_step(dAWAd, dAWr, q[2], q[3] + step.x * q[4], step, stop);
    } else if (have_pen) {
        dAWAd = sc->dAWAd;
        dAWr = sc->dAWr;
        step = sc->steps[2];
        ok = newton_step(dAWAd, dAWr, q[0], q[1], step, stop);
    } else {
        dAWAd = q[0];
        dAWr = q[1];
        step = make_double2(0, 0);
        for (int j = 1; j <= 3 && ok; ++j) ok = newton_step(dAWAd, dAWr, 0.0, 0.0, step, stop);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!have_pen) {
            sc->dAWAd = dAWAd;
            sc->dAWr = dAWr;
        }
        if (ok) {
            sc->steps[3] = step;
            sc->step = step;
            sc->pending = 1;
        } else {
            sc->stop = stop;
        }
    }
    if (!ok) return;
    const double sr = step.x, si = step.y;
    double v[1] = {0};
    while (i < n) {
        const long long j = i + stride;
        float2 a2 = make_float2(0.f, 0.f), y2 = a2, d2 = a2;
        double2 ax2 = make_double2(0, 0), x2 = ax2;
        float w2 = 1.f;
        if (j < M) {
            a2 = Adir[j];
            ax2 = Ax[j];
            y2 = y[j];
            if (W) w2 = W[j];
        }
        if (j < N) {
            d2 = ddir[j];
            x2 = x[j];
        }
        if (i < M) {
            ax.x += sr * a.x - si * a.y;
            ax.y += sr * a.y + si * a.x;
            Ax[i] = ax;
            const double rr = (double)yv.x - ax.x, ri = (double)yv.y - ax.y;
            v[0] += rr * rr + ri * ri;
            if (rout) {
                double wr = rr, wi = ri;
                if (W) {
                    const double wv = w;
                    wr *= wv;
                    wi *= wv;
                }
                rout[i] = make_float2((float)wr, (float)wi);
            }
        }
        if (i < N) {
            xo.x += sr * d.x - si * d.y;
            xo.y += sr * d.y + si * d.x;
            x[i] = xo;
            x32[i] = make_float2((float)xo.x, (float)xo.y);
        }
        i = j;
        a = a2;
        ax = ax2;
        yv = y2;
        w = w2;
        d = d2;
        xo = x2;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) cp.u[blockIdx.x] = v[0];
}

// closes the bookkeeping of the last completed iteration
__global__ void k_cg_finish(pgv::CgParts cp, CgScal* sc, double* resid)
{
    if (!sc->pending) return;
    double u[1];
    all_sum_parts<1>(cp.u, cp.nu, u);
    if (threadIdx.x == 0) {
        resid[sc->ii] = sqrt(u[0]);
        sc->iter += 1;
        sc->ii += 1;
        sc->pending = 0;
    }
}

// yi' W yi (solve_pwls_pcg.hpp:44)
__global__ void k_wnorm2(const float2* y, const float* W, long long n, double* part)
{
    double v[1] = {0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double w = W ? (double)W[i] : 1.0;
        v[0] += w * ((double)y[i].x * y[i].x + (double)y[i].y * y[i].y);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void k_cg_init(const double* part, int nparts, CgScal* sc, double tol)
{
    double v[1] = {0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) v[0] += part[b];
    block_sum<1>(v);
    if (threadIdx.x != 0) return;
    CgScal z = {};
    z.ywy = v[0];
    z.tol = tol;
    *sc = z;
}

inline unsigned nblk(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline unsigned nred(long long n)
{
    const long long b = (n + kThreads - 1) / kThreads;
    const long long cap = n >= pgv::kBigParts ? pgv::kMaxParts : pgv::kSmallParts;
    return (unsigned)(b < cap ? (b < 1 ? 1 : b) : cap);
}

Pen to_pen(const pgv::PenSpec& s)
{
    Pen p;
    p.kind = s.kind;
    p.nd = s.nd;
    p.n = s.n;
    p.off[0] = s.off[0];
    p.off[1] = s.off[1];
    p.off[2] = s.off[2];
    p.beta = s.beta;
    p.delta = s.delta;
    return p;
}

}  // namespace

namespace pgv {

void cmul(float2* out, const float2* a, const float2* b, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_cmul, dim3(nblk(n)), dim3(kThreads), 0, s, out, a, b, n);
}
void cmul_acc(float2* out, const float2* a, const float2* b, long long n, bool first, bool conj_b,
              hipStream_t s)
{
    hipLaunchKernelGGL(k_cmul_acc, dim3(nblk(n)), dim3(kThreads), 0, s, out, a, b, n, (int)first,
                       (int)conj_b);
}
void rmul_acc(float2* out, const float* r, const float2* a, long long n, bool first, hipStream_t s)
{
    hipLaunchKernelGGL(k_rmul_acc, dim3(nblk(n)), dim3(kThreads), 0, s, out, r, a, n, (int)first);
}
void coil_sum(float2* out, const float2* P, int nc, long long N, hipStream_t s)
{
    hipLaunchKernelGGL(k_coil_sum, dim3(nblk(N)), dim3(kThreads), 0, s, out, P, nc, N);
}
void fill(float2* out, long long n, float2 v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(kThreads), 0, s, out, n, v);
}
void resid(float2* r, const float2* y, const double2* Ax, const float* W, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_resid, dim3(nblk(n)), dim3(kThreads), 0, s, r, y, Ax, W, n);
}
void f2d(double2* out, const float2* in, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f2d, dim3(nblk(n)), dim3(kThreads), 0, s, out, in, n);
}

void pen_gradient(const PenSpec& ps, const float2* x, float2* g, hipStream_t s)
{
    hipLaunchKernelGGL(k_pen_gradient, dim3(nblk(ps.n)), dim3(kThreads), 0, s, to_pen(ps), x, g);
}
void pen_denom_pdot(const PenSpec& ps, const float2* ddir, const float2* x, const double2* step,
                    bool want_pdot, double* part, double* out2, const int* stop, hipStream_t s)
{
    const unsigned nb = nred(ps.n);
    hipLaunchKernelGGL(k_pen_denom_pdot, dim3(nb), dim3(kThreads), 0, s, to_pen(ps), ddir, x, step,
                       (int)want_pdot, part, stop);
    if (out2) hipLaunchKernelGGL(k_reduce_parts<2>, dim3(1), dim3(kThreads), 0, s, part, (int)nb, out2);
}
void pen_value(const PenSpec& ps, const float2* x, double* part, double* out, hipStream_t s)
{
    const unsigned nb = nred(ps.n);
    hipLaunchKernelGGL(k_pen_value, dim3(nb), dim3(kThreads), 0, s, to_pen(ps), x, part);
    hipLaunchKernelGGL(k_reduce_parts<1>, dim3(1), dim3(kThreads), 0, s, part, (int)nb, out);
}

CgParts cg_parts_alloc_sizes(long long N, long long M, int* total)
{
    CgParts c{};
    c.ng = (int)nred(N);
    c.nd = (int)nred(N);
    c.na = (int)nred(M);
    c.nl = (int)nred(N);
    c.nu = (int)nred(M > N ? M : N);
    *total = 2 * c.ng + c.nd + 2 * c.na + 3 * 2 * c.nl + c.nu;
    return c;
}

CgParts cg_parts(double* base, long long N, long long M)
{
    int total = 0;
    CgParts c = cg_parts_alloc_sizes(N, M, &total);
    double* q = base;
    c.g = q;
    q += 2 * c.ng;
    c.d = q;
    q += c.nd;
    c.a = q;
    q += 2 * c.na;
    for (int j = 0; j < 3; ++j) {
        c.l[j] = q;
        q += 2 * c.nl;
    }
    c.u = q;
    return c;
}

size_t cg_parts_doubles(long long N, long long M)
{
    int total = 0;
    cg_parts_alloc_sizes(N, M, &total);
    return (size_t)total;
}

void cg_init(const float2* y, const float* W, long long n, double* part, CgScal* sc, double tol,
             hipStream_t s)
{
    const unsigned nb = nred(n);
    hipLaunchKernelGGL(k_wnorm2, dim3(nb), dim3(kThreads), 0, s, y, W, n, part);
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(kThreads), 0, s, part, (int)nb, sc, tol);
}

void cg_grad(const float2* P, int nc, long long N, const PenSpec* ps, const float2* x,
             float2* ngrad, const CgParts& cp, CgScal* sc, double* resid, hipStream_t s)
{
    Pen p = ps ? to_pen(*ps) : Pen{};
    hipLaunchKernelGGL(k_cg_grad, dim3(cp.ng), dim3(kThreads), 0, s, P, nc, N, p, ps ? 1 : 0, x,
                       ngrad, cp, sc, resid);
}

void cg_dir(const float2* ngrad, float2* ddir, long long N, const CgParts& cp, CgScal* sc,
            hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_dir, dim3(cp.nd), dim3(kThreads), 0, s, ngrad, ddir, N, cp, sc);
}

void cg_adir(const float2* Adir, const float2* y, const double2* Ax, const float* W, long long n,
             const PenSpec* ps, const float2* ddir, const float2* x, const CgParts& cp,
             CgScal* sc, hipStream_t s)
{
    Pen p = ps ? to_pen(*ps) : Pen{};
    const int np = ps ? cp.nl : 0;
    hipLaunchKernelGGL(k_cg_adir, dim3(cp.na + np), dim3(kThreads), 0, s, Adir, y, Ax, W, n, cp.na,
                       p, ps ? 1 : 0, ddir, x, cp, sc);
}

void cg_linesearch(const PenSpec* ps, const float2* ddir, const float2* x, const CgParts& cp,
                   CgScal* sc, hipStream_t s)
{
    // without a penalty, or with a quadratic one, the update kernel takes all
    // three steps
    if (!ps || ps->kind == 0) return;
    for (int j = 1; j <= 2; ++j)
        hipLaunchKernelGGL(k_cg_ls, dim3(cp.nl), dim3(kThreads), 0, s, j, to_pen(*ps), ddir, x, cp,
                           sc);
}

void cg_update(double2* Ax, const float2* Adir, const float2* y, const float* W, long long M,
               double2* x, float2* x32, const float2* ddir, long long N, const PenSpec* ps,
               const CgParts& cp, float2* rout, CgScal* sc, hipStream_t s)
{
    const int pen_mode = !ps ? 0 : (ps->kind == 0 ? 2 : 1);
    hipLaunchKernelGGL(k_cg_update, dim3(cp.nu), dim3(kThreads), 0, s, Ax, Adir, y, W, M, x, x32,
                       ddir, N, pen_mode, cp, rout, sc);
}

void cg_finish(const CgParts& cp, CgScal* sc, double* resid, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_finish, dim3(1), dim3(kThreads), 0, s, cp, sc, resid);
}

}  // namespace pgv
