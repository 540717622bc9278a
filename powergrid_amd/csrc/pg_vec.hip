// pg_vec.hip -- elementwise, reduction, penalty and PWLS-PCG kernels.
//
// Vectors are FP32 complex; every inner product, norm, step size and the
// line-search scalars are FP64 and live on the device (struct CgScal), so a
// whole iteration can be captured in a HIP graph.  Reductions are two-stage
// (per-block partials, then one block summing them in a fixed order), so
// results do not depend on scheduling.
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

// gradient assembly: g_data = sum_c P[c] (the SENSE coil sum), partial
// ||g_data||^2; ngrad = g_data - R.Gradient(x); partial ||ngrad||^2
// (solve_pwls_pcg.hpp:83-95)
__global__ void k_cg_grad(const float2* P, int nc, long long N, Pen p, int have_pen,
                          const float2* x, float2* ngrad, double* part, const CgScal* sc)
{
    if (sc->stop) return;
    double v[2] = {0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x) {
        float2 g = P[i];
        for (int c = 1; c < nc; ++c) {
            const float2 q = P[(long long)c * N + i];
            g.x += q.x;
            g.y += q.y;
        }
        v[0] += (double)g.x * g.x + (double)g.y * g.y;
        if (have_pen) {
            const float2 r = pen_grad_at(p, x, nullptr, make_double2(0, 0), i);
            g.x -= r.x;
            g.y -= r.y;
        }
        ngrad[i] = g;
        v[1] += (double)g.x * g.x + (double)g.y * g.y;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// control 1: zero-gradient exit, gamma (solve_pwls_pcg.hpp:88-107)
__global__ void k_cg_ctl_grad(const double* part, int nparts, CgScal* sc)
{
    double v[2] = {0, 0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        v[0] += part[2 * b];
        v[1] += part[2 * b + 1];
    }
    block_sum<2>(v);
    if (threadIdx.x != 0 || sc->stop) return;
    const double gnorm = sqrt(v[0]);
    if (gnorm / sc->ywy < 1e-10) {  // norm_grad (:41-46); NaN compares false
        sc->stop = PG_EXIT_ZERO_GRAD;
        return;
    }
    const double newinprod = v[1];
    if (sc->ii == 0) {
        sc->gamma = 0;
        sc->first = 1;
    } else {
        sc->first = 0;
        sc->gamma = (fabs(sc->oldinprod) < 1e-10) ? 0.0 : newinprod / sc->oldinprod;
    }
    sc->oldinprod = newinprod;
}

// ddir = ngrad + gamma*ddir (ddir = ngrad on the first iteration); partial
// real(cdot(ddir, ngrad))
__global__ void k_cg_dir(const float2* ngrad, float2* ddir, long long N, double* part,
                         const CgScal* sc)
{
    if (sc->stop) return;
    const float gam = (float)sc->gamma;
    const int first = sc->first;
    double v[1] = {0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x) {
        const float2 g = ngrad[i];
        float2 d = g;
        if (!first) {
            const float2 o = ddir[i];
            d = make_float2(g.x + gam * o.x, g.y + gam * o.y);
        }
        ddir[i] = d;
        v[0] += (double)d.x * g.x + (double)d.y * g.y;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// control 2: descent check (solve_pwls_pcg.hpp:113-116)
__global__ void k_cg_ctl_dir(const double* part, int nparts, CgScal* sc)
{
    double v[1] = {0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) v[0] += part[b];
    block_sum<1>(v);
    if (threadIdx.x != 0 || sc->stop) return;
    if (v[0] < 0) sc->stop = PG_EXIT_NOT_DESCENT;
}

// partials of dAWAd = real(cdot(Adir, W Adir)) and
// dAWr = real(Adir' (W (y - Ax)))   (solve_pwls_pcg.hpp:123-125)
__global__ void k_cg_adir_dots(const float2* Adir, const float2* y, const double2* Ax,
                               const float* W, long long n, double* part, const CgScal* sc)
{
    if (sc->stop) return;
    double v[2] = {0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float2 a = Adir[i];
        const double w = W ? (double)W[i] : 1.0;
        const double rr = (double)y[i].x - Ax[i].x, ri = (double)y[i].y - Ax[i].y;
        v[0] += w * ((double)a.x * a.x + (double)a.y * a.y);
        v[1] += w * ((double)a.x * rr + (double)a.y * ri);
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

__global__ void k_cg_ctl_adir(const double* part, int nparts, CgScal* sc)
{
    double v[2] = {0, 0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        v[0] += part[2 * b];
        v[1] += part[2 * b + 1];
    }
    block_sum<2>(v);
    if (threadIdx.x != 0 || sc->stop) return;
    sc->dAWAd = v[0];
    sc->dAWr = v[1];
    sc->step = make_double2(0, 0);
}

// one Newton step of the line search (solve_pwls_pcg.hpp:129-152):
// denom = dAWAd + pdenom; step -= (-dAWr + step*dAWAd + pdot) / denom
__global__ void k_cg_ctl_ls(const double* part, int nparts, int have_pen, CgScal* sc)
{
    double v[2] = {0, 0};
    if (have_pen)
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
            v[0] += part[2 * b];
            v[1] += part[2 * b + 1];
        }
    block_sum<2>(v);
    if (threadIdx.x != 0 || sc->stop) return;
    const double dr = sc->dAWAd + v[0];  // Denom is real-valued here
    if (!(fabs(dr) >= 1e-20) || fabs(dr) > 1e25) {
        sc->stop = PG_EXIT_BAD_DENOM;
        return;
    }
    const double2 s = sc->step;
    const double numr = -sc->dAWr + s.x * sc->dAWAd + v[1];
    const double numi = s.y * sc->dAWAd;
    sc->step = make_double2(s.x - numr / dr, s.y - numi / dr);
}

// Ax += step*Adir, x += step*ddir in FP64 (x32 keeps an FP32 copy of x for
// the penalty stencils); partial ||y - Ax||^2 (solve_pwls_pcg.hpp:160-162)
__global__ void k_cg_update(double2* Ax, const float2* Adir, const float2* y, long long M,
                            double2* x, float2* x32, const float2* ddir, long long N, double* part,
                            const CgScal* sc)
{
    if (sc->stop) return;
    const double sr = sc->step.x, si = sc->step.y;
    double v[1] = {0};
    const long long n = M > N ? M : N;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < M) {
            const float2 a = Adir[i];
            double2 o = Ax[i];
            o.x += sr * a.x - si * a.y;
            o.y += sr * a.y + si * a.x;
            Ax[i] = o;
            const double rr = (double)y[i].x - o.x, ri = (double)y[i].y - o.y;
            v[0] += rr * rr + ri * ri;
        }
        if (i < N) {
            const float2 d = ddir[i];
            double2 o = x[i];
            o.x += sr * d.x - si * d.y;
            o.y += sr * d.y + si * d.x;
            x[i] = o;
            x32[i] = make_float2((float)o.x, (float)o.y);
        }
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void k_cg_ctl_update(const double* part, int nparts, CgScal* sc, double* resid)
{
    double v[1] = {0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) v[0] += part[b];
    block_sum<1>(v);
    if (threadIdx.x != 0) return;
    if (!sc->stop) {
        resid[sc->ii] = sqrt(v[0]);
        sc->iter += 1;
        sc->ii += 1;
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

__global__ void k_cg_init(const double* part, int nparts, CgScal* sc)
{
    double v[1] = {0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) v[0] += part[b];
    block_sum<1>(v);
    if (threadIdx.x != 0) return;
    CgScal z = {};
    z.ywy = v[0];
    *sc = z;
}

inline unsigned nblk(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline unsigned nred(long long n)
{
    long long b = (n + kThreads - 1) / kThreads;
    return (unsigned)(b < pgv::kMaxParts ? (b < 1 ? 1 : b) : pgv::kMaxParts);
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

void cg_init(const float2* y, const float* W, long long n, double* part, CgScal* sc, hipStream_t s)
{
    const unsigned nb = nred(n);
    hipLaunchKernelGGL(k_wnorm2, dim3(nb), dim3(kThreads), 0, s, y, W, n, part);
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(kThreads), 0, s, part, (int)nb, sc);
}

void cg_grad(const float2* P, int nc, long long N, const PenSpec* ps, const float2* x,
             float2* ngrad, double* part, CgScal* sc, hipStream_t s)
{
    const unsigned nb = nred(N);
    Pen p = ps ? to_pen(*ps) : Pen{};
    hipLaunchKernelGGL(k_cg_grad, dim3(nb), dim3(kThreads), 0, s, P, nc, N, p, ps ? 1 : 0, x, ngrad,
                       part, sc);
    hipLaunchKernelGGL(k_cg_ctl_grad, dim3(1), dim3(kThreads), 0, s, part, (int)nb, sc);
}

void cg_dir(const float2* ngrad, float2* ddir, long long N, double* part, CgScal* sc,
            hipStream_t s)
{
    const unsigned nb = nred(N);
    hipLaunchKernelGGL(k_cg_dir, dim3(nb), dim3(kThreads), 0, s, ngrad, ddir, N, part, sc);
    hipLaunchKernelGGL(k_cg_ctl_dir, dim3(1), dim3(kThreads), 0, s, part, (int)nb, sc);
}

void cg_adir(const float2* Adir, const float2* y, const double2* Ax, const float* W, long long n,
             double* part, CgScal* sc, hipStream_t s)
{
    const unsigned nb = nred(n);
    hipLaunchKernelGGL(k_cg_adir_dots, dim3(nb), dim3(kThreads), 0, s, Adir, y, Ax, W, n, part, sc);
    hipLaunchKernelGGL(k_cg_ctl_adir, dim3(1), dim3(kThreads), 0, s, part, (int)nb, sc);
}

void cg_linesearch(const PenSpec* ps, const float2* ddir, const float2* x, double* part,
                   CgScal* sc, hipStream_t s)
{
    for (int j = 0; j < 3; ++j) {
        unsigned nb = 1;
        if (ps) {
            nb = nred(ps->n);
            hipLaunchKernelGGL(k_pen_denom_pdot, dim3(nb), dim3(kThreads), 0, s, to_pen(*ps), ddir, x,
                               (const double2*)&sc->step, 1, part, &sc->stop);
        }
        hipLaunchKernelGGL(k_cg_ctl_ls, dim3(1), dim3(kThreads), 0, s, part, (int)nb, ps ? 1 : 0, sc);
    }
}

void cg_update(double2* Ax, const float2* Adir, const float2* y, long long M, double2* x,
               float2* x32, const float2* ddir, long long N, double* part, CgScal* sc,
               double* resid, hipStream_t s)
{
    const unsigned nb = nred(M > N ? M : N);
    hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(kThreads), 0, s, Ax, Adir, y, M, x, x32, ddir, N,
                       part, sc);
    hipLaunchKernelGGL(k_cg_ctl_update, dim3(1), dim3(kThreads), 0, s, part, (int)nb, sc, resid);
}

}  // namespace pgv
