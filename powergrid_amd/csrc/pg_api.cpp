// pg_api.cpp -- C ABI (include/powergrid_amd.h): operator handles with the
// reference's constructor arguments, forward/adjoint dispatch (batched or
// one NUFFT at a time), penalties and the device-resident PWLS-PCG.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "pg_dft.h"
#include "pg_gnufft.h"
#include "pg_internal.h"
#include "pg_recon.h"
#include "pg_herm.h"
#include "pg_tsint.h"
#include "pg_vec.h"

#include <complex>

namespace pg {

static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }

// ----------------------------------------------------------- profiling --
struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};
static bool g_prof = false;
static std::vector<ProfRec> g_prof_recs;
static std::mutex g_prof_mu;

// HIP refuses calls that use the legacy stream (hipMemcpy, null-stream
// launches and synchronisations) from any thread while a stream capture is
// open.  The library's own set-up copies and kernels therefore go on a
// per-thread non-blocking stream (setup_stream), and graph capture and the
// remaining legacy-stream calls (host-buffer operator calls, copies that the
// caller asks for on the null stream, stream and graph creation and
// destruction) hold this lock.  Threads may then build, solve and free their
// operators concurrently; the device work on their own streams (graph
// launches, asynchronous copies) does not take it.
static std::recursive_mutex g_legacy_mu;
using LegacyLock = std::lock_guard<std::recursive_mutex>;

// the calling thread's stream for set-up copies and kernels: non-blocking, so
// that set-up never touches the legacy stream (created on first use, kept
// for the thread's life)
static hipStream_t setup_stream()
{
    static thread_local hipStream_t st = nullptr;
    if (!st) {
        LegacyLock lk(g_legacy_mu);
        PG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    }
    return st;
}

// Tracing ranges, the reference's RANGE markers (Tracer.hpp:19-34, e.g.
// SENSE.cpp:54,78, TimeSegmentation.cpp:256,301, gridding.cpp:37-595): with
// pg_set_tuning("trace", 1) operator applications, their launch stages,
// solves and set-up steps push roctx ranges, which rocprofv3 --marker-trace
// records.  libroctx64 is opened at run time on first use; without it
// nothing is pushed.  Ranges are host-side: they bracket the enqueueing of
// work, which is what the reference's NVTX ranges mark as well.
static std::atomic<int> g_trace{0};
static std::atomic<unsigned long long> g_trace_pushed{0};
struct RoctxApi {
    int (*push)(const char*) = nullptr;
    int (*pop)() = nullptr;
};
static const RoctxApi& roctx_api()
{
    static const RoctxApi api = [] {
        RoctxApi a;
        for (const char* name : {"libroctx64.so.4", "libroctx64.so", "/opt/rocm/lib/libroctx64.so.4"}) {
            void* h = dlopen(name, RTLD_NOW | RTLD_LOCAL);
            if (!h) continue;
            a.push = reinterpret_cast<int (*)(const char*)>(dlsym(h, "roctxRangePushA"));
            a.pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
            if (a.push && a.pop) break;
            a = RoctxApi{};
        }
        return a;
    }();
    return api;
}
struct TraceRange {
    bool on = false;
    explicit TraceRange(const char* name)
    {
        if (!g_trace.load(std::memory_order_relaxed)) return;
        const RoctxApi& a = roctx_api();
        if (!a.push) return;
        a.push(name);
        g_trace_pushed.fetch_add(1, std::memory_order_relaxed);
        on = true;
    }
    ~TraceRange()
    {
        if (on) roctx_api().pop();
    }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

struct Prof {
    bool on = false;
    hipStream_t s;
    ProfRec r;
    TraceRange tr;
    Prof(const char* name, hipStream_t st) : s(st), tr(name)
    {
        if (!g_prof) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hipStreamIsCapturing(st, &cs);
        if (cs != hipStreamCaptureStatusNone) return;
        on = true;
        r.name = name;
        hipEventCreate(&r.a);
        hipEventCreate(&r.b);
        hipEventRecord(r.a, s);
    }
    ~Prof()
    {
        if (!on) return;
        hipEventRecord(r.b, s);
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_prof_recs.push_back(r);
    }
};

// ---------------------------------------------------------------- FFT --
bool fft_length_supported(int n)
{
    if (n < 1) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    for (int r : {5, 7, 11, 13})
        while (n % r == 0) n /= r;
    return n == 1;
}

FftPlan1D make_fft_plan(int n)
{
    PG_REQUIRE(fft_length_supported(n), PG_ERR_UNSUPPORTED,
               "FFT length " + std::to_string(n) + " has a prime factor above 13");
    FftPlan1D p;
    p.n = n;
    int rem = n;
    while (rem > 1) {
        PG_REQUIRE(p.nst < kMaxStages, PG_ERR_UNSUPPORTED, "FFT length too large");
        int r = 13;
        for (int c : {4, 2, 3, 5, 7, 11})
            if (rem % c == 0) {
                r = c;
                break;
            }
        p.radix[p.nst++] = r;
        rem /= r;
    }
    std::vector<float2> tw(n);
    for (int k = 0; k < n; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)n;
        tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    PG_HIP(hipMalloc(&p.tw, sizeof(float2) * n));
    const hipStream_t ss = setup_stream();
    PG_HIP(hipMemcpyAsync(p.tw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice, ss));
    PG_HIP(hipStreamSynchronize(ss));
    return p;
}

void free_fft_plan(FftPlan1D& p)
{
    if (p.tw) hipFree(p.tw);
    p.tw = nullptr;
}

// fn(begin, end) over [0, n) on up to 4 host threads (set-up loops whose
// iterations are independent), at least `grain` iterations each
template <class F>
static void host_parallel(size_t n, size_t grain, F fn)
{
    const size_t hw = std::max(1u, std::min(4u, std::thread::hardware_concurrency()));
    const size_t nth = std::max<size_t>(1, std::min(hw, n / std::max<size_t>(grain, 1)));
    std::vector<std::thread> th;
    for (size_t t = 1; t < nth; ++t) th.emplace_back(fn, n * t / nth, n * (t + 1) / nth);
    fn(0, n / nth);
    for (std::thread& x : th) x.join();
}

template <class T>
static T* dev_alloc(size_t n)
{
    if (n == 0) n = 1;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * n);
    if (e != hipSuccess)
        throw Error{PG_ERR_OOM, "hipMalloc of " + std::to_string(sizeof(T) * n) + " bytes failed"};
    return (T*)p;
}

template <class T>
static T* dev_copy(const T* h, size_t n)
{
    T* d = dev_alloc<T>(n);
    if (h && n) {
        const hipStream_t ss = setup_stream();
        const hipError_t e1 = hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, ss);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(ss) : e1;
        if (e2 != hipSuccess) {
            hipFree(d);
            PG_HIP(e2);
        }
    }
    return d;
}

template <class T>
static void dev_free(T*& p)
{
    if (p) hipFree((void*)p);
    p = nullptr;
}

// ------------------------------------------------------ Gnufft setup ----
// griddingSupport.cpp:33-48 (Numerical Recipes I0; evaluated in double, as
// the reference's float LUT build promotes to double)
static double bessi0(double x)
{
    double ax, ans, y;
    if ((ax = std::fabs(x)) < 3.75) {
        y = x / 3.75;
        y = y * y;
        ans = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
              y * (0.360768e-1 + y * 0.45813e-2)))));
    } else {
        y = 3.75 / ax;
        ans = (std::exp(ax) / std::sqrt(ax)) * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
              y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
              y * (-0.1647633e-1 + y * 0.392377e-2))))))));
    }
    return ans;
}

// KB transform psi(u) of griddingSupport.cpp:249-260, evaluated in double
static double kb_ft(double W, double beta, double u)
{
    const double e = M_PI * M_PI * W * W * u * u - beta * beta;
    if (e >= 0) return std::sin(std::sqrt(e)) / std::sqrt(e);
    return std::sinh(std::sqrt(-e)) / std::sqrt(-e);
}

// bumped whenever an operator workspace is (re)allocated: a cached solver
// graph that captured workspace addresses is rebuilt after that.  Atomic:
// operators of different threads grow their workspaces concurrently, and a
// lost or reordered update could leave a stale graph in place.
static std::atomic<uint64_t> g_ws_gen{1};

void Gnufft::ensure_batch(int B)
{
    if (B <= cap_batch) return;
    g_ws_gen.fetch_add(1);
    dev_free(grids);
    dev_free(tbuf);
    dev_free(tbuf2);
    const Geometry& g = geo;
    grids = dev_alloc<float2>((size_t)B * g.G());
    // (z-class planes: the y FFT of B*nzc image-sized planes)
    tbuf = dev_alloc<float2>((size_t)B * std::max(g.Nz, nzc) * g.Nx * g.Gy);
    if (g.dims == 3) tbuf2 = dev_alloc<float2>((size_t)B * g.Nz * g.Gx * g.Gy);
    dev_free(vals);
    vals = dev_alloc<float2>((size_t)M * B);
    dev_free(partial);
    if (n_slots > 0) partial = dev_alloc<float2>((size_t)((B + 63) / 64) * n_slots * 16 * 64);
    cap_batch = B;
}

void Gnufft::release()
{
    dev_free(lo);
    dev_free(w);
    dev_free(dfx);
    dev_free(dfy);
    dev_free(dfz);
    dev_free(dap);
    free_fft_plan(px);
    free_fft_plan(py);
    free_fft_plan(pz);
    dev_free(witems);
    dev_free(bgs);
    dev_free(comb);
    dev_free(partial);
    dev_free(entries);
    dev_free(iitems);
    dev_free(iitems4);
    dev_free(irec);
    dev_free(irec4);
    dev_free(ent_w);
    dev_free(tile_off);
    dev_free(tile_idx);
    dev_free(zmat);
    dev_free(dapxy);
    dev_free(zcls);
    nzc = 0;
    dev_free(grids);
    dev_free(tbuf);
    dev_free(tbuf2);
    dev_free(vals);
    dev_free(img_tmp);
    dev_free(dat_tmp);
    cap_batch = 0;
}

static float __int_as_float_host(int v)
{
    float f;
    std::memcpy(&f, &v, sizeof f);
    return f;
}

// Per-dimension window of one sample: unclamped start and the NT weights.
// gridding.cpp:64-97 evaluated in double: s = os (k + N/2); taps
// [max(0, ceil(s - W os/2)), min(os N - 1, floor(s + W os/2))]; weight
// LUT[(int)(d^2 4/W^2 S)], d = |s - n|/os, 0 past the end of the table.
struct Win1 {
    int lo;       // first tap (unclamped)
    int vlo, vhi; // valid tap range (inclusive; empty if vhi < vlo)
};

static Win1 window1(double k, int N, double os, double W, int NT, const std::vector<double>& lut,
                    double* wout)
{
    const double s = os * (k + (double)N / 2.0);
    const double hw = W * os / 2.0;
    Win1 r;
    r.lo = (int)std::ceil(s - hw);
    const int hi = (int)std::min(os * (double)N - 1.0, std::floor(s + hw));
    r.vlo = std::max(0, r.lo);
    r.vhi = std::min(hi, r.lo + NT - 1);
    const int S = (int)lut.size();
    for (int j = 0; j < NT; ++j) {
        const int n = r.lo + j;
        double wt = 0.0;
        if (n >= r.vlo && n <= r.vhi) {
            const double d = std::fabs(s - (double)n) / os;
            const long long k0 = (long long)((d * d * 4.0 / (W * W)) * (double)S);
            wt = (k0 >= S) ? 0.0 : lut[k0];
        }
        wout[j] = wt;
    }
    return r;
}

// Gnufft.cpp:34-122.  unit_taps: every tap weight is scaled by the power of
// two nearest 1/LUT[0] (exact in binary floating point, so every rounding is
// that of the unscaled weights, while sums of products of up to three taps,
// ~LUT[0]^6 ~ 1e44, stay inside the FP32 range); the deapodization is scaled
// back, so the operator is unchanged.
static void gnufft_setup(Gnufft& G, size_t M, float os, size_t Nx, size_t Ny, size_t Nz,
                         const float* kx, const float* ky, const float* kz, float W,
                         bool unit_taps = false)
{
    PG_REQUIRE(M > 0 && Nx > 0 && Ny > 0 && Nz > 0, PG_ERR_ARG, "empty Gnufft");
    PG_REQUIRE(kx && ky, PG_ERR_ARG, "kx/ky required");
    PG_REQUIRE(Nz == 1 || kz, PG_ERR_ARG, "kz required for 3-D");
    PG_REQUIRE(M < (1u << 30), PG_ERR_ARG, "too many samples");
    PG_REQUIRE(Nx == Ny, PG_ERR_UNSUPPORTED,
               "Nx != Ny: the reference's zero_pad2d/circshift2 index x-fastest and are only "
               "consistent for square images");
    Geometry& g = G.geo;
    g.Nx = (int)Nx;
    g.Ny = (int)Ny;
    g.Nz = (int)Nz;
    g.dims = Nz > 1 ? 3 : 2;
    g.os = os;
    // gridding.cpp:440-446
    g.Gx = (int)std::ceil(os * (float)Nx);
    if (g.Gx % 2) g.Gx += 1;
    g.Gy = (int)std::ceil(os * (float)Ny);
    if (g.Gy % 2) g.Gy += 1;
    g.Gz = (Nz == 1) ? 1 : (int)std::ceil(os * (float)Nz);
    PG_REQUIRE(std::fabs(os * (float)Nx - g.Gx) < 1e-6f && std::fabs(os * (float)Ny - g.Gy) < 1e-6f,
               PG_ERR_UNSUPPORTED, "gridos*N must be an even integer");
    // (the 4x4 spread patches tile the x-y grid)
    PG_REQUIRE(g.Gx % 4 == 0 && g.Gy % 4 == 0, PG_ERR_UNSUPPORTED,
               "gridos*N must be a multiple of 4");
    PG_REQUIRE(Nz == 1 || (std::fabs(os * (float)Nz - g.Gz) < 1e-6f && g.Gz % 4 == 0),
               PG_ERR_UNSUPPORTED, "gridos*Nz must be a multiple of 4");
    g.W = W;
    g.ntaps = (int)std::ceil(W * os - 1e-6f);
    PG_REQUIRE(g.ntaps == 8 || g.ntaps == 6 || g.ntaps == 12, PG_ERR_UNSUPPORTED,
               "kernelWidth*gridos must be 6, 8 or 12 taps");
    // Gnufft.cpp:85-88 (beta kept in double for the tables below)
    const double beta_d = M_PI * std::sqrt(((double)os - 0.5) * ((double)os - 0.5) *
                                           ((double)W * W * 4.0) / ((double)os * os) - 0.8);
    g.beta = (float)beta_d;
    // griddingSupport.cpp:469-470, 516-518
    g.offX = (int)(((float)Nx * os / 2.0f) - ((float)Nx / 2.0f));
    g.offY = (int)(((float)Ny * os / 2.0f) - ((float)Ny / 2.0f));
    g.offZ = (Nz > 1) ? (int)(((float)Nz * os / 2.0f) - ((float)Nz / 2.0f)) : 0;
    PG_REQUIRE(g.Ny <= 1024, PG_ERR_UNSUPPORTED, "Ny > 1024");
    G.M = (int)M;
    const int NT = g.ntaps;
    const int D = g.dims;

    // KB look-up table (griddingSupport.cpp:52-76), S = 10000 W, in double
    const int S = (int)(size_t)(10000 * W);
    std::vector<double> lut(S);
    for (int k = 0; k < S; ++k)
        lut[k] = bessi0(beta_d * std::sqrt(1.0 - (double)k / (double)S)) / (double)W;
    const int wexp = unit_taps ? -(int)std::lround(std::log2(lut[0])) : 0;
    for (int k = 0; k < S && wexp != 0; ++k) lut[k] = std::ldexp(lut[k], wexp);

    // per-sample windows and weights
    std::vector<int> lo((size_t)D * M);
    std::vector<float> w((size_t)D * NT * M);
    std::vector<double> wd((size_t)D * NT * M);  // the same in FP64, sample-major [i][d][j]
    std::vector<int> vlo((size_t)D * M), vhi((size_t)D * M);
    const float* kk[3] = {kx, ky, kz};
    const int Nd[3] = {g.Nx, g.Ny, g.Nz};
    host_parallel(M, 8192, [&](size_t b, size_t e) {
        std::vector<double> tmp(NT);
        for (int d = 0; d < D; ++d)
            for (size_t i = b; i < e; ++i) {
                Win1 r = window1((double)kk[d][i], Nd[d], os, W, NT, lut, tmp.data());
                lo[(size_t)d * M + i] = r.lo;
                vlo[(size_t)d * M + i] = r.vlo;
                vhi[(size_t)d * M + i] = r.vhi;
                for (int j = 0; j < NT; ++j) {
                    w[((size_t)d * NT + j) * M + i] = (float)tmp[j];
                    wd[((size_t)i * D + d) * NT + j] = tmp[j];
                }
            }
    });
    G.lo = dev_copy(lo.data(), lo.size());
    G.w = dev_copy(w.data(), w.size());

    // per-axis deapodization factors (griddingSupport.cpp:246-260, 327-348)
    std::vector<float> fx(Nx), fy(Ny), fz(Nz, 1.0f);
    for (size_t X = 0; X < Nx; ++X)
        fx[X] = (float)kb_ft(W, beta_d, ((double)X - (double)(Nx / 2)) / (double)Nx);
    for (size_t Y = 0; Y < Ny; ++Y)
        fy[Y] = (float)kb_ft(W, beta_d, ((double)Y - (double)(Ny / 2)) / (double)Ny);
    if (D == 3)
        for (size_t Z = 0; Z < Nz; ++Z)
            fz[Z] = (float)kb_ft(W, beta_d, ((double)Z - (double)Nz / 2.0) / (double)Nz);
    G.dfx = dev_copy(fx.data(), Nx);
    G.dfy = dev_copy(fy.data(), Ny);
    G.dfz = dev_copy(fz.data(), Nz);
    G.inv_osd = (float)(std::ldexp(1.0, -wexp * D) /
                        ((D == 2) ? (double)os * os : (double)os * os * os));
    {
        // the same factors in FP64, combined per voxel (y fastest)
        std::vector<double> ax(Nx), ay(Ny), az(Nz, 1.0);
        for (size_t X = 0; X < Nx; ++X)
            ax[X] = kb_ft(W, beta_d, ((double)X - (double)(Nx / 2)) / (double)Nx);
        for (size_t Y = 0; Y < Ny; ++Y)
            ay[Y] = kb_ft(W, beta_d, ((double)Y - (double)(Ny / 2)) / (double)Ny);
        if (D == 3)
            for (size_t Z = 0; Z < Nz; ++Z)
                az[Z] = kb_ft(W, beta_d, ((double)Z - (double)Nz / 2.0) / (double)Nz);
        const double osd = (D == 2) ? (double)os * os : (double)os * os * os;
        std::vector<float> dap(Nx * Ny * Nz);
        for (size_t Z = 0; Z < Nz; ++Z)
            for (size_t X = 0; X < Nx; ++X)
                for (size_t Y = 0; Y < Ny; ++Y) {
                    const double k = ax[X] * ay[Y] * az[Z];
                    dap[Y + X * Ny + Z * Nx * Ny] =
                        std::isnan(k) ? 0.0f : (float)(std::ldexp(1.0, -wexp * D) / (osd * k));
                }
        G.dap = dev_copy(dap.data(), dap.size());
    }

    G.fft_ok = fft_length_supported(g.Gx) && fft_length_supported(g.Gy) &&
               (D == 2 || fft_length_supported(g.Gz));
    if (G.fft_ok) {
        G.px = make_fft_plan(g.Gx);
        G.py = make_fft_plan(g.Gy);
        if (D == 3) G.pz = make_fft_plan(g.Gz);
    }

    auto empty = [&](size_t i) {
        for (int d = 0; d < D; ++d)
            if (vhi[(size_t)d * M + i] < vlo[(size_t)d * M + i]) return true;
        return false;
    };

    // 3-D: z classes, the distinct (z window start, FP64 z weights) of the
    // samples with a non-empty window.  A stack of 2-D trajectories (kz on a
    // Cartesian raster) has one class per partition; a plan uses them when
    // there are at most Gz (and at most 128, with classes x Nz <= 4096 for
    // the z matrix in LDS, Nz <= 128) and the y FFT runs in registers (Gy a
    // power of two from 16 to 1024).
    std::vector<int> pl;  // plane of every sample (2-D: none, all on plane 0)
    int npl = 1;
    G.nzc = 0;
    if (D == 3 && pgk::g_zclass.load() && G.fft_ok && (g.Gy & (g.Gy - 1)) == 0 && g.Gy >= 16 &&
        g.Gy <= 1024) {
        std::map<std::vector<uint64_t>, int> key2c;
        std::vector<int> cls(M, -1);
        std::vector<uint64_t> key(NT + 1), last;
        int lastc = -1;
        bool few = true;
        for (size_t i = 0; i < M && few; ++i) {
            if (empty(i)) continue;
            key[0] = (uint64_t)(uint32_t)lo[2 * M + i];
            for (int j = 0; j < NT; ++j) std::memcpy(&key[j + 1], &wd[(i * D + 2) * NT + j], 8);
            if (lastc >= 0 && key == last) {
                cls[i] = lastc;
                continue;
            }
            auto it = key2c.find(key);
            if (it == key2c.end()) {
                if ((int)key2c.size() >= g.Gz) {
                    few = false;
                    break;
                }
                it = key2c.emplace(key, (int)key2c.size()).first;
            }
            cls[i] = lastc = it->second;
            last = key;
        }
        if (few && !key2c.empty() && (long long)key2c.size() * g.Nz <= 4096 &&
            key2c.size() <= 128 && g.Nz <= 128) {
            // classes in order of (window start, weights): std::map's order
            const int nzc = (int)key2c.size();
            std::vector<int> remap(nzc);
            int k = 0;
            for (auto& kv : key2c) remap[kv.second] = k++;
            pl.assign(M, 0);
            for (size_t i = 0; i < M; ++i)
                if (cls[i] >= 0) pl[i] = remap[cls[i]];
            // The z pipeline of the forward transform for image plane z and
            // class k's taps, with the z part of the deapodization, as one
            // complex matrix built in FP64:
            //   Zm[k][z] = dz(z) sum_j wz_k[j] exp(-2 pi i p(lo_k + j) q(z) / Gz)
            // q(z) = (z + offZ + Gz/2) mod Gz places image plane z (zero
            // padding + fftshift), p(n) = (n + Gz/2) mod Gz is the FFT output
            // that grid plane n holds (ifftshift), taps off the grid weigh 0
            // (gridding.cpp:67-73), dz = 1/(os psi_z) (0 where psi_z is NaN,
            // griddingSupport.cpp:327-353).  The adjoint applies its
            // conjugate transpose.
            const int Gz = g.Gz, Nzi = g.Nz;
            std::vector<float2> zm((size_t)nzc * Nzi);
            k = 0;
            for (auto& kv : key2c) {
                const int l0 = (int)(int32_t)(uint32_t)kv.first[0];
                for (int z = 0; z < Nzi; ++z) {
                    const long long q = ((long long)z + g.offZ + Gz / 2) % Gz;
                    double re = 0.0, im = 0.0;
                    for (int j = 0; j < NT; ++j) {
                        const int n = l0 + j;
                        double wj;
                        std::memcpy(&wj, &kv.first[j + 1], 8);
                        if (n < 0 || n >= Gz || wj == 0.0) continue;
                        const long long p = ((long long)n + Gz / 2) % Gz;
                        const double ang = -2.0 * M_PI * (double)((p * q) % Gz) / (double)Gz;
                        re += wj * std::cos(ang);
                        im += wj * std::sin(ang);
                    }
                    const double pz =
                        kb_ft(W, beta_d, ((double)z - (double)Nzi / 2.0) / (double)Nzi);
                    const double dz = std::isnan(pz) ? 0.0 : std::ldexp(1.0, -wexp) / ((double)os * pz);
                    zm[(size_t)k * Nzi + z] = make_float2((float)(re * dz), (float)(im * dz));
                }
                ++k;
            }
            // the x-y part of the deapodization per image column (y fastest)
            std::vector<float> dxy((size_t)g.Nx * g.Ny);
            for (int X = 0; X < g.Nx; ++X) {
                const double ax = kb_ft(W, beta_d, ((double)X - (double)(g.Nx / 2)) / (double)g.Nx);
                for (int Y = 0; Y < g.Ny; ++Y) {
                    const double ay =
                        kb_ft(W, beta_d, ((double)Y - (double)(g.Ny / 2)) / (double)g.Ny);
                    const double kk = ax * ay;
                    dxy[(size_t)X * g.Ny + Y] =
                        std::isnan(kk) ? 0.0f
                                       : (float)(std::ldexp(1.0, -2 * wexp) / ((double)os * os * kk));
                }
            }
            G.nzc = nzc;
            npl = nzc;
            G.zmat = dev_copy(zm.data(), zm.size());
            G.dapxy = dev_copy(dxy.data(), dxy.size());
            G.zcls = dev_copy(pl.data(), pl.size());
        }
    }
    auto plane = [&](size_t i) { return pl.empty() ? 0 : pl[i]; };

    if (D == 2 || G.nzc > 0) {
        // x-y work lists over npl planes (2-D: one)
        if (NT <= 8) {
            // interp tiles: the window start is re-based into [-padl, G-1] (taps
            // moved off the start carry zero weight), binned by (plane, kIT x kIT
            // tile)
            const int padl = NT / 2;
            G.ipadl = padl;
            G.itx = (g.Gx + padl + kIT - 1) / kIT;
            G.ity = (g.Gy + padl + kIT - 1) / kIT;
            const int nt = G.itx * G.ity;
            const long long nbin = (long long)nt * npl;
            std::vector<int> rlo(2 * M);
            std::vector<float> rw((size_t)2 * 8 * M, 0.0f);
            std::vector<int> tcnt(nbin + 1, 0), tof(M);
            const int Gd[2] = {g.Gx, g.Gy};
            for (size_t i = 0; i < M; ++i) {
                int tt[2];
                for (int d = 0; d < 2; ++d) {
                    int l0 = lo[(size_t)d * M + i];
                    int sh = 0;
                    if (empty(i)) {
                        l0 = 0;
                        sh = NT;  // all weights zero
                    } else if (l0 < -padl) {
                        sh = -padl - l0;
                        l0 = -padl;
                    }
                    PG_REQUIRE(l0 <= Gd[d] - 1, PG_ERR_ARG, "window start past the grid");
                    rlo[(size_t)d * M + i] = l0;
                    for (int j = 0; j < 8; ++j) {
                        const int src = j + sh;
                        rw[((size_t)d * M + i) * 8 + j] =
                            (src < NT) ? w[((size_t)d * NT + src) * M + i] : 0.0f;
                    }
                    tt[d] = (l0 + padl) / kIT;
                }
                tof[i] = plane(i) * nt + tt[0] * G.ity + tt[1];
                tcnt[tof[i] + 1]++;
            }
            for (long long t = 0; t < nbin; ++t) tcnt[t + 1] += tcnt[t];
            std::vector<int> pos(tcnt.begin(), tcnt.end() - 1), order(M);
            for (size_t i = 0; i < M; ++i) order[pos[tof[i]]++] = (int)i;
            std::vector<float4> rec((size_t)5 * M), rec4((size_t)5 * M);
            std::vector<int4> its, its4;
            // LDS bank balance.  A wavefront of k_interp_lds reads, per tap,
            // one 8-byte point per (sample, segment) lane at index
            //   ((ox + x) LP + l) R + oy + y   (R = kIR, LP = 8 segment slots)
            // of its item's staged region; the 8-byte bank pair of an index
            // is its value mod 32 (64 banks; mod 16 with 32), so the 8
            // samples of a wavefront read distinct bank pairs at every tap
            // when their keys (LP R ox + oy) mod 32 differ.  One of
            // k_interp_lds4 reads 32 contiguous bytes per sample (its four
            // segment slots) at index ((ox + x) P + oy + y) 4 + l, P the
            // item's row pitch, so that 4 (8) consecutive samples cover the
            // 32 (64) banks once when their keys (P ox + oy) mod 4 (mod 8)
            // differ (gfx950 serves 8-byte reads a half-wavefront at a time
            // over 64 banks: measured conflict-free when the 8 keys of each
            // half differ).  Each item takes the pitch of kIR .. kIR + 2
            // whose dealt keys collide least.  Each item's
            // records are sorted by key and dealt round-robin to groups of
            // 8 samples, which spreads the keys of a group over their values
            // (ascending within a group).  The order of the samples changes
            // nothing else: each is interpolated on its own.
            auto deal = [&](int k0, int n, int group, int xmul, int kmask,
                            const std::vector<int>& idx, const std::vector<int>& oxs,
                            const std::vector<int>& oys, std::vector<int>& outp) {
                std::vector<int> srt(n);
                for (int j = 0; j < n; ++j) srt[j] = j;
                std::stable_sort(srt.begin(), srt.end(), [&](int a, int b) {
                    return ((xmul * oxs[k0 + a] + oys[k0 + a]) & kmask) <
                           ((xmul * oxs[k0 + b] + oys[k0 + b]) & kmask);
                });
                const int ng = (n + group - 1) / group;
                std::vector<int> fill(ng, 0);
                int g = 0;
                for (int j = 0; j < n; ++j) {
                    while (fill[g] >= std::min(group, n - g * group)) g = (g + 1) % ng;
                    outp[k0 + g * group + fill[g]++] = idx[k0 + srt[j]];
                    g = (g + 1) % ng;
                }
            };
            std::vector<int> oxs(M), oys(M), o8(M), o4(M), t4(M);
            auto put = [&](std::vector<float4>& r, int k, int i, int ox, int oy) {
                float4 h;
                h.x = __int_as_float_host(i);
                h.y = __int_as_float_host(ox);
                h.z = __int_as_float_host(oy);
                h.w = 0.0f;
                const float* wx = &rw[(size_t)i * 8];
                const float* wy = &rw[((size_t)M + i) * 8];
                r[(size_t)5 * k + 0] = h;
                r[(size_t)5 * k + 1] = make_float4(wx[0], wx[1], wx[2], wx[3]);
                r[(size_t)5 * k + 2] = make_float4(wx[4], wx[5], wx[6], wx[7]);
                r[(size_t)5 * k + 3] = make_float4(wy[0], wy[1], wy[2], wy[3]);
                r[(size_t)5 * k + 4] = make_float4(wy[4], wy[5], wy[6], wy[7]);
            };
            for (long long tb = 0; tb < nbin; ++tb) {
                const int p = (int)(tb / nt), t = (int)(tb - (long long)p * nt);
                const int tx = t / G.ity, ty = t % G.ity;
                const int rx0 = tx * kIT - padl, ry0 = ty * kIT - padl;
                for (int k = tcnt[tb]; k < tcnt[tb + 1]; ++k) {
                    const int i = order[k];
                    oxs[k] = rlo[i] - rx0;
                    oys[k] = rlo[M + i] - ry0;
                    PG_REQUIRE(oxs[k] >= 0 && oxs[k] < kIT && oys[k] >= 0 && oys[k] < kIT,
                               PG_ERR_ARG, "interp binning");
                }
                for (int k = tcnt[tb]; k < tcnt[tb + 1]; k += kIChunk) {
                    const int n = std::min(kIChunk, tcnt[tb + 1] - k);
                    deal(k, n, 8, 8 * kIR, 31, order, oxs, oys, o8);
                    its.push_back(make_int4(t, k, k + n, p));
                }
                for (int k = tcnt[tb]; k < tcnt[tb + 1]; k += kIChunk4) {
                    const int n = std::min(kIChunk4, tcnt[tb + 1] - k);
                    // the region's row pitch P (points) whose keys (P ox + oy)
                    // mod 8, once dealt, give the fewest bank conflicts: the
                    // sum over groups of 8 of the largest key multiplicity
                    int best = -1, bestc = 0;
                    for (int P = kIR; P < kIR + kI4Pitches; ++P) {
                        deal(k, n, 8, P, 7, order, oxs, oys, t4);
                        int cost = 0;
                        for (int h = 0; h < n; h += 8) {
                            int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0}, mx = 0;
                            for (int j = h; j < std::min(n, h + 8); ++j) {
                                const int i = t4[k + j];
                                const int key = (P * (rlo[i] - rx0) + rlo[M + i] - ry0) & 7;
                                mx = std::max(mx, ++cnt[key]);
                            }
                            cost += mx;
                        }
                        if (best < 0 || cost < bestc) {
                            best = P;
                            bestc = cost;
                            std::copy(t4.begin() + k, t4.begin() + k + n, o4.begin() + k);
                        }
                    }
                    its4.push_back(make_int4(t | (best - kIR) << 24, k, k + n, p));
                }
                for (int k = tcnt[tb]; k < tcnt[tb + 1]; ++k) {
                    const int i8 = o8[k], i4 = o4[k];
                    put(rec, k, i8, rlo[i8] - rx0, rlo[M + i8] - ry0);
                    put(rec4, k, i4, rlo[i4] - rx0, rlo[M + i4] - ry0);
                }
            }
            // items stay in (plane, tile raster) order: neighbouring tiles,
            // whose staged regions overlap, run at about the same time and
            // share L2 lines
            G.n_iitems = (int)its.size();
            G.iitems = dev_copy(its.data(), its.size());
            G.n_iitems4 = (int)its4.size();
            G.iitems4 = dev_copy(its4.data(), its4.size());
            G.irec = dev_copy(rec.data(), rec.size());
            G.irec4 = dev_copy(rec4.data(), rec4.size());
        }  // wider windows interpolate with the gather kernel

        // 4x4 patch work items with per-entry patch weights; patch q of plane
        // p is p * np + q
        const int npx = g.Gx / kPatch, npy = g.Gy / kPatch;
        const long long np = (long long)npx * npy;
        const long long npall = np * npl;
        PG_REQUIRE(npall < (1ll << 31), PG_ERR_UNSUPPORTED, "too many patches");
        G.npp = (int)np;
        std::vector<long long> cnt(npall + 1, 0);
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<long long> pos;
            std::vector<int> ent;
            if (pass == 1) {
                for (long long q = 0; q < npall; ++q) cnt[q + 1] += cnt[q];
                PG_REQUIRE(cnt[npall] < (1ll << 31), PG_ERR_UNSUPPORTED, "too many patch entries");
                pos.assign(cnt.begin(), cnt.end() - 1);
                ent.resize(std::max<long long>(1, cnt[npall]));
            }
            for (size_t i = 0; i < M; ++i) {
                if (empty(i)) continue;
                const long long pb = (long long)plane(i) * np;
                const int x0 = vlo[i] / kPatch, x1 = vhi[i] / kPatch;
                const int y0 = vlo[M + i] / kPatch, y1 = vhi[M + i] / kPatch;
                for (int a = x0; a <= x1; ++a)
                    for (int c = y0; c <= y1; ++c) {
                        const long long q = pb + (long long)a * npy + c;
                        if (pass == 0) cnt[q + 1]++;
                        else ent[pos[q]++] = (int)i;
                    }
            }
            if (pass == 1) {
                // weights of entry k (patch q, sample i) on the patch's 16 points,
                // wx * wy formed in FP64 and rounded once
                std::vector<float> ew((size_t)16 * std::max<long long>(1, cnt[npall]), 0.0f);
                auto fill = [&](long long q0, long long q1) {
                    for (long long q = q0; q < q1; ++q) {
                        const long long qq = q % np;
                        const int px0 = (int)(qq / npy) * kPatch, py0 = (int)(qq % npy) * kPatch;
                        for (long long k = cnt[q]; k < cnt[q + 1]; ++k) {
                            const size_t i = (size_t)ent[k];
                            const double* wx = &wd[(i * D + 0) * NT];
                            const double* wy = &wd[(i * D + 1) * NT];
                            double ax[kPatch], ay[kPatch];
                            for (int u = 0; u < kPatch; ++u) {
                                const int jx = px0 + u - lo[i], jy = py0 + u - lo[M + i];
                                ax[u] = (jx >= 0 && jx < NT) ? wx[jx] : 0.0;
                                ay[u] = (jy >= 0 && jy < NT) ? wy[jy] : 0.0;
                            }
                            for (int u = 0; u < kPatch; ++u)
                                for (int v = 0; v < kPatch; ++v)
                                    ew[16 * k + 4 * u + v] = (float)(ax[u] * ay[v]);
                        }
                    }
                };
                // a few host threads over runs of patches of equal entry count
                const int nth = (int)std::min<long long>(
                    std::max(1u, std::min(8u, std::thread::hardware_concurrency())),
                    std::max<long long>(1, cnt[npall] / 16384));
                std::vector<std::thread> th;
                long long q0 = 0;
                for (int t = 1; t <= nth; ++t) {
                    long long q1 = q0;
                    const long long target = cnt[npall] * t / nth;
                    while (q1 < npall && (cnt[q1] < target || t == nth)) ++q1;
                    if (t < nth) th.emplace_back(fill, q0, q1);
                    else fill(q0, q1);
                    q0 = q1;
                }
                for (std::thread& x : th) x.join();
                G.entries = dev_copy(ent.data(), ent.size());
                G.ent_w = dev_copy(ew.data(), ew.size());
                G.tile_entries = cnt[npall];
            }
        }
        // XCD groups: the patches in (plane, Hilbert) order over the npx x npy
        // patch grid, cut into 8 runs of equal work; group g's blocks get ids
        // congruent to g mod 8 (workgroups are dealt to the 8 XCDs
        // round-robin), so a sample's row of values is read into one L2
        const int kXcd = 8;
        int hn = 1;
        while (hn < std::max(npx, npy)) hn <<= 1;
        std::vector<long long> hkey(np);
        for (long long q = 0; q < np; ++q) {
            int x = (int)(q / npy), y = (int)(q % npy);
            long long d = 0;
            for (int sq = hn / 2; sq > 0; sq /= 2) {
                const int rx = (x & sq) > 0, ry = (y & sq) > 0;
                d += (long long)sq * sq * ((3 * rx) ^ ry);
                if (ry == 0) {
                    if (rx == 1) {
                        x = hn - 1 - x;
                        y = hn - 1 - y;
                    }
                    std::swap(x, y);
                }
            }
            hkey[q] = d;
        }
        std::vector<int> horder(np);
        for (long long q = 0; q < np; ++q) horder[q] = (int)q;
        std::stable_sort(horder.begin(), horder.end(),
                         [&](int u, int v) { return hkey[u] < hkey[v]; });
        auto pweight = [&](long long q) { return (double)(cnt[q + 1] - cnt[q]) + 16.0; };
        double wtot = 0;
        for (long long q = 0; q < npall; ++q) wtot += pweight(q);
        std::vector<std::vector<int>> grp(kXcd);
        {
            double run = 0;
            for (int p = 0; p < npl; ++p)
                for (int q0 : horder) {
                    const int q = (int)((long long)p * np + q0);
                    const double mid = run + 0.5 * pweight(q);
                    grp[std::min(kXcd - 1, (int)(mid * kXcd / wtot))].push_back(q);
                    run += pweight(q);
                }
        }
        // per group: blocks of kW = kSpreadWaves wavefront items {patch, entry
        // begin, entry end, slot}.  A patch of n <= 64 kW entries is one unit
        // taking gs = 1 (n <= 64) or the power of two >= n/64 wavefronts and
        // writes the grid (slot -1); a heavier patch is cut into ceil(n/(64
        // kW)) units of kW wavefronts, each writing a partial slot that a
        // second pass adds up in slot order.
        struct Unit {
            int q, e0, e1, slot, gs;
        };
        const int kW = kSpreadWaves, kUnitMax = kW * 64;
        int nlg = 0;  // gs = 1, 2, ..., kW
        while ((1 << nlg) < kW) ++nlg;
        ++nlg;
        struct SBlock {
            int gs;
            int cost;
            int4 w[kSpreadWaves];
        };
        std::vector<int4> comb;
        int nslots = 0;
        std::vector<std::vector<SBlock>> gblk(kXcd);
        for (int gi = 0; gi < kXcd; ++gi) {
            std::vector<std::vector<Unit>> byg(nlg);
            for (int q : grp[gi]) {
                const int b0 = (int)cnt[q], n = (int)(cnt[q + 1] - cnt[q]);
                if (n > kUnitMax) {
                    const int parts = (n + kUnitMax - 1) / kUnitMax;
                    comb.push_back(make_int4(q, nslots, parts, 0));
                    for (int k = 0; k < parts; ++k)
                        byg[nlg - 1].push_back(Unit{q, b0 + (int)((long long)n * k / parts),
                                                    b0 + (int)((long long)n * (k + 1) / parts),
                                                    nslots + k, kW});
                    nslots += parts;
                    continue;
                }
                int gs = 1, lg = 0;
                while ((long long)gs * 64 < n) {
                    gs *= 2;
                    ++lg;
                }
                byg[lg].push_back(Unit{q, b0, b0 + n, -1, gs});
            }
            for (int lg = 0; lg < nlg; ++lg) {
                const int gs = 1 << lg, per = kW / gs;
                auto& v = byg[lg];
                std::stable_sort(v.begin(), v.end(),
                                 [](const Unit& u, const Unit& w) { return u.e1 - u.e0 > w.e1 - w.e0; });
                for (size_t k0 = 0; k0 < v.size(); k0 += per) {
                    SBlock sb;
                    sb.gs = gs;
                    sb.cost = 0;
                    for (int w = 0; w < kW; ++w) sb.w[w] = make_int4(-1, 0, 0, -1);
                    for (int pk = 0; pk < per && k0 + pk < v.size(); ++pk) {
                        const Unit& un = v[k0 + pk];
                        const long long n = un.e1 - un.e0;
                        for (int w = 0; w < gs; ++w) {
                            const int e0 = (int)(un.e0 + n * w / gs), e1 = (int)(un.e0 + n * (w + 1) / gs);
                            sb.w[pk * gs + w] = make_int4(un.q, e0, e1, un.slot);
                            sb.cost = std::max(sb.cost, e1 - e0);
                        }
                    }
                    gblk[gi].push_back(sb);
                }
            }
            // the longest blocks first
            std::stable_sort(gblk[gi].begin(), gblk[gi].end(),
                             [](const SBlock& u, const SBlock& w) { return u.cost > w.cost; });
        }
        size_t nmax = 0;
        for (int gi = 0; gi < kXcd; ++gi) nmax = std::max(nmax, gblk[gi].size());
        std::vector<int4> witems(nmax * kXcd * kW);
        std::vector<int> bgs(nmax * kXcd, 1);
        for (size_t k = 0; k < nmax; ++k)
            for (int gi = 0; gi < kXcd; ++gi) {
                const size_t blk = k * kXcd + gi;
                for (int w = 0; w < kW; ++w)
                    witems[blk * kW + w] = (k < gblk[gi].size()) ? gblk[gi][k].w[w] : make_int4(-1, 0, 0, -1);
                if (k < gblk[gi].size()) bgs[blk] = gblk[gi][k].gs;
            }
        {
            // groups are aligned runs of gs wavefronts with one patch and slot
            // and consecutive ranges; all ranges of a patch tile its entries
            std::vector<std::vector<std::pair<int, int>>> rng(npall);
            for (size_t blk = 0; blk < nmax * kXcd; ++blk) {
                const int gs = bgs[blk];
                for (int w = 0; w < kW; ++w) {
                    const int4 it = witems[blk * kW + w];
                    const int4 ld = witems[blk * kW + (w / gs) * gs];
                    PG_REQUIRE(it.x == ld.x && it.w == ld.w, PG_ERR_STATE, "spread work groups");
                    if (w % gs) PG_REQUIRE(witems[blk * kW + w - 1].z == it.y, PG_ERR_STATE, "spread work groups");
                    if (it.x < 0) continue;
                    PG_REQUIRE(it.x < npall && it.z >= it.y, PG_ERR_STATE, "spread work list");
                    rng[it.x].push_back({it.y, it.z});
                }
            }
            for (long long q = 0; q < npall; ++q) {
                auto& r = rng[q];
                std::sort(r.begin(), r.end());
                long long at = cnt[q];
                for (auto& pr : r) {
                    PG_REQUIRE(pr.first == at, PG_ERR_STATE, "spread work list coverage");
                    at = pr.second;
                }
                PG_REQUIRE(at == cnt[q + 1] && !r.empty(), PG_ERR_STATE, "spread work list coverage");
            }
        }
        G.n_comb = (int)comb.size();
        G.n_slots = nslots;
        if (!comb.empty()) G.comb = dev_copy(comb.data(), comb.size());
        G.n_sblocks = (int)(nmax * kXcd);
        G.witems = dev_copy(witems.data(), witems.size());
        G.bgs = dev_copy(bgs.data(), bgs.size());
    }
    if (D == 3) {
        // the full 3-D grid (small batches, or no z classes): samples whose
        // window meets each 8x8x4 tile
        const int TX = 8, TY = 8, TZ = 4;
        G.ntx = (g.Gx + TX - 1) / TX;
        G.nty = (g.Gy + TY - 1) / TY;
        G.ntz = (g.Gz + TZ - 1) / TZ;
        const long long ntiles = (long long)G.ntx * G.nty * G.ntz;
        std::vector<long long> cnt(ntiles + 1, 0);
        for (int pass = 0; pass < 2 && G.nzc == 0; ++pass) {
            std::vector<long long> pos;
            std::vector<int> idx;
            if (pass == 1) {
                for (long long t = 0; t < ntiles; ++t) cnt[t + 1] += cnt[t];
                PG_REQUIRE(cnt[ntiles] < (1ll << 31), PG_ERR_UNSUPPORTED, "too many tile entries");
                pos.assign(cnt.begin(), cnt.end() - 1);
                idx.resize(std::max<long long>(1, cnt[ntiles]));
            }
            for (size_t i = 0; i < M; ++i) {
                if (empty(i)) continue;
                for (int tz = vlo[2 * M + i] / TZ; tz <= vhi[2 * M + i] / TZ; ++tz)
                    for (int ty = vlo[M + i] / TY; ty <= vhi[M + i] / TY; ++ty)
                        for (int tx = vlo[i] / TX; tx <= vhi[i] / TX; ++tx) {
                            const long long t = ((long long)tz * G.nty + ty) * G.ntx + tx;
                            if (pass == 0) cnt[t + 1]++;
                            else idx[pos[t]++] = (int)i;
                        }
            }
            if (pass == 1) {
                G.tile_entries = cnt[ntiles];
                std::vector<int> off(cnt.begin(), cnt.end());
                G.tile_off = dev_copy(off.data(), off.size());
                G.tile_idx = dev_copy(idx.data(), idx.size());
            }
        }
    }

    G.img_tmp = dev_alloc<float2>(g.N());
    G.dat_tmp = dev_alloc<float2>(M);
}

// ----------------------------------------------------------- operators --
enum Kind { K_NUFFT = 1, K_DFT = 2, K_TS = 3, K_SENSE = 4, K_PCSENSE = 5, K_PCSENSE_DFT = 6 };

struct SolveCache;
struct Op {
    int magic = 0x50474f50;
    Kind kind;
    size_t n_in = 0, n_out = 0;
    int mode = PG_MODE_BATCHED;
    SolveCache* scache = nullptr;  // PCG work buffers and graph of the last solve
    virtual ~Op();
};

struct NufftOp : Op {
    Gnufft G;
    float2* wtmp = nullptr;  // dcf-weighted coil data of pg_grid_coil_images
    size_t wtmp_cap = 0;
    ~NufftOp() override
    {
        G.release();
        dev_free(wtmp);
    }
};

struct DftOp : Op {
    pgd::DftData d{};
    float2* part = nullptr;
    size_t part_cap = 0;
    float2* tmp = nullptr;  // N (plain SENSE)
    void ensure_part(size_t n)
    {
        if (n <= part_cap) return;
        g_ws_gen.fetch_add(1);
        dev_free(part);
        part = dev_alloc<float2>(n);
        part_cap = n;
    }
    ~DftOp() override
    {
        dev_free(d.kx);
        dev_free(d.ky);
        dev_free(d.kz);
        dev_free(d.t);
        dev_free(d.ix);
        dev_free(d.iy);
        dev_free(d.iz);
        dev_free(d.fm);
        dev_free(d.gx);
        dev_free(d.gy);
        dev_free(d.gz);
        dev_free(part);
        dev_free(tmp);
    }
};

struct TsOp : Op {
    Op* inner = nullptr;
    int L = 1;
    float tau = 0, tmin = 0;
    float* AA = nullptr;    // L x M (Hanning; the real parts of AAc when that is set)
    float2* AAc = nullptr;  // L x M complex (min-max interpolators)
    float* AAi = nullptr;   // L x M, the imaginary parts of AAc, for the fused kernels
    float2* Wo = nullptr;   // L x N
    float2* WoH = nullptr;  // L x N
    float2* tmp_n = nullptr;
    float2* tmp_n2 = nullptr;
    float2* tmp_m = nullptr;
    float2* tmp_m2 = nullptr;
    ~TsOp() override
    {
        dev_free(AA);
        dev_free(AAc);
        dev_free(AAi);
        dev_free(Wo);
        dev_free(WoH);
        dev_free(tmp_n);
        dev_free(tmp_n2);
        dev_free(tmp_m);
        dev_free(tmp_m2);
    }
};

struct SenseOp : Op {
    Op* inner = nullptr;
    int nc = 1;
    size_t M = 0, N = 0;
    float2* S = nullptr;    // nc x N
    float2* P = nullptr;    // nc x N partial images
    float2* tmp_n = nullptr;
    float2* tmp_n2 = nullptr;
    ~SenseOp() override
    {
        dev_free(S);
        dev_free(P);
        dev_free(tmp_n);
        dev_free(tmp_n2);
    }
};

// pcSenseTimeSeg.cpp:40-171: Ns shots, each with its own NUFFT plan and
// time segmentation over its Nd samples, and the tasks (shot, coil) the
// operator covers, shot-major.  pcSenseTimeSeg covers all Ns*Nc tasks and
// its data is Nd x (Ns*Nc), column s + c*Ns; one MPI rank's share of
// mpipcSENSETimeSeg (MPI/mpipcSENSETimeSeg.cpp:101-262) covers every
// nranks-th task and its data is Nd x ntask, column = task index.  The task
// maps S_c exp(-i P_s) are stored in task order ([task][n]), so a shot's
// tasks are one block of maps and of partial images.
struct PcSenseOp : Op {
    int Ns = 1, Nc = 1, L = 1;
    size_t Nd = 0, Ni = 0;
    bool rank_share = false;
    std::vector<int> tshot, tcoil, tcol;  // per task: shot, coil, data column
    std::vector<int> task0, ntask;        // per shot: first task, task count
    std::vector<int> adj_order;           // task order of the adjoint's sum
    std::vector<Op*> G;   // NufftOp per shot with tasks (owned; null otherwise)
    std::vector<Op*> A;   // TsOp per shot with tasks (owned; null otherwise)
    float2* SS = nullptr;  // ntask x Ni
    float2* P = nullptr;   // ntask x Ni partial images
    float2* tmp_n = nullptr;
    int tasks() const { return (int)tshot.size(); }
    // first data column of shot s and the stride between its tasks' columns
    long long col0(int s) const { return rank_share ? task0[s] : s; }
    long long col_stride() const { return rank_share ? 1 : Ns; }
    bool batched() const { return mode == PG_MODE_BATCHED; }
    ~PcSenseOp() override
    {
        for (Op* o : A) delete o;
        for (Op* o : G) delete o;
        dev_free(SS);
        dev_free(P);
        dev_free(tmp_n);
    }
};

// pcSENSE.cpp:39-153: Ns shots, each an exact DFT (Gdft) over its Nd
// samples.  The reference's maps are S_c exp(+i P_s) in the forward and
// conj(S_c) exp(+i P_s) in the adjoint (expiPMap = conj(exp(-i P)),
// pcSENSE.cpp:97-98,117,142).  Stored [s][c][n]: SF = S_c exp(+i P_s), and
// SA = S_c exp(-i P_s), whose conjugate the DFT adjoint kernel applies.
struct PcDftOp : Op {
    int Ns = 1, Nc = 1;
    size_t Nd = 0, Ni = 0;
    std::vector<DftOp*> D;  // per shot (owned)
    DftOp* all = nullptr;   // every shot as one DFT of Ns*Nd samples (owned)
    float2* SF = nullptr;
    float2* SA = nullptr;
    float2* tmp_n = nullptr;
    ~PcDftOp() override
    {
        for (DftOp* d : D) delete d;
        delete all;
        dev_free(SF);
        dev_free(SA);
        dev_free(tmp_n);
    }
};

static Op* as_op(pg_op h)
{
    Op* op = reinterpret_cast<Op*>(h);
    PG_REQUIRE(op && op->magic == 0x50474f50, PG_ERR_ARG, "invalid operator handle");
    return op;
}

// Description of a composite that maps onto the fused batched kernels.
struct Fused {
    Gnufft* G = nullptr;
    int nc = 1, L = 1;
    const float2* S = nullptr;
    const float2* Wo = nullptr;
    const float2* WoH = nullptr;
    const float* AA = nullptr;
    const float* AAi = nullptr;  // imaginary parts of complex weights, or null
};

static bool fusable(Op* op, Fused& f)
{
    if (op->mode != PG_MODE_BATCHED) return false;
    Op* in = op;
    if (in->kind == K_SENSE) {
        SenseOp* s = static_cast<SenseOp*>(in);
        f.nc = s->nc;
        f.S = s->S;
        in = s->inner;
        if (in->mode != PG_MODE_BATCHED) return false;
    }
    if (in->kind == K_TS) {
        TsOp* t = static_cast<TsOp*>(in);
        f.L = t->L;
        f.Wo = t->Wo;
        f.WoH = t->WoH;
        f.AA = t->AA;
        f.AAi = t->AAi;
        in = t->inner;
        if (in->mode != PG_MODE_BATCHED) return false;
    }
    if (in->kind != K_NUFFT) return false;
    f.G = &static_cast<NufftOp*>(in)->G;
    return true;
}

static void op_forward(Op* op, const float2* in, float2* out, hipStream_t s);
static void op_adjoint(Op* op, const float2* in, float2* out, hipStream_t s);

static void require_fft(const Gnufft& G)
{
    PG_REQUIRE(G.fft_ok, PG_ERR_UNSUPPORTED,
               "grid lengths " + std::to_string(G.geo.Gx) + "x" + std::to_string(G.geo.Gy) + "x" +
                   std::to_string(G.geo.Gz) +
                   " have a prime factor above 13: only the spatial interpolation is available");
}

// pg_set_tuning("autotune", 1): plans pick their kernel variants by timing
static std::atomic<int> g_autotune{0};

// Times each variant of plan G's batched interpolation (LDS-staged or
// gather) and spread (vector units or matrix cores) for nc x L grids, on the
// plan's own workspaces (their contents do not matter for timing), and keeps
// the faster of each.  Set-up work on the thread's set-up stream, before any
// capture; a plan is tuned once per grid count.
static void autotune(Gnufft& G, int nc, int L, const float* AA, const float* AAi)
{
    const int B = nc * L;
    const bool planes = pgk::zclass_active(G, B);
    if ((G.geo.dims != 2 && !planes) || !G.fft_ok || G.tuned_for(B)) return;
    G.ensure_batch(B);
    const hipStream_t ss = setup_stream();
    hipEvent_t e[2] = {nullptr, nullptr};
    {
        LegacyLock lk(g_legacy_mu);
        PG_HIP(hipEventCreate(&e[0]));
        if (hipEventCreate(&e[1]) != hipSuccess) {
            hipEventDestroy(e[0]);
            throw Error{PG_ERR_HIP, "hipEventCreate failed"};
        }
    }
    struct Ev {
        hipEvent_t* e;
        ~Ev()
        {
            LegacyLock lk(g_legacy_mu);
            hipEventDestroy(e[0]);
            hipEventDestroy(e[1]);
        }
    } ev{e};
    // timing inputs: zeroed workspaces (a small fill kernel, not a memset node)
    pgv::fill(G.grids, (long long)B * G.geo.G(), make_float2(0.f, 0.f), ss);
    pgv::fill(G.vals, (long long)B * G.M, make_float2(0.f, 0.f), ss);
    auto timed = [&](auto&& launch) {
        launch();  // warm-up
        PG_HIP(hipEventRecord(e[0], ss));
        for (int r = 0; r < 3; ++r) launch();
        PG_HIP(hipEventRecord(e[1], ss));
        PG_HIP(hipEventSynchronize(e[1]));
        float ms = 0;
        PG_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
        return ms;
    };
    G.tuned.push_back(Gnufft::Tuned{B, 1, 0});
    Gnufft::Tuned& t = G.tuned.back();
    if (pgk::interp_lds_possible(G, L, AA != nullptr)) {
        // gather (0), all segments staged over bins of <= 128 samples (1),
        // four segments per stage over bins of <= 64 samples (2)
        auto interp = [&] {
            pgk::launch_interp(G, G.grids, nc, L, AA, AAi, G.vals, G.M, ss, planes);
        };
        float best = 0.f;
        int pick = 1;
        for (int v = 0; v <= 2; ++v) {
            if (v == 2 && G.n_iitems4 == 0) continue;
            t.lds = v;
            const float ms = timed(interp);
            if (v == 0 || ms < best) {
                best = ms;
                pick = v;
            }
        }
        t.lds = pick;
    } else {
        t.lds = 0;
    }
    if (B % 4 == 0) {
        auto spread = [&] {
            pgk::launch_spread(G, nullptr, G.M, nc, L, AA, AAi, nullptr, nullptr, nullptr, G.grids,
                               true, ss, planes);
        };
        t.mfma = 0;
        const float ms_vec = timed(spread);
        t.mfma = 1;
        const float ms_mfma = timed(spread);
        t.mfma = ms_mfma < ms_vec ? 1 : 0;
    }
    PG_HIP(hipGetLastError());
    // the fills (and any timing launch) are done before the caller's stream
    // uses the workspaces again, whatever was tunable
    PG_HIP(hipStreamSynchronize(ss));
    // solve graphs captured before this plan was tuned ran the defaults
    g_ws_gen.fetch_add(1);
}

// autotune the plan of a fused launch on stream s, unless already tuned for
// its grid count or s is capturing; s is drained first, since the timing
// runs overwrite the plan's workspaces
static void maybe_autotune(const Fused& f, hipStream_t s)
{
    if (!g_autotune.load() || f.G->tuned_for(f.nc * f.L)) return;
    if (s) {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        PG_HIP(hipStreamIsCapturing(s, &st));
        if (st != hipStreamCaptureStatusNone) return;
        PG_HIP(hipStreamSynchronize(s));
    } else {
        LegacyLock lk(g_legacy_mu);
        PG_HIP(hipDeviceSynchronize());
    }
    autotune(*f.G, f.nc, f.L, f.AA, f.AAi);
}

static void fused_forward(const Fused& f, const float2* x, float2* y, hipStream_t s,
                          long long y_cstride = -1)
{
    Gnufft& G = *f.G;
    require_fft(G);
    const int B = f.nc * f.L;
    G.ensure_batch(B);
    maybe_autotune(f, s);
    pgk::WeightArgs w{f.S, f.Wo, f.WoH, f.AA};
    const bool planes = pgk::zclass_active(G, B);
    {
        Prof p("fwd_prep_fft", s);
        pgk::launch_fwd_prep_fft(G, x, 0, w, f.nc, f.L, G.tbuf, G.tbuf2, G.grids, s, planes);
    }
    {
        Prof p("interp", s);
        pgk::launch_interp(G, G.grids, f.nc, f.L, f.AA, f.AAi, y, y_cstride < 0 ? G.M : y_cstride,
                           s, planes);
    }
}

// adjoint into per-coil partial images P (nc x N), optionally of the CG
// residual W (y - Ax) formed on the fly
static void fused_adjoint_partial(const Fused& f, const float2* in, const float2* resid_y,
                                  const double2* ax64, const float* Wd, float2* P, hipStream_t s,
                                  bool vals_ready = false, long long in_cstride = -1)
{
    Gnufft& G = *f.G;
    require_fft(G);
    const int B = f.nc * f.L;
    G.ensure_batch(B);
    maybe_autotune(f, s);
    pgk::WeightArgs w{f.S, f.Wo, f.WoH, f.AA};
    const long long ics = in_cstride < 0 ? G.M : in_cstride;
    const bool planes = pgk::zclass_active(G, B);
    if (!vals_ready && pgk::spread_uses_values(G, B)) {
        Prof p("spread_values", s);
        pgk::launch_spread_values(G, in, ics, f.nc, f.L, f.AA, f.AAi, resid_y, ax64, Wd, s);
        vals_ready = true;
    }
    {
        Prof p("spread", s);
        pgk::launch_spread(G, in, ics, f.nc, f.L, f.AA, f.AAi, resid_y, ax64, Wd, G.grids,
                           vals_ready, s, planes);
    }
    {
        Prof p("adj_fft_post", s);
        pgk::launch_adj_fft_post(G, G.grids, G.tbuf, G.tbuf2, w, f.nc, f.L, P, G.geo.N(), s,
                                 planes);
    }
}

// allocate every workspace a forward/adjoint pair can touch (so that a
// captured graph never allocates)
static void op_prepare(Op* op)
{
    Fused f;
    if (fusable(op, f)) {
        f.G->ensure_batch(f.nc * f.L);
        return;
    }
    switch (op->kind) {
    case K_NUFFT:
        static_cast<NufftOp*>(op)->G.ensure_batch(1);
        return;
    case K_DFT: {
        DftOp* d = static_cast<DftOp*>(op);
        d->ensure_part(std::max(pgd::part_elems(d->d.M, d->d.N, 1, false),
                                pgd::part_elems(d->d.M, d->d.N, 1, true)));
        return;
    }
    case K_TS:
        op_prepare(static_cast<TsOp*>(op)->inner);
        return;
    case K_PCSENSE: {
        PcSenseOp* pc = static_cast<PcSenseOp*>(op);
        for (int sh = 0; sh < pc->Ns; ++sh) {
            Op* ts = pc->A[sh];
            if (!ts) continue;
            if (pc->mode == PG_MODE_BATCHED)
                static_cast<NufftOp*>(static_cast<TsOp*>(ts)->inner)
                    ->G.ensure_batch(pc->ntask[sh] * pc->L);
            op_prepare(ts);
        }
        return;
    }
    case K_PCSENSE_DFT: {
        PcDftOp* pc = static_cast<PcDftOp*>(op);
        if (pc->mode == PG_MODE_BATCHED) {
            DftOp* d = pc->all;
            d->ensure_part(std::max(pgd::part_elems(d->d.M, d->d.N, pc->Nc, false, pc->Ns),
                                    pgd::part_elems(d->d.M, d->d.N, pc->Nc, true, pc->Ns)));
            return;
        }
        for (DftOp* d : pc->D)
            d->ensure_part(std::max(pgd::part_elems(d->d.M, d->d.N, 1, false),
                                    pgd::part_elems(d->d.M, d->d.N, 1, true)));
        return;
    }
    case K_SENSE: {
        SenseOp* se = static_cast<SenseOp*>(op);
        if (se->inner->kind == K_DFT) {
            DftOp* d = static_cast<DftOp*>(se->inner);
            d->ensure_part(std::max(pgd::part_elems(d->d.M, d->d.N, se->nc, false),
                                    pgd::part_elems(d->d.M, d->d.N, se->nc, true)));
        }
        op_prepare(se->inner);
        return;
    }
    }
}

// the reference's class name of an operator, for its tracing ranges
static const char* op_range_name(const Op* op, bool fwd)
{
    switch (op->kind) {
    case K_NUFFT: return fwd ? "Gnufft::operator*" : "Gnufft::operator/";
    case K_DFT: return fwd ? "Gdft::operator*" : "Gdft::operator/";
    case K_TS: return fwd ? "TimeSegmentation::operator*" : "TimeSegmentation::operator/";
    case K_SENSE: return fwd ? "SENSE::operator*" : "SENSE::operator/";
    case K_PCSENSE: return fwd ? "pcSenseTimeSeg::operator*" : "pcSenseTimeSeg::operator/";
    case K_PCSENSE_DFT: return fwd ? "pcSENSE::operator*" : "pcSENSE::operator/";
    }
    return fwd ? "operator*" : "operator/";
}

static void op_forward(Op* op, const float2* in, float2* out, hipStream_t s)
{
    TraceRange tr(op_range_name(op, true));
    Fused f;
    if (fusable(op, f)) {
        fused_forward(f, in, out, s);
        return;
    }
    switch (op->kind) {
    case K_NUFFT: {
        Fused g;
        g.G = &static_cast<NufftOp*>(op)->G;
        fused_forward(g, in, out, s);
        return;
    }
    case K_DFT: {
        DftOp* d = static_cast<DftOp*>(op);
        d->ensure_part(pgd::part_elems(d->d.M, d->d.N, 1, false));
        Prof p("gdft_fwd", s);
        pgd::forward(d->d, in, nullptr, 1, out, d->part, s);
        return;
    }
    case K_TS: {
        // TimeSegmentation.cpp:252-296, one segment at a time
        TsOp* t = static_cast<TsOp*>(op);
        const long long N = (long long)t->n_in, M = (long long)t->n_out;
        for (int l = 0; l < t->L; ++l) {
            pgv::cmul(t->tmp_n, t->Wo + (size_t)l * N, in, N, s);
            op_forward(t->inner, t->tmp_n, t->tmp_m, s);
            if (t->AAc)
                pgv::cmul_acc(out, t->tmp_m, t->AAc + (size_t)l * M, M, l == 0, false, s);
            else
                pgv::rmul_acc(out, t->AA + (size_t)l * M, t->tmp_m, M, l == 0, s);
        }
        return;
    }
    case K_SENSE: {
        SenseOp* se = static_cast<SenseOp*>(op);
        const long long N = (long long)se->N, M = (long long)se->M;
        if (se->inner->kind == K_DFT && op->mode == PG_MODE_BATCHED) {
            DftOp* d = static_cast<DftOp*>(se->inner);
            d->ensure_part(pgd::part_elems(d->d.M, d->d.N, se->nc, false));
            Prof p("gdft_fwd", s);
            pgd::forward(d->d, in, se->S, se->nc, out, d->part, s);
            return;
        }
        // SENSE.cpp:59-67, one coil at a time
        for (int c = 0; c < se->nc; ++c) {
            pgv::cmul(se->tmp_n, se->S + (size_t)c * N, in, N, s);
            op_forward(se->inner, se->tmp_n, out + (size_t)c * M, s);
        }
        return;
    }
    case K_PCSENSE: {
        // pcSenseTimeSeg.cpp:129-148, MPI/mpipcSENSETimeSeg.cpp:163-171
        PcSenseOp* pc = static_cast<PcSenseOp*>(op);
        const long long Ni = (long long)pc->Ni, Nd = (long long)pc->Nd;
        if (pc->batched()) {
            // every task x segment of a shot in one fused launch; the shot's
            // task columns are col_stride*Nd apart
            for (int sh = 0; sh < pc->Ns; ++sh) {
                if (!pc->ntask[sh]) continue;
                TsOp* ts = static_cast<TsOp*>(pc->A[sh]);
                Fused f;
                f.G = &static_cast<NufftOp*>(ts->inner)->G;
                f.nc = pc->ntask[sh];
                f.L = ts->L;
                f.S = pc->SS + (size_t)pc->task0[sh] * Ni;
                f.Wo = ts->Wo;
                f.WoH = ts->WoH;
                f.AA = ts->AA;
                f.AAi = ts->AAi;
                fused_forward(f, in, out + pc->col0(sh) * Nd, s, pc->col_stride() * Nd);
            }
            return;
        }
        for (int k = 0; k < pc->tasks(); ++k) {
            pgv::cmul(pc->tmp_n, pc->SS + (size_t)k * Ni, in, Ni, s);
            op_forward(pc->A[pc->tshot[k]], pc->tmp_n, out + (size_t)pc->tcol[k] * Nd, s);
        }
        return;
    }
    case K_PCSENSE_DFT: {
        // pcSENSE.cpp:108-127
        PcDftOp* pc = static_cast<PcDftOp*>(op);
        const long long Ni = (long long)pc->Ni, Nd = (long long)pc->Nd;
        Prof p("gdft_fwd", s);
        if (op->mode == PG_MODE_BATCHED) {
            // every shot and coil in one launch: column s + c*Ns of Nd samples
            // is sample range [s*Nd, (s+1)*Nd) of coil c's Ns*Nd
            DftOp* d = pc->all;
            d->ensure_part(pgd::part_elems(d->d.M, d->d.N, pc->Nc, false, pc->Ns));
            pgd::forward(d->d, in, pc->SF, pc->Nc, out, d->part, s, -1, pc->Ns, pc->Nc * Ni);
            return;
        }
        for (int sh = 0; sh < pc->Ns; ++sh) {
            DftOp* d = pc->D[sh];
            const float2* SFs = pc->SF + (size_t)sh * pc->Nc * Ni;
            d->ensure_part(pgd::part_elems(d->d.M, d->d.N, 1, false));
            for (int c = 0; c < pc->Nc; ++c) {
                pgv::cmul(pc->tmp_n, SFs + (size_t)c * Ni, in, Ni, s);
                pgd::forward(d->d, pc->tmp_n, nullptr, 1,
                             out + ((size_t)sh + (size_t)c * pc->Ns) * Nd, d->part, s);
            }
        }
        return;
    }
    }
}

static void op_adjoint(Op* op, const float2* in, float2* out, hipStream_t s)
{
    TraceRange tr(op_range_name(op, false));
    Fused f;
    if (fusable(op, f)) {
        if (f.nc == 1) {
            fused_adjoint_partial(f, in, nullptr, nullptr, nullptr, out, s);
        } else {
            SenseOp* se = static_cast<SenseOp*>(op);
            fused_adjoint_partial(f, in, nullptr, nullptr, nullptr, se->P, s);
            pgv::coil_sum(out, se->P, se->nc, (long long)se->N, s);
        }
        return;
    }
    switch (op->kind) {
    case K_NUFFT: {
        Fused g;
        g.G = &static_cast<NufftOp*>(op)->G;
        fused_adjoint_partial(g, in, nullptr, nullptr, nullptr, out, s);
        return;
    }
    case K_DFT: {
        DftOp* d = static_cast<DftOp*>(op);
        d->ensure_part(pgd::part_elems(d->d.M, d->d.N, 1, true));
        Prof p("gdft_adj", s);
        pgd::adjoint(d->d, in, nullptr, 1, out, d->part, s);
        return;
    }
    case K_TS: {
        // TimeSegmentation.cpp:297-324
        TsOp* t = static_cast<TsOp*>(op);
        const long long N = (long long)t->n_in, M = (long long)t->n_out;
        for (int l = 0; l < t->L; ++l) {
            if (t->AAc)  // conj(AA) % d
                pgv::cmul_acc(t->tmp_m, in, t->AAc + (size_t)l * M, M, true, true, s);
            else
                pgv::rmul_acc(t->tmp_m, t->AA + (size_t)l * M, in, M, true, s);
            op_adjoint(t->inner, t->tmp_m, t->tmp_n, s);
            pgv::cmul_acc(out, t->tmp_n, t->WoH + (size_t)l * N, N, l == 0, false, s);
        }
        return;
    }
    case K_SENSE: {
        SenseOp* se = static_cast<SenseOp*>(op);
        const long long N = (long long)se->N, M = (long long)se->M;
        if (se->inner->kind == K_DFT && op->mode == PG_MODE_BATCHED) {
            DftOp* d = static_cast<DftOp*>(se->inner);
            d->ensure_part(pgd::part_elems(d->d.M, d->d.N, se->nc, true));
            Prof p("gdft_adj", s);
            pgd::adjoint(d->d, in, se->S, se->nc, out, d->part, s);
            return;
        }
        // SENSE.cpp:84-96
        for (int c = 0; c < se->nc; ++c) {
            op_adjoint(se->inner, in + (size_t)c * M, se->tmp_n, s);
            pgv::cmul_acc(out, se->tmp_n, se->S + (size_t)c * N, N, c == 0, true, s);
        }
        return;
    }
    case K_PCSENSE: {
        // pcSenseTimeSeg.cpp:152-171, MPI/mpipcSENSETimeSeg.cpp:218-228
        PcSenseOp* pc = static_cast<PcSenseOp*>(op);
        const long long Ni = (long long)pc->Ni, Nd = (long long)pc->Nd;
        if (pc->batched()) {
            // per shot, every task x segment in one fused launch, into that
            // shot's block of partial images; then one sum over the tasks
            for (int sh = 0; sh < pc->Ns; ++sh) {
                if (!pc->ntask[sh]) continue;
                TsOp* ts = static_cast<TsOp*>(pc->A[sh]);
                Fused f;
                f.G = &static_cast<NufftOp*>(ts->inner)->G;
                f.nc = pc->ntask[sh];
                f.L = ts->L;
                f.S = pc->SS + (size_t)pc->task0[sh] * Ni;
                f.Wo = ts->Wo;
                f.WoH = ts->WoH;
                f.AA = ts->AA;
                f.AAi = ts->AAi;
                fused_adjoint_partial(f, in + pc->col0(sh) * Nd, nullptr, nullptr, nullptr,
                                      pc->P + (size_t)pc->task0[sh] * Ni, s, false,
                                      pc->col_stride() * Nd);
            }
            pgv::coil_sum(out, pc->P, pc->tasks(), Ni, s);
            return;
        }
        bool first = true;
        for (int k : pc->adj_order) {
            op_adjoint(pc->A[pc->tshot[k]], in + (size_t)pc->tcol[k] * Nd, pc->tmp_n, s);
            pgv::cmul_acc(out, pc->tmp_n, pc->SS + (size_t)k * Ni, Ni, first, true, s);
            first = false;
        }
        return;
    }
    case K_PCSENSE_DFT: {
        // pcSENSE.cpp:131-153
        PcDftOp* pc = static_cast<PcDftOp*>(op);
        const long long Ni = (long long)pc->Ni, Nd = (long long)pc->Nd;
        Prof p("gdft_adj", s);
        if (op->mode == PG_MODE_BATCHED) {
            // every shot and coil in one launch; the shots' sums are added in
            // shot order
            DftOp* d = pc->all;
            d->ensure_part(pgd::part_elems(d->d.M, d->d.N, pc->Nc, true, pc->Ns));
            pgd::adjoint(d->d, in, pc->SA, pc->Nc, out, d->part, s, -1, false, pc->Ns,
                         pc->Nc * Ni);
            return;
        }
        // the reference's order: coils outer, shots inner
        for (int c = 0; c < pc->Nc; ++c)
            for (int sh = 0; sh < pc->Ns; ++sh) {
                DftOp* d = pc->D[sh];
                d->ensure_part(pgd::part_elems(d->d.M, d->d.N, 1, true));
                pgd::adjoint(d->d, in + ((size_t)sh + (size_t)c * pc->Ns) * Nd, nullptr, 1,
                             pc->tmp_n, d->part, s);
                pgv::cmul_acc(out, pc->tmp_n, pc->SA + ((size_t)sh * pc->Nc + c) * Ni, Ni,
                              c == 0 && sh == 0, true, s);
            }
        return;
    }
    }
}

// ----------------------------------------------------------- penalty ----
struct Penalty {
    int magic = 0x5047504e;
    pgv::PenSpec spec;
    double* part = nullptr;
    double* out = nullptr;
    ~Penalty()
    {
        dev_free(part);
        dev_free(out);
    }
};

static Penalty* as_pen(pg_penalty h)
{
    Penalty* p = reinterpret_cast<Penalty*>(h);
    PG_REQUIRE(p && p->magic == 0x5047504e, PG_ERR_ARG, "invalid penalty handle");
    return p;
}

// --------------------------------------------------------------- PCG ----
struct CgWork {
    double2 *x = nullptr, *Ax = nullptr;  // FP64 CG state (solution, A x)
    float2* x32 = nullptr;                // FP32 copy of x (penalty stencils, output)
    float2 *ngrad = nullptr, *ddir = nullptr, *Adir = nullptr;
    float2 *g = nullptr, *r = nullptr, *tmp = nullptr;
    double* part = nullptr;
    pgv::CgParts cp{};
    double* resid = nullptr;
    CgScal* sc = nullptr;
    ~CgWork()
    {
        dev_free(x);
        dev_free(Ax);
        dev_free(x32);
        dev_free(ngrad);
        dev_free(ddir);
        dev_free(Adir);
        dev_free(g);
        dev_free(r);
        dev_free(tmp);
        dev_free(part);
        dev_free(resid);
        dev_free(sc);
    }
};

// one PCG iteration (solve_pwls_pcg.hpp:80-164) as a launch sequence
// Work buffers and the instantiated graph of a PCG solve, kept on the
// operator and reused while the solve's key is unchanged (the graph bakes in
// the addresses of y, W and the operator workspaces, and the penalty's
// parameters).
struct SolveCache {
    const Penalty* R = nullptr;
    pgv::PenSpec spec{};
    const float2* y = nullptr;
    const float* W = nullptr;
    size_t niter = 0;
    double tol = 0.0;
    uint64_t ws_gen = 0;
    uint64_t uid = 0;  // unique per instantiated graph (0: none)
    CgWork w;
    hipStream_t own = nullptr;  // stream for graph solves without a user stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~SolveCache()
    {
        LegacyLock lk(g_legacy_mu);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (own) hipStreamDestroy(own);
    }
    void drop_graph()
    {
        LegacyLock lk(g_legacy_mu);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

Op::~Op() { delete scache; }

static void cg_iteration(Op* A, Penalty* R, const float2* y, const float* W, CgWork& w,
                         hipStream_t s, const Fused* f)
{
    TraceRange tr("solve_pwls_pcg iteration");
    const long long N = (long long)A->n_in, M = (long long)A->n_out;
    const pgv::PenSpec* ps = R ? &R->spec : nullptr;
    // ngrad = A / (W (y - Ax)) - R.Gradient(x)
    if (f) {
        SenseOp* se = (A->kind == K_SENSE) ? static_cast<SenseOp*>(A) : nullptr;
        float2* P = se ? se->P : w.g;
        // the spread values come from the previous update (or the set-up)
        fused_adjoint_partial(*f, nullptr, y, w.Ax, W, P, s, true);
        pgv::cg_grad(P, f->nc, N, ps, w.x32, w.ngrad, w.cp, w.sc, w.resid, s);
    } else {
        pgv::resid(w.r, y, w.Ax, W, M, s);
        op_adjoint(A, w.r, w.g, s);
        pgv::cg_grad(w.g, 1, N, ps, w.x32, w.ngrad, w.cp, w.sc, w.resid, s);
    }
    pgv::cg_dir(w.ngrad, w.ddir, N, w.cp, w.sc, s);
    op_forward(A, w.ddir, w.Adir, s);
    pgv::cg_adir(w.Adir, y, w.Ax, W, M, ps, w.ddir, w.x32, w.cp, w.sc, s);
    pgv::cg_linesearch(ps, w.ddir, w.x32, w.cp, w.sc, s);
    // fused: the update also writes W (y - Ax) in FP32, from which the next
    // adjoint's per-sample spread values are formed (coalesced along grids)
    pgv::cg_update(w.Ax, w.Adir, y, W, M, w.x, w.x32, w.ddir, N, ps, w.cp, f ? w.r : nullptr,
                   w.sc, s);
    if (f)
        pgk::launch_spread_values(*f->G, w.r, f->G->M, f->nc, f->L, f->AA, f->AAi, nullptr,
                                  nullptr, nullptr, s);
}

static bool same_spec(const pgv::PenSpec& a, const pgv::PenSpec& b)
{
    return a.kind == b.kind && a.nd == b.nd && a.n == b.n && a.off[0] == b.off[0] &&
           a.off[1] == b.off[1] && a.off[2] == b.off[2] && a.beta == b.beta && a.delta == b.delta;
}

// x = x0 (or 0), Ax = A x (:55), the CG scalars, and the first adjoint's
// spread values
static void cg_setup(Op* A, const float2* y, const float* W, CgWork& w, const Fused* f,
                     double tol, hipStream_t s)
{
    const long long N = (long long)A->n_in, M = (long long)A->n_out;
    pgv::f2d(w.x, w.x32, N, s);
    op_forward(A, w.x32, w.tmp, s);
    pgv::f2d(w.Ax, w.tmp, M, s);
    pgv::cg_init(y, W, M, w.part, w.sc, tol, s);
    if (f)  // spread values of the first adjoint, from the initial Ax
        pgk::launch_spread_values(*f->G, nullptr, f->G->M, f->nc, f->L, f->AA, f->AAi, y, w.Ax, W,
                                  s);
}

// the solve cache of operator A for this key (R, W, y, niter, tol), made
// anew when the key changed; its graph is dropped when a workspace moved
static SolveCache* solve_cache(Op* A, Penalty* R, const float* W, const float2* y, size_t niter,
                               double tol)
{
    const long long N = (long long)A->n_in, M = (long long)A->n_out;
    op_prepare(A);  // may grow workspaces (and bump g_ws_gen)
    pgv::PenSpec spec{};
    if (R) spec = R->spec;
    SolveCache* c = A->scache;
    const bool same = c && c->R == R && (!R || same_spec(c->spec, spec)) && c->y == y && c->W == W &&
                      c->niter == niter && c->tol == tol;
    if (!same) {
        delete c;
        A->scache = nullptr;
        c = new SolveCache();
        A->scache = c;
        c->R = R;
        c->spec = spec;
        c->y = y;
        c->W = W;
        c->niter = niter;
        c->tol = tol;
        CgWork& w = c->w;
        w.x = dev_alloc<double2>(N);
        w.Ax = dev_alloc<double2>(M);
        w.x32 = dev_alloc<float2>(N);
        w.tmp = dev_alloc<float2>(M);
        w.ngrad = dev_alloc<float2>(N);
        w.ddir = dev_alloc<float2>(N);
        w.Adir = dev_alloc<float2>(M);
        w.g = dev_alloc<float2>(N);
        w.r = dev_alloc<float2>(M);
        w.part = dev_alloc<double>(std::max<size_t>(pgv::cg_parts_doubles(N, M), 2 * pgv::kMaxParts));
        w.cp = pgv::cg_parts(w.part, N, M);
        w.resid = dev_alloc<double>(std::max<size_t>(niter, 1));
        w.sc = dev_alloc<CgScal>(1);
    }
    const uint64_t gen = g_ws_gen.load();
    if (c->ws_gen != gen) c->drop_graph();
    c->ws_gen = gen;
    return c;
}

// capture the whole solve (set-up, every iteration, bookkeeping) on s into
// c->graph and instantiate it, unless that is done; every kernel checks the
// device-side stop flag, so an early exit leaves the remaining iterations as
// no-ops
static void solve_capture(SolveCache* c, Op* A, Penalty* R, const float* W, const float2* y,
                          const Fused* f, hipStream_t s)
{
    if (c->exec) return;
    CgWork& w = c->w;
    LegacyLock lk(g_legacy_mu);
    PG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        cg_setup(A, y, W, w, f, c->tol, s);
        for (size_t it = 0; it < c->niter; ++it) cg_iteration(A, R, y, W, w, s, f);
        pgv::cg_finish(w.cp, w.sc, w.resid, s);
    } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(s, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    PG_HIP(hipStreamEndCapture(s, &c->graph));
    PG_HIP(hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
    static std::atomic<uint64_t> next_uid{1};
    c->uid = next_uid.fetch_add(1);
}

// the result of a finished solve: x_out, the residual history and info
static void solve_results(SolveCache* c, long long N, float2* x_out, double* resid_h,
                          pg_cg_info* info, hipStream_t s)
{
    CgWork& w = c->w;
    CgScal hs{};
    PG_HIP(hipMemcpyAsync(&hs, w.sc, sizeof(CgScal), hipMemcpyDeviceToHost, s));
    PG_HIP(hipMemcpyAsync(x_out, w.x32, sizeof(float2) * N, hipMemcpyDeviceToDevice, s));
    std::vector<double> rh(std::max<size_t>(c->niter, 1), 0.0);
    PG_HIP(hipMemcpyAsync(rh.data(), w.resid, sizeof(double) * rh.size(), hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    if (resid_h)
        for (int i = 0; i < hs.iter && (size_t)i < c->niter; ++i) resid_h[i] = rh[i];
    if (info) {
        info->iterations = hs.iter;
        info->exit_reason = hs.stop;
        info->final_resid = hs.iter > 0 ? rh[hs.iter - 1] : 0.0;
    }
}

static void solve_pcg(Op* A, Penalty* R, const float2* x0, const float* W, const float2* y,
                      size_t niter, double tol, int mode, float2* x_out, double* resid_h,
                      pg_cg_info* info, hipStream_t user_stream)
{
    const long long N = (long long)A->n_in;
    Fused fz;
    const bool fused = fusable(A, fz);
    if (!(tol > 0.0)) tol = 0.0;
    SolveCache* c = solve_cache(A, R, W, y, niter, tol);
    CgWork& w = c->w;

    hipStream_t s = user_stream;
    if (mode == PG_CG_GRAPH && s == nullptr) {
        LegacyLock lk(g_legacy_mu);
        PG_HIP(hipStreamSynchronize(nullptr));
        if (!c->own) PG_HIP(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
        s = c->own;
    }

    if (fused) maybe_autotune(fz, s);  // before any capture
    if (x0) PG_HIP(hipMemcpyAsync(w.x32, x0, sizeof(float2) * N, hipMemcpyDeviceToDevice, s));
    else pgv::fill(w.x32, N, make_float2(0.f, 0.f), s);

    if (mode == PG_CG_GRAPH) {
        solve_capture(c, A, R, W, y, fused ? &fz : nullptr, s);
        PG_HIP(hipGraphLaunch(c->exec, s));
    } else {
        cg_setup(A, y, W, w, fused ? &fz : nullptr, tol, s);
        for (size_t it = 0; it < niter; ++it) {
            cg_iteration(A, R, y, W, w, s, fused ? &fz : nullptr);
            // host-side branch on the device scalars, as the reference loop does
            CgScal hs{};
            PG_HIP(hipMemcpyAsync(&hs, w.sc, sizeof(CgScal), hipMemcpyDeviceToHost, s));
            PG_HIP(hipStreamSynchronize(s));
            if (hs.stop) break;
        }
        pgv::cg_finish(w.cp, w.sc, w.resid, s);
    }
    solve_results(c, N, x_out, resid_h, info, s);
}

// every operator object under A (the composite and its inner operators)
static void collect_ops(Op* A, std::vector<const Op*>& out)
{
    if (!A) return;
    out.push_back(A);
    switch (A->kind) {
    case K_TS:
        collect_ops(static_cast<TsOp*>(A)->inner, out);
        break;
    case K_SENSE:
        collect_ops(static_cast<SenseOp*>(A)->inner, out);
        break;
    case K_PCSENSE: {
        PcSenseOp* pc = static_cast<PcSenseOp*>(A);
        for (Op* o : pc->A) collect_ops(o, out);
        break;
    }
    case K_PCSENSE_DFT: {
        PcDftOp* pc = static_cast<PcDftOp*>(A);
        for (DftOp* d : pc->D) out.push_back(d);
        out.push_back(pc->all);
        break;
    }
    default:
        break;
    }
}

// pg_set_tuning("batch_streams", n): the solves of a batch run on n of the
// library's streams at once (0: as one parent graph)
static std::atomic<int> g_batch_streams{4};

// the library's streams for batches on the current device (default
// priority, non-blocking; created on first use and kept)
static hipStream_t batch_stream(int i)
{
    static std::map<int, std::vector<hipStream_t>>* per_dev =
        new std::map<int, std::vector<hipStream_t>>();  // kept for the process
    int dev = 0;
    PG_HIP(hipGetDevice(&dev));
    LegacyLock lk(g_legacy_mu);
    std::vector<hipStream_t>& v = (*per_dev)[dev];
    while ((int)v.size() <= i) {
        hipStream_t st = nullptr;
        PG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        v.push_back(st);
    }
    return v[i];
}

// the captured solve graphs of a batch launched on nst streams at once,
// the most expensive first, each onto the stream with the least work so far
// (work ~ grids x (grid points + samples) of the fused launches); the
// streams wait for everything enqueued on s before, and s waits for them
static void solve_batch_streams(size_t K, Op* const* A, SolveCache* const* cs, int nst,
                                hipStream_t s, std::vector<hipEvent_t>& ev)
{
    const int S = (int)std::min<size_t>((size_t)nst, K);
    std::vector<double> cost(K);
    for (size_t k = 0; k < K; ++k) {
        Fused fz;
        cost[k] = fusable(A[k], fz)
                      ? (double)fz.nc * fz.L * ((double)fz.G->geo.G() + (double)fz.G->M)
                      : (double)A[k]->n_in + (double)A[k]->n_out;
    }
    std::vector<size_t> order(K);
    for (size_t k = 0; k < K; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    ev.assign(S + 1, nullptr);
    {
        LegacyLock lk(g_legacy_mu);
        for (int i = 0; i <= S; ++i) PG_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
    std::vector<hipStream_t> st(S);
    for (int i = 0; i < S; ++i) st[i] = batch_stream(i);
    PG_HIP(hipEventRecord(ev[S], s));
    for (int i = 0; i < S; ++i) PG_HIP(hipStreamWaitEvent(st[i], ev[S], 0));
    std::vector<double> load(S, 0.0);
    for (size_t k : order) {
        const int i = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        PG_HIP(hipGraphLaunch(cs[k]->exec, st[i]));
        load[i] += cost[k];
    }
    for (int i = 0; i < S; ++i) {
        PG_HIP(hipEventRecord(ev[i], st[i]));
        PG_HIP(hipStreamWaitEvent(s, ev[i], 0));
    }
}

// K independent reconstructions (reconSolve each: W = ones, x0 = 0) as one
// graph launch: each solve's captured graph is a child of one parent graph,
// with no edges between them, so the device runs them concurrently
static void solve_batch(size_t K, Op* const* A, Penalty* const* R, const float2* const* y,
                        size_t niter, double tol, float2* const* x_out, pg_cg_info* infos,
                        hipStream_t s)
{
    if (!(tol > 0.0)) tol = 0.0;
    std::vector<SolveCache*> cs(K);
    for (size_t k = 0; k < K; ++k) {
        Fused fz;
        const bool fused = fusable(A[k], fz);
        cs[k] = solve_cache(A[k], R[k], nullptr, y[k], niter, tol);
        if (fused) maybe_autotune(fz, s);
        solve_capture(cs[k], A[k], R[k], nullptr, y[k], fused ? &fz : nullptr, s);
    }
    for (size_t k = 0; k < K; ++k)
        pgv::fill(cs[k]->w.x32, (long long)A[k]->n_in, make_float2(0.f, 0.f), s);
    const int nst = g_batch_streams.load();
    if (nst > 0) {
        // the fork/join events are destroyed once s has drained (solve_results
        // synchronizes it), also when a launch fails
        std::vector<hipEvent_t> ev;
        struct Evs {
            std::vector<hipEvent_t>& e;
            hipStream_t s;
            ~Evs()
            {
                LegacyLock lk(g_legacy_mu);
                hipStreamSynchronize(s);
                for (hipEvent_t x : e)
                    if (x) hipEventDestroy(x);
            }
        } guard{ev, s};
        solve_batch_streams(K, A, cs.data(), nst, s, ev);
        for (size_t k = 0; k < K; ++k)
            solve_results(cs[k], (long long)A[k]->n_in, x_out[k], nullptr,
                          infos ? infos + k : nullptr, s);
        return;
    }
    // the parent graph of these K solve graphs, from a small cache keyed by
    // the children's unique ids (a recaptured or destroyed solve never
    // matches a stale entry)
    struct Parent {
        std::vector<uint64_t> key;
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
    };
    static std::mutex mu;
    static std::vector<Parent>* cache = new std::vector<Parent>();  // kept for the process
    std::vector<uint64_t> key(K);
    for (size_t k = 0; k < K; ++k) key[k] = cs[k]->uid;
    hipGraphExec_t exec = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const Parent& p : *cache)
            if (p.key == key) exec = p.e;
        if (!exec) {
            LegacyLock ll(g_legacy_mu);
            Parent p;
            p.key = key;
            PG_HIP(hipGraphCreate(&p.g, 0));
            try {
                for (size_t k = 0; k < K; ++k) {
                    hipGraphNode_t node;
                    PG_HIP(hipGraphAddChildGraphNode(&node, p.g, nullptr, 0, cs[k]->graph));
                }
                PG_HIP(hipGraphInstantiate(&p.e, p.g, nullptr, nullptr, 0));
            } catch (...) {
                hipGraphDestroy(p.g);
                throw;
            }
            if (cache->size() >= 4) {
                hipGraphExecDestroy(cache->front().e);
                hipGraphDestroy(cache->front().g);
                cache->erase(cache->begin());
            }
            cache->push_back(p);
            exec = p.e;
        }
        // launched and read back under the cache lock, so that no other
        // thread evicts this graph while it runs
        PG_HIP(hipGraphLaunch(exec, s));
        for (size_t k = 0; k < K; ++k)
            solve_results(cs[k], (long long)A[k]->n_in, x_out[k], nullptr,
                          infos ? infos + k : nullptr, s);
    }
}

// TimeSegmentation.cpp:36-221 over the operator G (not owned)
using cd = std::complex<double>;

// TimeSegmentation.cpp:94-150 (interptype 2), in FP64: the exact
// least-squares min-max interpolator.  g_j = exp(i fm_j tau), G = [1, g,
// ..., g^(L-1)] (N x L), GTG = G^H G built as the reference does (N on the
// diagonal, sum_j conj(g_j^k) on the k-th sub-diagonal, conjugates above);
// W = inv(GTG) G^H, or pinv(GTG) G^H when 1/cond(GTG) <= 2e-15 (pinv
// cutoff: L * max eigenvalue * double eps); AA[l][i] = conj(W exp(i fm t_i))
// on the device.  Returns the L x M complex weights.
static float2* minmax_ls_weights(const float* fieldmap, const float* tvec, size_t M, size_t N,
                                 int L, size_t shots, double tau, double tmin)
{
    std::vector<double> fm(N);
    for (size_t j = 0; j < N; ++j) fm[j] = fieldmap[j];
    std::vector<cd> glsum(L, cd(0, 0));
    for (size_t j = 0; j < N; ++j)
        for (int k = 0; k < L - 1; ++k) glsum[k] += std::exp(cd(0, -(double)(k + 1) * fm[j] * tau));
    std::vector<cd> GTG((size_t)L * L, cd(0, 0));
    for (int r = 0; r < L; ++r) GTG[(size_t)r * L + r] = (double)N;
    for (int k = 0; k < L - 1; ++k)
        for (int r = k + 1; r < L; ++r) {
            GTG[(size_t)r * L + (r - k - 1)] += glsum[k];
            GTG[(size_t)(r - k - 1) * L + r] += std::conj(glsum[k]);
        }
    std::vector<double> ev;
    std::vector<cd> V;
    herm_eig(L, GTG, ev, V);
    double smax = 0, smin = 1e300;
    for (double e : ev) {
        smax = std::max(smax, std::fabs(e));
        smin = std::min(smin, std::fabs(e));
    }
    const double rcn = smax > 0 ? smin / smax : 0.0;
    const double cut = rcn > 10 * 2e-16 ? 0.0 : (double)L * smax * 2.220446049250313e-16;
    std::vector<cd> inv((size_t)L * L, cd(0, 0));
    for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
            cd acc(0, 0);
            for (int k = 0; k < L; ++k)
                if (std::fabs(ev[k]) > cut && ev[k] != 0.0)
                    acc += V[(size_t)a * L + k] * std::conj(V[(size_t)b * L + k]) / ev[k];
            inv[(size_t)a * L + b] = acc;
        }
    // W[l][j] = sum_c inv[l][c] conj(g_j^c)
    std::vector<double2> W((size_t)L * N);
    std::vector<cd> gc(L);
    for (size_t j = 0; j < N; ++j) {
        for (int c = 0; c < L; ++c) gc[c] = std::exp(cd(0, -(double)c * fm[j] * tau));
        for (int l = 0; l < L; ++l) {
            cd acc(0, 0);
            for (int c = 0; c < L; ++c) acc += inv[(size_t)l * L + c] * gc[c];
            W[(size_t)l * N + j] = make_double2(acc.real(), acc.imag());
        }
    }
    const size_t none = M / shots;
    std::vector<double> tv(none);
    for (size_t i = 0; i < none; ++i) tv[i] = (double)tvec[i] - tmin;
    struct Tmp {
        double* fm = nullptr;
        double* tv = nullptr;
        double2* W = nullptr;
        ~Tmp()
        {
            dev_free(fm);
            dev_free(tv);
            dev_free(W);
        }
    } d;
    d.fm = dev_copy(fm.data(), N);
    d.tv = dev_copy(tv.data(), none);
    d.W = dev_copy(W.data(), W.size());
    float2* AA = dev_alloc<float2>((size_t)L * M);
    pgt::minmax_weights(d.fm, d.tv, d.W, (int)N, (int)none, L, (int)shots, AA,
                        setup_stream());
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(setup_stream());
    if (e1 != hipSuccess || e2 != hipSuccess) {
        dev_free(AA);
        PG_HIP(e1 != hipSuccess ? e1 : e2);
    }
    return AA;
}

// TimeSegmentation.cpp:157-207 (interptype 3), in FP64: the approximate
// min-max interpolator, which replaces the voxels by the field map's
// 256-bin histogram (Armadillo hist: centres min + (n + 1/2) range/256, each
// value at its nearest centre, the lower on a tie).  KK = floor(2 pi /
// ((range/256) tau)), dwn = 2 pi / (KK tau); bin n stands for the frequency
// om_n = min + (n + 1/2) dwn, and bins n >= KK drop out (the KK-point FFT
// input is trimmed).  The (L+1) x (L+1) matrix inverted is GTGap.st(),
// entry (r, c) = sum_n h_n exp(-i om_n tau (r - c)); inv when its 1-norm
// reciprocal condition exceeds 10 double eps, else pinv (cut-off (L+1) *
// max eigenvalue * double eps).  The reference assigns the L+1
// coefficients of each sample to an L-column row, which keeps the first L
// (TimeSegmentation.cpp:204 under ARMA_NO_DEBUG).  Returns the L x M
// complex weights.
static float2* histogram_ls_weights(const float* fieldmap, const float* tvec, size_t M, size_t N,
                                    int L, size_t shots, double tau, double tmin)
{
    constexpr int nbins = pgt::kHistBins;
    PG_REQUIRE(L + 1 <= pgt::kHistCoef, PG_ERR_UNSUPPORTED,
               "interptype 3 supports at most 16 time segments");
    double fmin = 0, fmax = 0;
    for (size_t j = 0; j < N; ++j) {
        const double v = fieldmap[j];
        PG_REQUIRE(std::isfinite(v), PG_ERR_ARG, "interptype 3 needs a finite field map");
        if (j == 0 || v < fmin) fmin = v;
        if (j == 0 || v > fmax) fmax = v;
    }
    const double range = fmax - fmin;
    PG_REQUIRE(range > 0, PG_ERR_UNSUPPORTED, "interptype 3 needs a field map with a nonzero range");
    std::vector<double> h(nbins, 0.0);
    auto centre = [&](int n) { return fmin + range * ((0.5 + n) / nbins); };
    for (size_t j = 0; j < N; ++j) {
        const double v = fieldmap[j];
        int k0 = (int)std::floor((v - fmin) / range * nbins);
        k0 = std::min(std::max(k0, 0), nbins - 1);
        int best = std::max(k0 - 1, 0);
        double bd = std::fabs(centre(best) - v);
        for (int n = best + 1; n <= std::min(k0 + 1, nbins - 1); ++n) {
            const double d = std::fabs(centre(n) - v);
            if (d < bd) {
                bd = d;
                best = n;
            }
        }
        h[best] += 1.0;
    }
    const double kkd = std::floor(2.0 * M_PI / ((range / nbins) * tau));
    PG_REQUIRE(kkd >= 2.0 * L + 1 && kkd < 1e8, PG_ERR_UNSUPPORTED,
               "interptype 3: the histogram FFT length is outside [2L+1, 1e8)");
    const int KK = (int)kkd;
    const double dwn = 2.0 * M_PI / (KK * tau);
    const int nb = std::min(nbins, KK);
    std::vector<double> om(nb);
    std::vector<double2> z(nb);
    for (int n = 0; n < nb; ++n) {
        om[n] = fmin + (n + 0.5) * dwn;
        z[n] = make_double2(std::cos(om[n] * tau), -std::sin(om[n] * tau));
    }
    const int nc = L + 1;
    // f[d] = sum_n h_n exp(-i om_n tau d), d = 0..L; entry (r, c) is f[r - c],
    // or conj(f[c - r]) above the diagonal
    std::vector<cd> f(nc, cd(0, 0));
    for (int d = 0; d < nc; ++d)
        for (int n = 0; n < nb; ++n) f[d] += h[n] * std::exp(cd(0, -om[n] * tau * d));
    std::vector<cd> A((size_t)nc * nc);
    for (int r = 0; r < nc; ++r)
        for (int c = 0; c < nc; ++c) A[(size_t)r * nc + c] = r >= c ? f[r - c] : std::conj(f[c - r]);
    std::vector<double> ev;
    std::vector<cd> V;
    herm_eig(nc, A, ev, V);
    double smax = 0, smin = 1e300;
    for (double e : ev) {
        smax = std::max(smax, std::fabs(e));
        smin = std::min(smin, std::fabs(e));
    }
    auto inverse = [&](double cut) {
        std::vector<cd> inv((size_t)nc * nc, cd(0, 0));
        for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) {
                cd acc(0, 0);
                for (int k = 0; k < nc; ++k)
                    if (std::fabs(ev[k]) > cut && ev[k] != 0.0)
                        acc += V[(size_t)a * nc + k] * std::conj(V[(size_t)b * nc + k]) / ev[k];
                inv[(size_t)a * nc + b] = acc;
            }
        return inv;
    };
    auto norm1 = [&](const std::vector<cd>& X) {
        double m = 0;
        for (int c = 0; c < nc; ++c) {
            double sacc = 0;
            for (int r = 0; r < nc; ++r) sacc += std::abs(X[(size_t)r * nc + c]);
            m = std::max(m, sacc);
        }
        return m;
    };
    const double eps = 2.220446049250313e-16;
    std::vector<cd> inv = inverse(0.0);
    const double rcond = (smin > 0 && smax > 0) ? 1.0 / (norm1(A) * norm1(inv)) : 0.0;
    if (!(rcond > 10 * eps)) inv = inverse((double)nc * smax * eps);
    std::vector<double2> iG((size_t)nc * nc);
    for (size_t e = 0; e < iG.size(); ++e) iG[e] = make_double2(inv[e].real(), inv[e].imag());
    const size_t none = M / shots;
    std::vector<double> tv(none);
    for (size_t i = 0; i < none; ++i) tv[i] = (double)tvec[i] - tmin;
    struct Tmp {
        double* h = nullptr;
        double* om = nullptr;
        double2* z = nullptr;
        double* tv = nullptr;
        double2* iG = nullptr;
        ~Tmp()
        {
            dev_free(h);
            dev_free(om);
            dev_free(z);
            dev_free(tv);
            dev_free(iG);
        }
    } d;
    d.h = dev_copy(h.data(), nb);
    d.om = dev_copy(om.data(), nb);
    d.z = dev_copy(z.data(), nb);
    d.tv = dev_copy(tv.data(), none);
    d.iG = dev_copy(iG.data(), iG.size());
    float2* AA = dev_alloc<float2>((size_t)L * M);
    pgt::histogram_weights(d.h, d.om, d.z, nb, d.tv, (int)none, d.iG, L, (int)shots, AA,
                        setup_stream());
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(setup_stream());
    if (e1 != hipSuccess || e2 != hipSuccess) {
        dev_free(AA);
        PG_HIP(e1 != hipSuccess ? e1 : e2);
    }
    return AA;
}

static TsOp* ts_build(Op* G, const float* fieldmap, const float* tvec, size_t M, size_t N, size_t L,
                      size_t interptype, size_t shots)
{
    PG_REQUIRE(G->n_in == N && G->n_out == M, PG_ERR_ARG, "TimeSegmentation sizes do not match G");
    PG_REQUIRE(interptype >= 1 && interptype <= 3, PG_ERR_UNSUPPORTED,
               "interptype 1 (Hanning), 2 (exact least-squares min-max) and 3 (histogram "
               "min-max) are implemented");
    PG_REQUIRE(N < (1u << 31) && M < (1u << 31), PG_ERR_ARG, "sizes too large");
    if (shots < 1) shots = 1;
    PG_REQUIRE(M % shots == 0, PG_ERR_ARG, "M must be a multiple of shots");
    std::unique_ptr<TsOp> op(new TsOp());
    op->kind = K_TS;
    op->inner = G;
    op->n_in = N;
    op->n_out = M;
    // TimeSegmentation.cpp:46-93, 213-218, evaluated in double and stored
    // as FP32 tables
    const int Lc = (L == 0 || L == 1) ? 1 : (int)L;
    op->L = Lc;
    double tmin = tvec[0], tmax = tvec[0];
    for (size_t j = 1; j < M; ++j) {
        tmin = std::min(tmin, (double)tvec[j]);
        tmax = std::max(tmax, (double)tvec[j]);
    }
    const double rangt = tmax - tmin;
    double tau = (rangt + 2.220446049250313e-16) / (double)Lc;
    std::vector<float> AA((size_t)Lc * M, 0.0f);
    std::vector<float2> Wo((size_t)Lc * N), WoH((size_t)Lc * N);
    if (Lc <= 1) {
        tau = 0;
        std::fill(AA.begin(), AA.end(), 1.0f);
        std::fill(Wo.begin(), Wo.end(), make_float2(1.f, 0.f));
        std::fill(WoH.begin(), WoH.end(), make_float2(1.f, 0.f));
    } else {
        const size_t none = M / shots;
        for (int l = 0; l < Lc; ++l)
            for (size_t j = 0; j < none; ++j) {
                const double d = ((double)tvec[j] - tmin) - (double)l * tau;
                float a = 0.0f;
                if (std::fabs(d) <= tau) a = (float)(0.5 + 0.5 * std::cos(M_PI * d / tau));
                for (size_t sh = 0; sh < shots; ++sh) AA[(size_t)l * M + j + sh * none] = a;
            }
        host_parallel((size_t)Lc * N, 16384, [&](size_t b, size_t e) {
            for (size_t k = b; k < e; ++k) {
                const size_t l = k / N, n = k - l * N;
                const double ph = (double)fieldmap[n] * ((double)l * tau + tmin);
                const float c = (float)std::cos(ph), sn = (float)std::sin(ph);
                Wo[k] = make_float2(c, -sn);
                WoH[k] = make_float2(c, sn);
            }
        });
    }
    op->tau = (float)tau;
    op->tmin = (float)tmin;
    if (interptype == 2 && Lc > 1)
        op->AAc = minmax_ls_weights(fieldmap, tvec, M, N, Lc, shots, tau, tmin);
    if (interptype == 3 && Lc > 1)
        op->AAc = histogram_ls_weights(fieldmap, tvec, M, N, Lc, shots, tau, tmin);
    std::vector<float> AAim;
    if (op->AAc) {
        // the fused kernels take complex weights as real and imaginary tables
        std::vector<float2> c(AA.size());
        const hipStream_t ss = setup_stream();
        PG_HIP(hipMemcpyAsync(c.data(), op->AAc, c.size() * sizeof(float2), hipMemcpyDeviceToHost,
                              ss));
        PG_HIP(hipStreamSynchronize(ss));
        AAim.resize(AA.size());
        for (size_t e = 0; e < c.size(); ++e) {
            AA[e] = c[e].x;
            AAim[e] = c[e].y;
        }
        op->AAi = dev_copy(AAim.data(), AAim.size());
    }
    op->AA = dev_copy(AA.data(), AA.size());
    op->Wo = dev_copy(Wo.data(), Wo.size());
    op->WoH = dev_copy(WoH.data(), WoH.size());
    op->tmp_n = dev_alloc<float2>(N);
    op->tmp_n2 = dev_alloc<float2>(N);
    op->tmp_m = dev_alloc<float2>(M);
    op->tmp_m2 = dev_alloc<float2>(M);
    return op.release();
}

}  // namespace pg

// ===================================================================== ABI
using namespace pg;

#define PG_TRY try {
#define PG_CATCH                                                     \
    }                                                                \
    catch (const pg::Error& e)                                       \
    {                                                                \
        pg::set_last_error(e.msg);                                   \
        return e.code;                                               \
    }                                                                \
    catch (const std::bad_alloc&)                                    \
    {                                                                \
        pg::set_last_error("host allocation failed");                \
        return PG_ERR_OOM;                                           \
    }                                                                \
    catch (...)                                                      \
    {                                                                \
        pg::set_last_error("unknown error");                         \
        return PG_ERR_STATE;                                         \
    }                                                                \
    return PG_OK;

extern "C" {

const char* pg_status_string(int st)
{
    switch (st) {
    case PG_OK: return "ok";
    case PG_ERR_ARG: return "invalid argument";
    case PG_ERR_UNSUPPORTED: return "unsupported";
    case PG_ERR_HIP: return "HIP runtime error";
    case PG_ERR_OOM: return "out of device memory";
    case PG_ERR_STATE: return "invalid state";
    case PG_ERR_NUMERIC: return "numerical error";
    }
    return "unknown status";
}

const char* pg_last_error(void) { return pg::g_last_error.c_str(); }
const char* pg_version(void) { return "0.1.0"; }

int pg_set_device(int device)
{
    PG_TRY PG_HIP(hipSetDevice(device));
    PG_CATCH
}
int pg_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    PG_TRY PG_REQUIRE(free_bytes && total_bytes, PG_ERR_ARG, "null argument");
    PG_HIP(hipMemGetInfo(free_bytes, total_bytes));
    PG_CATCH
}

int pg_malloc(void** ptr, size_t bytes)
{
    PG_TRY PG_REQUIRE(ptr, PG_ERR_ARG, "null ptr");
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) throw Error{PG_ERR_OOM, "hipMalloc failed"};
    PG_CATCH
}
int pg_free(void* ptr)
{
    PG_TRY if (ptr) PG_HIP(hipFree(ptr));
    PG_CATCH
}
// With stream NULL the copies are ordered after all work on the device: the
// library's own streams are non-blocking, so the null stream alone would not
// wait for them.
// A null stream means the legacy stream: the device is drained first, and
// the copy and its synchronisation hold the legacy lock throughout, so that
// no other thread's stream capture is open meanwhile.
static void copy_sync(void* d, const void* src, size_t n, hipMemcpyKind k, void* s, bool sync)
{
    std::unique_lock<std::recursive_mutex> lk(g_legacy_mu, std::defer_lock);
    if (!s) {
        lk.lock();
        PG_HIP(hipDeviceSynchronize());
    }
    PG_HIP(hipMemcpyAsync(d, src, n, k, (hipStream_t)s));
    if (sync) PG_HIP(hipStreamSynchronize((hipStream_t)s));
}
int pg_memcpy_htod(void* d, const void* h, size_t n, void* s)
{
    PG_TRY copy_sync(d, h, n, hipMemcpyHostToDevice, s, true);
    PG_CATCH
}
int pg_memcpy_dtoh(void* h, const void* d, size_t n, void* s)
{
    PG_TRY copy_sync(h, d, n, hipMemcpyDeviceToHost, s, true);
    PG_CATCH
}
int pg_memcpy_dtod(void* d, const void* src, size_t n, void* s)
{
    PG_TRY copy_sync(d, src, n, hipMemcpyDeviceToDevice, s, s == nullptr);
    PG_CATCH
}
int pg_stream_create(void** s)
{
    PG_TRY PG_REQUIRE(s, PG_ERR_ARG, "null ptr");
    hipStream_t st;
    LegacyLock lk(g_legacy_mu);
    PG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = (void*)st;
    PG_CATCH
}
int pg_stream_destroy(void* s)
{
    PG_TRY LegacyLock lk(g_legacy_mu);
    if (s) PG_HIP(hipStreamDestroy((hipStream_t)s));
    PG_CATCH
}
int pg_stream_synchronize(void* s)
{
    PG_TRY std::unique_lock<std::recursive_mutex> lk(g_legacy_mu, std::defer_lock);
    if (!s) lk.lock();
    PG_HIP(hipStreamSynchronize((hipStream_t)s));
    PG_CATCH
}

int pg_gnufft_create(pg_op* out, size_t M, float gridos, size_t Nx, size_t Ny, size_t Nz,
                     const float* kx, const float* ky, const float* kz, const float* ix,
                     const float* iy, const float* iz, float kernelWidth)
{
    (void)ix;
    (void)iy;
    (void)iz;
    PG_TRY TraceRange tr_("Gnufft::Gnufft");
    PG_REQUIRE(out, PG_ERR_ARG, "null out");
    *out = nullptr;
    std::unique_ptr<NufftOp> op(new NufftOp());
    op->kind = K_NUFFT;
    try {
        gnufft_setup(op->G, M, gridos, Nx, Ny, Nz, kx, ky, kz, kernelWidth);
    } catch (...) {
        op->G.release();
        throw;
    }
    op->n_in = Nx * Ny * Nz;
    op->n_out = M;
    *out = reinterpret_cast<pg_op>(op.release());
    PG_CATCH
}

int pg_gdft_create(pg_op* out, size_t M, size_t N, const float* kx, const float* ky,
                   const float* kz, const float* ix, const float* iy, const float* iz,
                   const float* fm, const float* t, int precision)
{
    PG_TRY TraceRange tr_("Gdft::Gdft");
    PG_REQUIRE(out && kx && ky && ix && iy, PG_ERR_ARG, "null argument");
    PG_REQUIRE(M > 0 && N > 0 && M < (1u << 31) && N < (1u << 31), PG_ERR_ARG, "bad sizes");
    PG_REQUIRE(precision == 0 || precision == 1, PG_ERR_ARG, "precision must be 0 or 1");
    *out = nullptr;
    std::unique_ptr<DftOp> op(new DftOp());
    op->kind = K_DFT;
    op->n_in = N;
    op->n_out = M;
    std::vector<float> zm(M, 0.0f), zn(N, 0.0f);
    op->d.M = (int)M;
    op->d.N = (int)N;
    op->d.precision = precision;
    op->d.kx = dev_copy(kx, M);
    op->d.ky = dev_copy(ky, M);
    op->d.kz = dev_copy(kz ? kz : zm.data(), M);
    op->d.t = dev_copy(t ? t : zm.data(), M);
    op->d.ix = dev_copy(ix, N);
    op->d.iy = dev_copy(iy, N);
    op->d.iz = dev_copy(iz ? iz : zn.data(), N);
    op->d.fm = dev_copy(fm ? fm : zn.data(), N);
    op->tmp = dev_alloc<float2>(N);
    *out = reinterpret_cast<pg_op>(op.release());
    PG_CATCH
}

int pg_gdftr2_create(pg_op* out, size_t M, size_t N, const float* kx, const float* ky,
                     const float* kz, const float* ix, const float* iy, const float* iz,
                     const float* fm, const float* t, size_t Nx, size_t Ny, size_t Nz)
{
    PG_TRY PG_REQUIRE(out && fm && Nx > 0 && Ny > 0 && Nz > 0, PG_ERR_ARG, "null argument");
    PG_REQUIRE(Nx * Ny * Nz == N, PG_ERR_ARG, "Nx*Ny*Nz must equal N");
    *out = nullptr;
    pg_op h = nullptr;
    const int st = pg_gdft_create(&h, M, N, kx, ky, kz, ix, iy, iz, fm, t, 0);
    if (st != PG_OK) throw pg::Error{st, pg::g_last_error};
    std::unique_ptr<DftOp> op(reinterpret_cast<DftOp*>(h));
    // GdftR2.cpp:60-107: Cd(FM, dim) in float, offsets 1, Ny, Nx*Ny of the
    // y-fastest image.  In 2-D the z offset is the whole map and the
    // reference indexes outside it; that map is zero here.
    std::vector<float> g(N);
    float** dst[3] = {&op->d.gx, &op->d.gy, &op->d.gz};
    const size_t offs[3] = {1, Ny, Nx * Ny};
    for (int dim = 0; dim < 3; ++dim) {
        const size_t o = offs[dim];
        if (o >= N) {
            std::fill(g.begin(), g.end(), 0.0f);
        } else {
            for (size_t i = o; i < N - o; ++i) g[i] = fm[i + o] - fm[i - o];
            for (size_t i = 0; i < o; ++i) g[i] = fm[i + o] - fm[i];
            for (size_t i = N - o; i < N; ++i) g[i] = fm[i] - fm[i - o];
        }
        *dst[dim] = dev_copy(g.data(), N);
    }
    op->d.numx = (int)Nx;
    op->d.numy = (int)Ny;
    op->d.numz = (int)Nz;
    *out = reinterpret_cast<pg_op>(op.release());
    PG_CATCH
}

int pg_ts_create(pg_op* out, pg_op Gh, const float* fieldmap, const float* tvec, size_t M,
                 size_t N, size_t L, size_t interptype, size_t shots)
{
    PG_TRY TraceRange tr_("TimeSegmentation::TimeSegmentation");
    PG_REQUIRE(out && fieldmap && tvec, PG_ERR_ARG, "null argument");
    *out = nullptr;
    Op* G = as_op(Gh);
    *out = reinterpret_cast<pg_op>(ts_build(G, fieldmap, tvec, M, N, L, interptype, shots));
    PG_CATCH
}


// pcSenseTimeSeg (all tasks, rank_share false) or one MPI rank's share of
// mpipcSENSETimeSeg (rank_share true: tasks rank, rank + nranks, ...)
static PcSenseOp* pcsense_build(const float* kx, const float* ky, const float* kz, size_t Nx,
                                size_t Ny, size_t Nz, size_t nc, const float* tvec, size_t L,
                                size_t interptype, const float* sens, const float* fieldmap,
                                const float* shotphase, size_t n_shots, size_t Nd,
                                float kernelWidth, bool rank_share, size_t rank, size_t nranks)
{
    PG_REQUIRE(kx && ky && tvec && sens && fieldmap && shotphase, PG_ERR_ARG, "null argument");
    PG_REQUIRE(Nz == 1 || kz, PG_ERR_ARG, "kz required for 3-D");
    PG_REQUIRE(n_shots > 0 && Nd > 0 && nc > 0, PG_ERR_ARG, "empty pcSenseTimeSeg");
    PG_REQUIRE(n_shots * nc < (1u << 30), PG_ERR_ARG, "too many tasks");
    const size_t Ni = Nx * Ny * Nz;
    std::unique_ptr<PcSenseOp> op(new PcSenseOp());
    op->kind = K_PCSENSE;
    op->Ns = (int)n_shots;
    op->Nc = (int)nc;
    op->Nd = Nd;
    op->Ni = Ni;
    op->rank_share = rank_share;
    const size_t ntot = n_shots * nc;
    // MPI/mpipcSENSETimeSeg.cpp:106-124: task t is shot t / Nc, coil t % Nc,
    // dealt round robin to the ranks
    const size_t first = rank_share ? rank : 0, step = rank_share ? nranks : 1;
    op->task0.assign(n_shots, 0);
    op->ntask.assign(n_shots, 0);
    for (size_t t = first; t < ntot; t += step) {
        const int sh = (int)(t / nc), c = (int)(t % nc);
        if (op->ntask[sh] == 0) op->task0[sh] = (int)op->tshot.size();
        ++op->ntask[sh];
        op->tcol.push_back(rank_share ? (int)op->tshot.size() : sh + c * (int)n_shots);
        op->tshot.push_back(sh);
        op->tcoil.push_back(c);
    }
    const int nt = op->tasks();
    PG_REQUIRE(nt > 0, PG_ERR_ARG, "this rank has no tasks (rank >= shots x coils)");
    // adjoint sum order: coils outer, shots inner (pcSenseTimeSeg.cpp:160-167),
    // or the rank's task order (MPI/mpipcSENSETimeSeg.cpp:218-228)
    for (int k = 0; k < nt; ++k) op->adj_order.push_back(k);
    if (!rank_share)
        std::sort(op->adj_order.begin(), op->adj_order.end(), [&](int a, int b) {
            return op->tcoil[a] != op->tcoil[b] ? op->tcoil[a] < op->tcoil[b]
                                                : op->tshot[a] < op->tshot[b];
        });
    op->n_in = Ni;
    op->n_out = Nd * (size_t)nt;
    // the interpolator's shots argument: 1 in pcSenseTimeSeg.cpp:99-100, Ns in
    // MPI/mpipcSENSETimeSeg.cpp:146-148 (there sample i of a shot takes the
    // weights of sample i mod Nd/Ns)
    const size_t ts_shots = rank_share ? n_shots : 1;
    PG_REQUIRE(Nd % ts_shots == 0, PG_ERR_ARG,
               "mpipcSENSETimeSeg needs the samples per shot to be a multiple of the shot count");
    op->G.assign(n_shots, nullptr);
    op->A.assign(n_shots, nullptr);
    std::vector<float> zk(Nd, 0.0f);
    for (size_t sh = 0; sh < n_shots; ++sh) {
        if (!op->ntask[sh]) continue;
        // Gnufft(Nd, 2.0, ...) and TimeSegmentation over the shot's samples
        std::unique_ptr<NufftOp> g(new NufftOp());
        g->kind = K_NUFFT;
        const size_t o = sh * Nd;
        try {
            gnufft_setup(g->G, Nd, 2.0f, Nx, Ny, Nz, kx + o, ky + o, kz ? kz + o : zk.data(),
                         kernelWidth);
        } catch (...) {
            g->G.release();
            throw;
        }
        g->n_in = Ni;
        g->n_out = Nd;
        op->G[sh] = g.release();
        op->A[sh] = ts_build(op->G[sh], fieldmap, tvec + o, Nd, Ni, L, interptype, ts_shots);
        op->L = static_cast<TsOp*>(op->A[sh])->L;
    }
    // pcSenseTimeSeg.cpp:104-115, in double: S_c exp(-i P_s), stored per task
    std::vector<float2> ss((size_t)nt * Ni);
    for (int k = 0; k < nt; ++k) {
        const size_t c = (size_t)op->tcoil[k], sh = (size_t)op->tshot[k];
        for (size_t n = 0; n < Ni; ++n) {
            const double sr = sens[2 * (c * Ni + n)], si = sens[2 * (c * Ni + n) + 1];
            const double ph = shotphase[sh * Ni + n];
            const double er = std::cos(ph), ei = -std::sin(ph);
            ss[(size_t)k * Ni + n] = make_float2((float)(sr * er - si * ei),
                                                 (float)(sr * ei + si * er));
        }
    }
    op->SS = dev_copy(ss.data(), ss.size());
    op->P = dev_alloc<float2>((size_t)nt * Ni);
    op->tmp_n = dev_alloc<float2>(Ni);
    return op.release();
}

int pg_pcsense_create(pg_op* out, const float* kx, const float* ky, const float* kz, size_t Nx,
                      size_t Ny, size_t Nz, size_t nc, const float* tvec, size_t L,
                      size_t interptype, const float* sens, const float* fieldmap,
                      const float* shotphase, size_t n_shots, size_t Nd, float kernelWidth)
{
    PG_TRY PG_REQUIRE(out, PG_ERR_ARG, "null argument");
    *out = nullptr;
    *out = reinterpret_cast<pg_op>(pcsense_build(kx, ky, kz, Nx, Ny, Nz, nc, tvec, L, interptype,
                                                 sens, fieldmap, shotphase, n_shots, Nd,
                                                 kernelWidth, false, 0, 1));
    PG_CATCH
}

int pg_pcsense_rank_create(pg_op* out, const float* kx, const float* ky, const float* kz,
                           size_t Nx, size_t Ny, size_t Nz, size_t nc, const float* tvec, size_t L,
                           size_t interptype, const float* sens, const float* fieldmap,
                           const float* shotphase, size_t n_shots, size_t Nd, float kernelWidth,
                           size_t rank, size_t nranks)
{
    PG_TRY PG_REQUIRE(out, PG_ERR_ARG, "null argument");
    PG_REQUIRE(nranks > 0 && rank < nranks, PG_ERR_ARG, "rank must be below nranks");
    *out = nullptr;
    *out = reinterpret_cast<pg_op>(pcsense_build(kx, ky, kz, Nx, Ny, Nz, nc, tvec, L, interptype,
                                                 sens, fieldmap, shotphase, n_shots, Nd,
                                                 kernelWidth, true, rank, nranks));
    PG_CATCH
}

int pg_pcsense_tasks(pg_op h, size_t* ntask, uint32_t* shot, uint32_t* coil, uint32_t* column)
{
    PG_TRY Op* op = as_op(h);
    PG_REQUIRE(op->kind == K_PCSENSE, PG_ERR_ARG, "not a pcSenseTimeSeg operator");
    PG_REQUIRE(ntask, PG_ERR_ARG, "null argument");
    const PcSenseOp* pc = static_cast<const PcSenseOp*>(op);
    *ntask = (size_t)pc->tasks();
    for (int k = 0; k < pc->tasks(); ++k) {
        if (shot) shot[k] = (uint32_t)pc->tshot[k];
        if (coil) coil[k] = (uint32_t)pc->tcoil[k];
        if (column) column[k] = (uint32_t)pc->tcol[k];
    }
    PG_CATCH
}

int pg_pcsense_gdft_create(pg_op* out, const float* kx, const float* ky, const float* kz,
                           size_t Nx, size_t Ny, size_t Nz, size_t nc, const float* tvec,
                           const float* sens, const float* fieldmap, const float* shotphase,
                           size_t n_shots, size_t Nd)
{
    PG_TRY PG_REQUIRE(out && kx && ky && tvec && sens && fieldmap && shotphase, PG_ERR_ARG,
                      "null argument");
    PG_REQUIRE(n_shots > 0 && Nd > 0 && nc > 0, PG_ERR_ARG, "empty pcSENSE");
    const size_t Ni = Nx * Ny * Nz;
    PG_REQUIRE(Ni > 0 && Ni < (1u << 31) && Nd < (1u << 31), PG_ERR_ARG, "bad sizes");
    *out = nullptr;
    std::unique_ptr<PcDftOp> op(new PcDftOp());
    op->kind = K_PCSENSE_DFT;
    op->Ns = (int)n_shots;
    op->Nc = (int)nc;
    op->Nd = Nd;
    op->Ni = Ni;
    op->n_in = Ni;
    op->n_out = Nd * n_shots * nc;
    // pcSENSE.cpp:62-84: (x - N/2)/N in double, rounded to float, y fastest
    std::vector<float> ix(Ni), iy(Ni), iz(Ni), zk(Nd, 0.0f);
    for (size_t z = 0; z < Nz; ++z)
        for (size_t x = 0; x < Nx; ++x)
            for (size_t y = 0; y < Ny; ++y) {
                const size_t n = y + x * Ny + z * Nx * Ny;
                ix[n] = (float)(((double)(float)x - (double)(float)Nx / 2.0) / (double)(float)Nx);
                iy[n] = (float)(((double)(float)y - (double)(float)Ny / 2.0) / (double)(float)Ny);
                iz[n] = (float)(((double)(float)z - (double)(float)Nz / 2.0) / (double)(float)Nz);
            }
    // pcSENSE.cpp:88-93: Gdft(Nd, Ni, Kx.col(s), ..., FMap, Tvec.col(s))
    for (size_t sh = 0; sh < n_shots; ++sh) {
        const size_t o = sh * Nd;
        pg_op h = nullptr;
        const int st = pg_gdft_create(&h, Nd, Ni, kx + o, ky + o, kz ? kz + o : zk.data(), ix.data(),
                                      iy.data(), iz.data(), fieldmap, tvec + o, 0);
        if (st != PG_OK) throw pg::Error{st, pg::g_last_error};
        op->D.push_back(reinterpret_cast<DftOp*>(h));
    }
    {
        // the same shots as one DFT (batched mode)
        std::vector<float> zall;
        if (!kz) zall.assign(Nd * n_shots, 0.0f);
        pg_op h = nullptr;
        const int st = pg_gdft_create(&h, Nd * n_shots, Ni, kx, ky, kz ? kz : zall.data(), ix.data(),
                                      iy.data(), iz.data(), fieldmap, tvec, 0);
        if (st != PG_OK) throw pg::Error{st, pg::g_last_error};
        op->all = reinterpret_cast<DftOp*>(h);
    }
    std::vector<float2> sf(n_shots * nc * Ni), sa(n_shots * nc * Ni);
    for (size_t sh = 0; sh < n_shots; ++sh)
        for (size_t c = 0; c < nc; ++c)
            for (size_t n = 0; n < Ni; ++n) {
                const double sr = sens[2 * (c * Ni + n)], si = sens[2 * (c * Ni + n) + 1];
                const double ph = shotphase[sh * Ni + n];
                const double er = std::cos(ph), ei = std::sin(ph);
                const size_t k = (sh * nc + c) * Ni + n;
                sf[k] = make_float2((float)(sr * er - si * ei), (float)(sr * ei + si * er));
                sa[k] = make_float2((float)(sr * er + si * ei), (float)(si * er - sr * ei));
            }
    op->SF = dev_copy(sf.data(), sf.size());
    op->SA = dev_copy(sa.data(), sa.size());
    op->tmp_n = dev_alloc<float2>(Ni);
    *out = reinterpret_cast<pg_op>(op.release());
    PG_CATCH
}

int pg_sense_create(pg_op* out, pg_op Gh, const float* sens, size_t M, size_t N, size_t nc)
{
    PG_TRY TraceRange tr_("SENSE::SENSE");
    PG_REQUIRE(out && sens && nc > 0, PG_ERR_ARG, "null argument");
    *out = nullptr;
    Op* G = as_op(Gh);
    PG_REQUIRE(G->n_in == N && G->n_out == M, PG_ERR_ARG, "SENSE sizes do not match G");
    std::unique_ptr<SenseOp> op(new SenseOp());
    op->kind = K_SENSE;
    op->inner = G;
    op->nc = (int)nc;
    op->M = M;
    op->N = N;
    op->n_in = N;
    op->n_out = M * nc;
    op->S = dev_copy(reinterpret_cast<const float2*>(sens), N * nc);
    op->P = dev_alloc<float2>(N * nc);
    op->tmp_n = dev_alloc<float2>(N);
    op->tmp_n2 = dev_alloc<float2>(N);
    *out = reinterpret_cast<pg_op>(op.release());
    PG_CATCH
}

int pg_op_destroy(pg_op h)
{
    PG_TRY if (!h) return PG_OK;
    Op* op = as_op(h);
    op->magic = 0;
    delete op;
    PG_CATCH
}

int pg_op_shape(pg_op h, size_t* n_in, size_t* n_out)
{
    PG_TRY Op* op = as_op(h);
    if (n_in) *n_in = op->n_in;
    if (n_out) *n_out = op->n_out;
    PG_CATCH
}

int pg_op_set_mode(pg_op h, int mode)
{
    PG_TRY Op* op = as_op(h);
    PG_REQUIRE(mode == PG_MODE_BATCHED || mode == PG_MODE_PLAIN, PG_ERR_ARG, "bad mode");
    // solve graphs captured before (on this operator or on a composite over
    // it) ran the old mode's kernels: every cached graph is recaptured
    if (op->mode != mode) g_ws_gen.fetch_add(1);
    op->mode = mode;
    if (op->kind == K_PCSENSE) {
        // the per-shot operators are internal: they follow the composite
        PcSenseOp* pc = static_cast<PcSenseOp*>(op);
        for (Op* o : pc->A)
            if (o) o->mode = mode;
        for (Op* o : pc->G)
            if (o) o->mode = mode;
    }
    PG_CATCH
}

int pg_op_forward(pg_op h, const void* in, void* out, void* stream)
{
    PG_TRY Op* op = as_op(h);
    PG_REQUIRE(in && out && in != out, PG_ERR_ARG, "bad buffers");
    // launches on the legacy (null) stream hold the legacy lock
    std::unique_lock<std::recursive_mutex> lk(g_legacy_mu, std::defer_lock);
    if (!stream) lk.lock();
    op_forward(op, (const float2*)in, (float2*)out, (hipStream_t)stream);
    PG_HIP(hipGetLastError());
    PG_CATCH
}

int pg_op_adjoint(pg_op h, const void* in, void* out, void* stream)
{
    PG_TRY Op* op = as_op(h);
    PG_REQUIRE(in && out && in != out, PG_ERR_ARG, "bad buffers");
    // launches on the legacy (null) stream hold the legacy lock
    std::unique_lock<std::recursive_mutex> lk(g_legacy_mu, std::defer_lock);
    if (!stream) lk.lock();
    op_adjoint(op, (const float2*)in, (float2*)out, (hipStream_t)stream);
    PG_HIP(hipGetLastError());
    PG_CATCH
}

int pg_dft_apply_static(pg_op h, int adjoint, const void* in, void* out, void* stream)
{
    PG_TRY Op* op = as_op(h);
    PG_REQUIRE(in && out && in != out, PG_ERR_ARG, "bad buffers");
    DftOp* d = nullptr;
    const float2* S = nullptr;
    int nc = 1;
    if (op->kind == K_DFT) {
        d = static_cast<DftOp*>(op);
    } else if (op->kind == K_SENSE && static_cast<SenseOp*>(op)->inner->kind == K_DFT) {
        SenseOp* se = static_cast<SenseOp*>(op);
        PG_REQUIRE(se->mode == PG_MODE_BATCHED, PG_ERR_STATE,
                   "pg_dft_apply_static runs the all-coil kernels: the SENSE operator must be in "
                   "PG_MODE_BATCHED");
        d = static_cast<DftOp*>(se->inner);
        S = se->S;
        nc = se->nc;
    }
    PG_REQUIRE(d, PG_ERR_UNSUPPORTED,
               "pg_dft_apply_static takes a Gdft or GdftR2 handle, or a SENSE handle over one");
    const size_t need = pgd::part_elems(d->d.M, d->d.N, nc, adjoint != 0);
    const size_t cap = pgd::static_scratch_elems();
    PG_REQUIRE(need <= cap, PG_ERR_ARG,
               "this operator's " + std::string(adjoint ? "adjoint" : "forward") + " needs " +
                   std::to_string(need) + " complex partial sums; the static scratch holds " +
                   std::to_string(cap));
    std::unique_lock<std::recursive_mutex> lk(g_legacy_mu, std::defer_lock);
    if (!stream) lk.lock();
    float2* part = pgd::static_scratch();
    const hipStream_t s = (hipStream_t)stream;
    if (adjoint)
        pgd::adjoint(d->d, (const float2*)in, S, nc, (float2*)out, part, s);
    else
        pgd::forward(d->d, (const float2*)in, S, nc, (float2*)out, part, s);
    PG_HIP(hipGetLastError());
    PG_CATCH
}

static int host_apply(pg_op h, const void* in, void* out, bool fwd)
{
    PG_TRY LegacyLock lk(g_legacy_mu);
    Op* op = as_op(h);
    PG_REQUIRE(in && out, PG_ERR_ARG, "bad buffers");
    const size_t ni = fwd ? op->n_in : op->n_out, no = fwd ? op->n_out : op->n_in;
    float2* di = dev_copy((const float2*)in, ni);
    float2* dout = nullptr;
    try {
        dout = dev_alloc<float2>(no);
        if (fwd) op_forward(op, di, dout, nullptr);
        else op_adjoint(op, di, dout, nullptr);
        PG_HIP(hipGetLastError());
        PG_HIP(hipMemcpy(out, dout, sizeof(float2) * no, hipMemcpyDeviceToHost));
    } catch (...) {
        dev_free(di);
        dev_free(dout);
        throw;
    }
    dev_free(di);
    dev_free(dout);
    PG_CATCH
}

int pg_op_forward_host(pg_op h, const void* in, void* out) { return host_apply(h, in, out, true); }
int pg_op_adjoint_host(pg_op h, const void* in, void* out) { return host_apply(h, in, out, false); }

// ------------------------------------- gridding alone, direct recon ----

static NufftOp* as_nufft(pg_op h)
{
    Op* op = as_op(h);
    PG_REQUIRE(op->kind == K_NUFFT, PG_ERR_ARG, "not a Gnufft handle");
    return static_cast<NufftOp*>(op);
}

int pg_gnufft_grid_size(pg_op h, size_t* Gx, size_t* Gy, size_t* Gz)
{
    PG_TRY const pg::Geometry& g = as_nufft(h)->G.geo;
    if (Gx) *Gx = (size_t)g.Gx;
    if (Gy) *Gy = (size_t)g.Gy;
    if (Gz) *Gz = (size_t)g.Gz;
    PG_CATCH
}

int pg_gnufft_forward_spatial(pg_op h, const void* grid, void* samples, void* stream)
{
    PG_TRY Gnufft& G = as_nufft(h)->G;
    PG_REQUIRE(grid && samples && grid != samples, PG_ERR_ARG, "bad buffers");
    hipStream_t s = (hipStream_t)stream;
    {
        Prof p("interp", s);
        pgk::launch_interp(G, (const float2*)grid, 1, 1, nullptr, nullptr, (float2*)samples, G.M, s);
    }
    PG_HIP(hipGetLastError());
    PG_CATCH
}

int pg_gnufft_adjoint_spatial(pg_op h, const void* samples, void* grid, void* stream)
{
    PG_TRY Gnufft& G = as_nufft(h)->G;
    PG_REQUIRE(grid && samples && grid != samples, PG_ERR_ARG, "bad buffers");
    hipStream_t s = (hipStream_t)stream;
    G.ensure_batch(1);
    {
        Prof p("spread", s);
        pgk::launch_spread(G, (const float2*)samples, G.M, 1, 1, nullptr, nullptr, nullptr,
                           nullptr, nullptr, (float2*)grid, false, s);
    }
    PG_HIP(hipGetLastError());
    PG_CATCH
}

int pg_density_compensation(const float* kx, const float* ky, const float* kz, size_t M,
                            size_t Ninplane, size_t Nz, float* dcf, void* stream)
{
    PG_TRY PG_REQUIRE(kx && ky && dcf && M > 0 && Ninplane > 0 && Nz > 0, PG_ERR_ARG,
                      "bad arguments");
    hipStream_t s = (hipStream_t)stream;
    // directRecon.cpp:31-35: a Gnufft on a 3x oversampled grid; its kernel
    // width is the reference Gnufft's fixed 4.0 (Gnufft.cpp:84)
    std::unique_ptr<NufftOp> op(new NufftOp);
    op->kind = K_NUFFT;
    // (the taps scaled by a power of two: the result is invariant to the
    // kernel's scale, and the float instantiation's sums, ~1e40 in 3-D,
    // would overflow unscaled)
    gnufft_setup(op->G, M, 3.0f, Ninplane, Ninplane, Nz, kx, ky, kz, 4.0f, true);
    Gnufft& G = op->G;
    G.ensure_batch(1);
    struct Tmp {
        float2* d = nullptr;
        float2* c1 = nullptr;
        float2* c2 = nullptr;
        double* part = nullptr;
        ~Tmp()
        {
            dev_free(d);
            dev_free(c1);
            dev_free(c2);
            dev_free(part);
        }
    } t;
    const long long Mm = (long long)M;
    t.d = dev_alloc<float2>(M);
    t.c1 = dev_alloc<float2>((size_t)G.geo.G());
    t.c2 = dev_alloc<float2>(M);
    t.part = dev_alloc<double>(pgr::kMaxParts);
    // directRecon.cpp:38-39: ones / (M * M), in float
    const float den = (float)M * (float)M;
    pgv::fill(t.d, Mm, make_float2(1.0f / den, 0.0f), s);
    // directRecon.cpp:46-52, Niter = 2
    for (int it = 0; it < 2; ++it) {
        {
            Prof p("spread", s);
            pgk::launch_spread(G, t.d, Mm, 1, 1, nullptr, nullptr, nullptr, nullptr, nullptr, t.c1,
                               false, s);
        }
        {
            Prof p("interp", s);
            pgk::launch_interp(G, t.c1, 1, 1, nullptr, nullptr, t.c2, Mm, s);
        }
        pgr::dcf_update(t.d, t.c2, Mm, s);
    }
    // directRecon.cpp:54-56
    pgr::dcf_finish(t.d, Mm, t.part, dcf, s);
    PG_HIP(hipGetLastError());
    // the temporaries are freed on return: wait for the kernels that use them
    PG_HIP(hipStreamSynchronize(s));
    PG_CATCH
}

int pg_grid_coil_images(pg_op h, const float* dcf, const void* data, size_t ncoils, void* images,
                        void* stream)
{
    PG_TRY NufftOp* no = as_nufft(h);
    PG_REQUIRE(dcf && data && images && ncoils > 0 && ncoils <= 4096, PG_ERR_ARG, "bad arguments");
    hipStream_t s = (hipStream_t)stream;
    Gnufft& G = no->G;
    const long long M = G.M, N = G.geo.N();
    const size_t need = ncoils * (size_t)M;
    if (need > no->wtmp_cap) {
        g_ws_gen.fetch_add(1);
        dev_free(no->wtmp);
        no->wtmp = dev_alloc<float2>(need);
        no->wtmp_cap = need;
    }
    // directRecon.cpp:123: G / (dcf % data_c) for every coil
    pgr::weight_columns(no->wtmp, dcf, (const float2*)data, M, (int)ncoils, s);
    Fused f;
    f.G = &G;
    if (no->mode == PG_MODE_BATCHED) {
        f.nc = (int)ncoils;
        fused_adjoint_partial(f, no->wtmp, nullptr, nullptr, nullptr, (float2*)images, s);
    } else {
        for (size_t c = 0; c < ncoils; ++c)
            fused_adjoint_partial(f, no->wtmp + c * (size_t)M, nullptr, nullptr, nullptr,
                                  (float2*)images + c * (size_t)N, s);
    }
    PG_HIP(hipGetLastError());
    PG_CATCH
}

int pg_sum_of_squares(const void* x, size_t nvox, size_t ncoils, float* out, void* stream)
{
    PG_TRY PG_REQUIRE(x && out && nvox > 0 && ncoils > 0 && ncoils < (1u << 20), PG_ERR_ARG,
                      "bad arguments");
    hipStream_t s = (hipStream_t)stream;
    {
        Prof p("sum_of_squares", s);
        pgr::sum_of_magnitudes((const float2*)x, (long long)nvox, (int)ncoils, out, s);
    }
    PG_HIP(hipGetLastError());
    PG_CATCH
}

// processNIFTI.hpp:41-101: a single-file NIfTI-1 image of float32 values.
// Header fields at the NIfTI-1 offsets; those the reference leaves
// uninitialised are zero here.
static void write_nifti_real(const std::string& path, const float* data, size_t Nx, size_t Ny,
                             size_t Nz, size_t Nslab, size_t NTR)
{
    unsigned char h[352];
    std::memset(h, 0, sizeof h);
    auto put32 = [&](int off, int32_t v) { std::memcpy(h + off, &v, 4); };
    auto put16 = [&](int off, int16_t v) { std::memcpy(h + off, &v, 2); };
    auto putf = [&](int off, float v) { std::memcpy(h + off, &v, 4); };
    put32(0, 348);  // sizeof_hdr
    const int16_t dim[8] = {3, (int16_t)Nx, (int16_t)Ny, (int16_t)Nz, (int16_t)Nslab,
                            (int16_t)NTR, 1, 1};
    for (int k = 0; k < 8; ++k) put16(40 + 2 * k, dim[k]);
    put16(70, 16);  // datatype NIFTI_TYPE_FLOAT32
    put16(72, 32);  // bitpix
    for (int k = 0; k < 4; ++k) putf(76 + 4 * k, 1.0f);  // pixdim[0..3]
    putf(108, 352.0f);  // vox_offset
    putf(112, 100.0f);  // scl_slope (as the reference writes it)
    h[123] = 2;         // xyzt_units NIFTI_UNITS_MM
    std::memcpy(h + 344, "n+1\0", 4);  // magic; bytes 348-351: the zero extender
    const size_t n = Nx * Ny * Nz * Nslab * NTR;
    FILE* f = std::fopen(path.c_str(), "wb");
    PG_REQUIRE(f, PG_ERR_ARG, "cannot open " + path + " for writing");
    const bool ok = std::fwrite(h, 1, sizeof h, f) == sizeof h &&
                    std::fwrite(data, sizeof(float), n, f) == n;
    const bool closed = std::fclose(f) == 0;
    PG_REQUIRE(ok && closed, PG_ERR_STATE, "short write to " + path);
}

int pg_write_nifti_real(const char* filename, const float* data, size_t Nx, size_t Ny, size_t Nz,
                        size_t Nslab, size_t NTR)
{
    PG_TRY PG_REQUIRE(filename && data, PG_ERR_ARG, "null argument");
    PG_REQUIRE(Nx && Ny && Nz && Nslab && NTR, PG_ERR_ARG, "empty image");
    PG_REQUIRE(Nx < 32768 && Ny < 32768 && Nz < 32768 && Nslab < 32768 && NTR < 32768,
               PG_ERR_ARG, "NIfTI-1 dimensions are 16-bit");
    write_nifti_real(std::string(filename) + ".nii", data, Nx, Ny, Nz, Nslab, NTR);
    PG_CATCH
}

int pg_write_nifti_mag_phase(const char* filename, const void* image_dev, size_t Nx, size_t Ny,
                             size_t Nz, size_t Nslab, size_t NTR, void* stream)
{
    PG_TRY PG_REQUIRE(filename && image_dev, PG_ERR_ARG, "null argument");
    PG_REQUIRE(Nx && Ny && Nz && Nslab && NTR, PG_ERR_ARG, "empty image");
    PG_REQUIRE(Nx < 32768 && Ny < 32768 && Nz < 32768 && Nslab < 32768 && NTR < 32768,
               PG_ERR_ARG, "NIfTI-1 dimensions are 16-bit");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = Nx * Ny * Nz * Nslab * NTR;
    struct Tmp {
        float* d = nullptr;
        ~Tmp() { dev_free(d); }
    } t;
    t.d = dev_alloc<float>(2 * n);
    {
        Prof p("mag_phase", s);
        pgr::mag_phase((const float2*)image_dev, (long long)n, t.d, t.d + n, s);
    }
    PG_HIP(hipGetLastError());
    std::vector<float> h(2 * n);
    PG_HIP(hipMemcpyAsync(h.data(), t.d, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    // processNIFTI.hpp:106-119
    write_nifti_real(std::string(filename) + "_mag.nii", h.data(), Nx, Ny, Nz, Nslab, NTR);
    write_nifti_real(std::string(filename) + "_phs.nii", h.data() + n, Nx, Ny, Nz, Nslab, NTR);
    PG_CATCH
}

int pg_penalty_create(pg_penalty* out, int kind, size_t Nx, size_t Ny, size_t Nz, float beta,
                      float delta, size_t dims2penalize)
{
    PG_TRY PG_REQUIRE(out, PG_ERR_ARG, "null out");
    PG_REQUIRE(kind == PG_PENALTY_QUAD || kind == PG_PENALTY_TV, PG_ERR_ARG, "bad penalty kind");
    PG_REQUIRE(Nx > 0 && Ny > 0 && Nz > 0, PG_ERR_ARG, "bad sizes");
    PG_REQUIRE(kind == PG_PENALTY_QUAD || delta != 0.0f, PG_ERR_ARG, "TV needs delta != 0");
    *out = nullptr;
    std::unique_ptr<Penalty> p(new Penalty());
    p->spec.kind = kind;
    p->spec.nd = ((Nz == 1) || (dims2penalize == 2)) ? 2 : 3;
    p->spec.n = (long long)(Nx * Ny * Nz);
    p->spec.off[0] = 1;
    p->spec.off[1] = (long long)Ny;
    p->spec.off[2] = (long long)(Nx * Ny);
    p->spec.beta = beta;
    p->spec.delta = delta;
    PG_REQUIRE(p->spec.n > p->spec.off[p->spec.nd - 1], PG_ERR_ARG, "image too small for the penalty");
    p->part = dev_alloc<double>(2 * pgv::kMaxParts);
    p->out = dev_alloc<double>(2);
    *out = reinterpret_cast<pg_penalty>(p.release());
    PG_CATCH
}

int pg_penalty_destroy(pg_penalty h)
{
    PG_TRY if (!h) return PG_OK;
    Penalty* p = as_pen(h);
    p->magic = 0;
    delete p;
    PG_CATCH
}

int pg_penalty_gradient(pg_penalty h, const void* x, void* g, void* stream)
{
    PG_TRY TraceRange tr_("Robject::Gradient");
    Penalty* p = as_pen(h);
    pgv::pen_gradient(p->spec, (const float2*)x, (float2*)g, (hipStream_t)stream);
    PG_HIP(hipGetLastError());
    PG_CATCH
}

int pg_penalty_denom(pg_penalty h, const void* ddir, const void* x, double* out2, void* stream)
{
    PG_TRY TraceRange tr_("Robject::Denom");
    Penalty* p = as_pen(h);
    PG_REQUIRE(out2, PG_ERR_ARG, "null out2");
    hipStream_t s = (hipStream_t)stream;
    pgv::pen_denom_pdot(p->spec, (const float2*)ddir, (const float2*)x, nullptr, false, p->part,
                        p->out, nullptr, s);
    double h2[2];
    PG_HIP(hipMemcpyAsync(h2, p->out, sizeof(h2), hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    out2[0] = h2[0];
    out2[1] = 0.0;
    PG_CATCH
}

int pg_penalty_value(pg_penalty h, const void* x, double* out, void* stream)
{
    PG_TRY TraceRange tr_("Robject::Penalty");
    Penalty* p = as_pen(h);
    PG_REQUIRE(out, PG_ERR_ARG, "null out");
    hipStream_t s = (hipStream_t)stream;
    pgv::pen_value(p->spec, (const float2*)x, p->part, p->out, s);
    double v;
    PG_HIP(hipMemcpyAsync(&v, p->out, sizeof(double), hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    *out = v * p->spec.beta;
    PG_CATCH
}

int pg_solve_pwls_pcg_tol(pg_op Ah, pg_penalty Rh, const void* x0, const float* W, const void* y,
                          size_t niter, double tol, int mode, void* x_out, double* resid,
                          pg_cg_info* info, void* stream)
{
    PG_TRY TraceRange tr_("solve_pwls_pcg");
    Op* A = as_op(Ah);
    Penalty* R = Rh ? as_pen(Rh) : nullptr;
    PG_REQUIRE(y && x_out, PG_ERR_ARG, "null y or x_out");
    PG_REQUIRE(mode == PG_CG_GRAPH || mode == PG_CG_HOST, PG_ERR_ARG, "bad CG mode");
    if (R) PG_REQUIRE((size_t)R->spec.n == A->n_in, PG_ERR_ARG, "penalty size != image size");
    solve_pcg(A, R, (const float2*)x0, W, (const float2*)y, niter, tol, mode, (float2*)x_out, resid,
              info, (hipStream_t)stream);
    PG_CATCH
}

int pg_solve_pwls_pcg(pg_op A, pg_penalty R, const void* x0, const float* W, const void* y,
                      size_t niter, int mode, void* x_out, double* resid, pg_cg_info* info,
                      void* stream)
{
    return pg_solve_pwls_pcg_tol(A, R, x0, W, y, niter, 0.0, mode, x_out, resid, info, stream);
}

int pg_recon_solve(pg_op A, pg_penalty R, const void* y, size_t niter, int mode, void* x_out,
                   double* resid, pg_cg_info* info, void* stream)
{
    return pg_solve_pwls_pcg_tol(A, R, nullptr, nullptr, y, niter, 0.0, mode, x_out, resid, info,
                                 stream);
}

int pg_recon_solve_tol(pg_op A, pg_penalty R, const void* y, size_t niter, double tol, int mode,
                       void* x_out, double* resid, pg_cg_info* info, void* stream)
{
    return pg_solve_pwls_pcg_tol(A, R, nullptr, nullptr, y, niter, tol, mode, x_out, resid, info,
                                 stream);
}

int pg_recon_solve_batch(size_t K, const pg_op* A, const pg_penalty* R, const void* const* y,
                         size_t niter, double tol, void* const* x_out, pg_cg_info* infos,
                         void* stream)
{
    PG_TRY TraceRange tr_("reconSolve batch");
    PG_REQUIRE(K > 0 && A && R && y && x_out, PG_ERR_ARG, "null argument");
    PG_REQUIRE(stream, PG_ERR_ARG, "a batch runs on a stream of the caller (pg_stream_create)");
    std::vector<Op*> ops(K);
    std::vector<Penalty*> pens(K);
    std::vector<const float2*> ys(K);
    std::vector<float2*> xs(K);
    std::vector<const Op*> seen;
    for (size_t k = 0; k < K; ++k) {
        ops[k] = as_op(A[k]);
        pens[k] = R[k] ? as_pen(R[k]) : nullptr;
        PG_REQUIRE(y[k] && x_out[k], PG_ERR_ARG, "null y or x_out");
        if (pens[k])
            PG_REQUIRE((size_t)pens[k]->spec.n == ops[k]->n_in, PG_ERR_ARG,
                       "penalty size != image size");
        std::vector<const Op*> mine;
        collect_ops(ops[k], mine);
        for (const Op* o : mine)
            PG_REQUIRE(std::find(seen.begin(), seen.end(), o) == seen.end(), PG_ERR_ARG,
                       "the reconstructions of a batch must not share operators (their "
                       "workspaces run concurrently)");
        seen.insert(seen.end(), mine.begin(), mine.end());
        ys[k] = (const float2*)y[k];
        xs[k] = (float2*)x_out[k];
    }
    solve_batch(K, ops.data(), pens.data(), ys.data(), niter, tol, xs.data(), infos,
                (hipStream_t)stream);
    PG_CATCH
}

int pg_op_tuning(pg_op h, int B, int* interp_lds, int* spread_mfma)
{
    PG_TRY Op* op = as_op(h);
    PG_REQUIRE(interp_lds && spread_mfma, PG_ERR_ARG, "null argument");
    if (op->kind == K_SENSE) op = static_cast<SenseOp*>(op)->inner;
    if (op->kind == K_TS) op = static_cast<TsOp*>(op)->inner;
    PG_REQUIRE(op->kind == K_NUFFT, PG_ERR_ARG, "not a Gnufft or an operator over one");
    const Gnufft::Tuned* t = static_cast<NufftOp*>(op)->G.tuned_for(B);
    *interp_lds = t ? t->lds : -1;
    *spread_mfma = t ? t->mfma : -1;
    PG_CATCH
}

int pg_gnufft_zclasses(pg_op h, size_t* nzc)
{
    PG_TRY Op* op = as_op(h);
    PG_REQUIRE(nzc, PG_ERR_ARG, "null argument");
    if (op->kind == K_SENSE) op = static_cast<SenseOp*>(op)->inner;
    if (op->kind == K_TS) op = static_cast<TsOp*>(op)->inner;
    PG_REQUIRE(op->kind == K_NUFFT, PG_ERR_ARG, "not a Gnufft or an operator over one");
    *nzc = (size_t)static_cast<NufftOp*>(op)->G.nzc;
    PG_CATCH
}

int pg_set_tuning(const char* key, int value)
{
    PG_TRY PG_REQUIRE(key, PG_ERR_ARG, "null key");
    const std::string k(key);
    if (k == "spread_mfma")
        pgk::g_spread_mfma.store(value ? 1 : 0);
    else if (k == "zclass")
        pgk::g_zclass.store(value ? 1 : 0);
    else if (k == "spread_rows")
        pgk::g_spread_rows.store(value ? 1 : 0);
    else if (k == "trace")
        g_trace.store(value ? 1 : 0);
    else if (k == "zfuse")
        pgk::g_zfuse.store(value ? 1 : 0);
    else if (k == "zstage_mfma")
        pgk::g_zmfma.store(value ? 1 : 0);
    else if (k == "zadj_mfma")
        pgk::g_zamfma.store(value ? 1 : 0);
    else if (k == "zfwd_fuse")
        pgk::g_zfwd.store(value ? 1 : 0);


    else if (k == "spread_stage")
        pgk::g_spread_stage.store(value ? 1 : 0);

    else if (k == "batch_streams")
        g_batch_streams.store(value < 0 ? 0 : (value > 8 ? 8 : value));
    else if (k == "interp_lds4")
        pgk::g_interp4.store(value < 0 ? 0 : (value > 2 ? 2 : value));
    else if (k == "autotune")
        g_autotune.store(value ? 1 : 0);
    else
        throw pg::Error{PG_ERR_ARG, "unknown tuning key " + k};
    // cached solve graphs captured the previous kernel variants
    g_ws_gen.fetch_add(1);
    PG_CATCH
}

int pg_trace_ranges(int* available, unsigned long long* pushed)
{
    PG_TRY PG_REQUIRE(available && pushed, PG_ERR_ARG, "null argument");
    *available = roctx_api().push ? 1 : 0;
    *pushed = g_trace_pushed.load();
    PG_CATCH
}

int pg_profile_enable(int on)
{
    g_prof = on != 0;
    return PG_OK;
}

int pg_profile_reset(void)
{
    PG_TRY std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& r : g_prof_recs) {
        hipEventSynchronize(r.b);
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    g_prof_recs.clear();
    PG_CATCH
}

int pg_profile_query(const char* name, int* count, double* total_ms)
{
    PG_TRY PG_REQUIRE(name, PG_ERR_ARG, "null name");
    std::lock_guard<std::mutex> lk(g_prof_mu);
    int n = 0;
    double tot = 0;
    for (auto& r : g_prof_recs) {
        if (r.name != name) continue;
        PG_HIP(hipEventSynchronize(r.b));
        float ms = 0;
        PG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        tot += ms;
        ++n;
    }
    if (count) *count = n;
    if (total_ms) *total_ms = tot;
    PG_CATCH
}

}  // extern "C"
