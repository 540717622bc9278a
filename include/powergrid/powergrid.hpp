// powergrid/powergrid.hpp -- the reference's C++ operator API, rebuilt on the
// C ABI of libpowergrid_amd.so (include/powergrid_amd.h).
//
// Code written against PowerGrid's operators
//     Gnufft<float> G(M, 2.0, Nx, Ny, Nz, kx, ky, kz, ix, iy, iz);
//     TimeSegmentation<float, Gnufft<float>> A(G, fm, t, M, N, L, 1, 1);
//     SENSE<float, TimeSegmentation<float, Gnufft<float>>> Sg(A, sen, M, N, nc);
//     QuadPenalty<float> R(Nx, Ny, Nz, beta, dims2penalize);
//     Col<complex<float>> img = reconSolve(data, Sg, R, kx, ky, kz, Nx, Ny, Nz, t, niter);
// (PowerGridIsmrmrd.cpp:242-251) compiles against this header with the same
// class names, constructor arguments and operator* / operator/ semantics
// (Gnufft.h:51-53, TimeSegmentation.h:70-71, SENSE.h:64, Gdft.h:47-49,
// QuadPenalty.h:43, TVPenalty.h:44, solve_pwls_pcg.hpp:48-51,
// reconSolve.h:63-66, reconSolve.cpp:30-54).  Differences:
//   * Col<T> is a small contiguous host vector (Armadillo is not required);
//   * operators are single precision (float), as the device path is;
//   * composites keep a non-owning reference to the inner operator, which
//     must outlive them (SENSE.h:53, TimeSegmentation.h:55), as before;
//   * solve_pwls_pcg runs the whole PCG loop on the GPU (FP64 scalars);
//   * failures throw powergrid::Error instead of printing and continuing.
#ifndef POWERGRID_AMD_POWERGRID_HPP
#define POWERGRID_AMD_POWERGRID_HPP

#include <algorithm>
#include <complex>
#include <cstddef>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <utility>
#include <vector>

#include "../powergrid_amd.h"

namespace powergrid {

using uword = std::size_t;
using std::complex;

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

inline void check(int st)
{
    if (st != PG_OK) throw Error(st, std::string(pg_status_string(st)) + ": " + pg_last_error());
}

// --------------------------------------------------------------- Col<T> ---
template <typename T>
class Col {
public:
    Col() = default;
    explicit Col(uword n) : v_(n), n_rows(n) {}
    Col(uword n, const T& fill) : v_(n, fill), n_rows(n) {}
    Col(std::initializer_list<T> l) : v_(l), n_rows(l.size()) {}
    Col(const Col& o) : v_(o.v_), n_rows(o.v_.size()) {}
    Col(Col&& o) noexcept : v_(std::move(o.v_)), n_rows(v_.size()) { o.n_rows = 0; }
    Col& operator=(const Col& o)
    {
        v_ = o.v_;
        n_rows = v_.size();
        return *this;
    }
    Col& operator=(Col&& o) noexcept
    {
        v_ = std::move(o.v_);
        n_rows = v_.size();
        o.n_rows = 0;
        return *this;
    }
    uword size() const { return v_.size(); }
    uword n_elem() const { return v_.size(); }
    T* memptr() { return v_.data(); }
    const T* memptr() const { return v_.data(); }
    T& operator()(uword i) { return v_[i]; }
    const T& operator()(uword i) const { return v_[i]; }
    T& operator[](uword i) { return v_[i]; }
    const T& operator[](uword i) const { return v_[i]; }
    void set_size(uword n)
    {
        v_.resize(n);
        n_rows = n;
    }
    Col& zeros()
    {
        std::fill(v_.begin(), v_.end(), T(0));
        return *this;
    }
    Col& zeros(uword n)
    {
        v_.assign(n, T(0));
        n_rows = n;
        return *this;
    }
    Col& ones()
    {
        std::fill(v_.begin(), v_.end(), T(1));
        return *this;
    }
    Col& ones(uword n)
    {
        v_.assign(n, T(1));
        n_rows = n;
        return *this;
    }
    typename std::vector<T>::iterator begin() { return v_.begin(); }
    typename std::vector<T>::iterator end() { return v_.end(); }
    typename std::vector<T>::const_iterator begin() const { return v_.begin(); }
    typename std::vector<T>::const_iterator end() const { return v_.end(); }

private:
    std::vector<T> v_;

public:
    uword n_rows = 0;  // length, as Armadillo's Col::n_rows
};

// ------------------------------------------------------------ operators ---
namespace detail {

// a device buffer owned by the library's runtime (pg_malloc / pg_free)
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { check(pg_malloc(&p, bytes ? bytes : 1)); }
    DevBuf(const void* host, size_t bytes) : DevBuf(bytes)
    {
        if (bytes) check(pg_memcpy_htod(p, host, bytes, nullptr));
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf()
    {
        if (p) pg_free(p);
    }
};

class OpHandle {
public:
    OpHandle() = default;
    explicit OpHandle(pg_op h) : h_(h) { check(pg_op_shape(h_, &n_in_, &n_out_)); }
    OpHandle(const OpHandle&) = delete;
    OpHandle& operator=(const OpHandle&) = delete;
    OpHandle(OpHandle&& o) noexcept { *this = std::move(o); }
    OpHandle& operator=(OpHandle&& o) noexcept
    {
        std::swap(h_, o.h_);
        std::swap(n_in_, o.n_in_);
        std::swap(n_out_, o.n_out_);
        return *this;
    }
    ~OpHandle()
    {
        if (h_) pg_op_destroy(h_);
    }
    pg_op get() const { return h_; }
    uword n_in() const { return n_in_; }
    uword n_out() const { return n_out_; }

    Col<complex<float>> forward(const Col<complex<float>>& d) const
    {
        if (d.size() != n_in_) throw Error(PG_ERR_ARG, "operator*: input has the wrong length");
        Col<complex<float>> out(n_out_);
        check(pg_op_forward_host(h_, d.memptr(), out.memptr()));
        return out;
    }
    Col<complex<float>> adjoint(const Col<complex<float>>& d) const
    {
        if (d.size() != n_out_) throw Error(PG_ERR_ARG, "operator/: input has the wrong length");
        Col<complex<float>> out(n_in_);
        check(pg_op_adjoint_host(h_, d.memptr(), out.memptr()));
        return out;
    }

private:
    pg_op h_ = nullptr;
    uword n_in_ = 0, n_out_ = 0;
};

template <class Derived>
class OpBase {
public:
    // forward transform (Gnufft.cpp:153-175, SENSE.cpp:51-71, ...)
    Col<complex<float>> operator*(const Col<complex<float>>& d) const { return h_.forward(d); }
    // adjoint transform (Gnufft.cpp:178-211, SENSE.cpp:75-97, ...)
    Col<complex<float>> operator/(const Col<complex<float>>& d) const { return h_.adjoint(d); }
    pg_op handle() const { return h_.get(); }
    uword n_in() const { return h_.n_in(); }
    uword n_out() const { return h_.n_out(); }
    // PG_MODE_BATCHED (default) or PG_MODE_PLAIN (one NUFFT at a time)
    void set_mode(int mode) { check(pg_op_set_mode(h_.get(), mode)); }

protected:
    detail::OpHandle h_;
};

template <typename T>
void require_float()
{
    static_assert(std::is_same<T, float>::value,
                  "the device path implements the single-precision operators");
}

}  // namespace detail

template <typename T1>
class Gnufft : public detail::OpBase<Gnufft<T1>> {
public:
    // Gnufft.h:51-53; kernelWidth is fixed to 4.0 in the reference (Gnufft.cpp:84)
    Gnufft(uword dataLength, T1 gridos, uword nx, uword ny, uword nz, const Col<T1>& k1,
           const Col<T1>& k2, const Col<T1>& k3, const Col<T1>& i1, const Col<T1>& i2,
           const Col<T1>& i3, T1 kernelWidth = T1(4.0))
    {
        detail::require_float<T1>();
        pg_op h = nullptr;
        check(pg_gnufft_create(&h, dataLength, gridos, nx, ny, nz, k1.memptr(), k2.memptr(),
                               k3.size() ? k3.memptr() : nullptr, i1.size() ? i1.memptr() : nullptr,
                               i2.size() ? i2.memptr() : nullptr, i3.size() ? i3.memptr() : nullptr,
                               kernelWidth));
        this->h_ = detail::OpHandle(h);
        Nx = nx;
        Ny = ny;
        Nz = nz;
        n1 = nx * ny * nz;
        n2 = dataLength;
        size_t gx = 0, gy = 0, gz = 0;
        check(pg_gnufft_grid_size(h, &gx, &gy, &gz));
        gridNumElems = gx * gy * gz;
    }
    // Gnufft.cpp:214-263: the oversampled grid (gridNumElems, y fastest,
    // unshifted) interpolated at the samples
    Col<complex<T1>> forwardSpatialInterp(const Col<complex<T1>>& d) const
    {
        return spatial(d, gridNumElems, n2, true);
    }
    // Gnufft.cpp:266-337: the samples spread onto the zeroed grid
    Col<complex<T1>> adjointSpatialInterp(const Col<complex<T1>>& d) const
    {
        return spatial(d, n2, gridNumElems, false);
    }
    uword n1 = 0, n2 = 0, Nx = 0, Ny = 0, Nz = 0, gridNumElems = 0;

private:
    Col<complex<T1>> spatial(const Col<complex<T1>>& d, uword nin, uword nout, bool fwd) const
    {
        if (d.size() != nin) throw Error(PG_ERR_ARG, "spatial interpolation: wrong input length");
        detail::DevBuf in(d.memptr(), sizeof(complex<T1>) * nin);
        detail::DevBuf out(sizeof(complex<T1>) * nout);
        check(fwd ? pg_gnufft_forward_spatial(this->handle(), in.p, out.p, nullptr)
                  : pg_gnufft_adjoint_spatial(this->handle(), in.p, out.p, nullptr));
        Col<complex<T1>> r(nout);
        check(pg_memcpy_dtoh(r.memptr(), out.p, sizeof(complex<T1>) * nout, nullptr));
        return r;
    }
};

template <typename T1>
class Gdft : public detail::OpBase<Gdft<T1>> {
public:
    // Gdft.h:47-49; precision 1 evaluates phase and trig in FP64
    Gdft(uword a, uword b, const Col<T1>& k1, const Col<T1>& k2, const Col<T1>& k3,
         const Col<T1>& i1, const Col<T1>& i2, const Col<T1>& i3, const Col<T1>& f1,
         const Col<T1>& t1, int precision = 0)
    {
        detail::require_float<T1>();
        pg_op h = nullptr;
        check(pg_gdft_create(&h, a, b, k1.memptr(), k2.memptr(), k3.memptr(), i1.memptr(),
                             i2.memptr(), i3.memptr(), f1.memptr(), t1.memptr(), precision));
        this->h_ = detail::OpHandle(h);
        n1 = a;
        n2 = b;
    }
    uword n1 = 0, n2 = 0;
};

template <typename T1, typename Tobj>
class TimeSegmentation : public detail::OpBase<TimeSegmentation<T1, Tobj>> {
public:
    // TimeSegmentation.h:70-71 (interptype 1 Hanning, 2 exact min-max, 3 histogram min-max)
    TimeSegmentation(Tobj& G, Col<T1> map_in, Col<T1> timeVec_in, uword a, uword b,
                     uword nTimeSegs, uword interptype = 1, uword shots = 1)
        : obj(&G)
    {
        detail::require_float<T1>();
        pg_op h = nullptr;
        check(pg_ts_create(&h, G.handle(), map_in.memptr(), timeVec_in.memptr(), a, b, nTimeSegs,
                           interptype, shots));
        this->h_ = detail::OpHandle(h);
        n1 = a;
        n2 = b;
        L = (nTimeSegs == 0 || nTimeSegs == 1) ? 1 : nTimeSegs;
    }
    Tobj* obj;  // non-owning (TimeSegmentation.h:55)
    uword n1 = 0, n2 = 0, L = 1;
};

template <typename T1, typename Tobj>
class SENSE : public detail::OpBase<SENSE<T1, Tobj>> {
public:
    // SENSE.h:64
    SENSE(Tobj& G, Col<complex<T1>> SENSEmap, uword a, uword b, uword c) : G_obj(&G)
    {
        detail::require_float<T1>();
        pg_op h = nullptr;
        check(pg_sense_create(&h, G.handle(), reinterpret_cast<const float*>(SENSEmap.memptr()), a,
                              b, c));
        this->h_ = detail::OpHandle(h);
        n1 = a;
        n2 = b;
        nc = c;
    }
    Tobj* G_obj;  // non-owning (SENSE.h:53)
    uword n1 = 0, n2 = 0, nc = 0;
};

// pcSenseTimeSeg.h:48-50: kx, ky, kz and t hold the shots' samples
// shot-major; ShotPhaseMap holds one real phase map per shot (its length sets
// the shot count, as in pcSenseTimeSeg.cpp:45-46); the per-shot Gnufft and
// TimeSegmentation objects are owned by the device handle.
template <typename T1>
class pcSenseTimeSeg : public detail::OpBase<pcSenseTimeSeg<T1>> {
public:
    pcSenseTimeSeg(Col<T1> kx, Col<T1> ky, Col<T1> kz, uword nx, uword ny, uword nz, uword nc,
                   Col<T1> t, uword lSegs, uword intType, Col<complex<T1>> SENSEmap,
                   Col<T1> FieldMap, Col<T1> ShotPhaseMap, float kernelWidth = 4.0f)
    {
        detail::require_float<T1>();
        Ni = nx * ny * nz;
        Nc = nc;
        Ns = Ni ? ShotPhaseMap.n_rows / Ni : 0;
        Nd = Ns ? kx.n_rows / Ns : 0;
        if (Ns * Ni != ShotPhaseMap.n_rows || Nd * Ns != kx.n_rows)
            throw Error(PG_ERR_ARG, "pcSenseTimeSeg: array lengths do not give whole shots");
        pg_op h = nullptr;
        check(pg_pcsense_create(&h, kx.memptr(), ky.memptr(), kz.n_rows ? kz.memptr() : nullptr, nx,
                                ny, nz, nc, t.memptr(), lSegs, intType,
                                reinterpret_cast<const float*>(SENSEmap.memptr()),
                                FieldMap.memptr(), ShotPhaseMap.memptr(), Ns, Nd, kernelWidth));
        this->h_ = detail::OpHandle(h);
    }
    uword Ni = 0, Nc = 0, Ns = 0, Nd = 0;
};

// MPI/mpipcSENSETimeSeg.h: one rank's share, with the rank and rank count in
// place of the boost::mpi environment and communicator.  operator* returns
// the rank's columns in task order and operator/ takes them (the reference's
// gather, reduce and broadcasts are left to the caller); tasks() gives each
// column's shot and coil.
template <typename T1>
class mpipcSENSETimeSeg : public detail::OpBase<mpipcSENSETimeSeg<T1>> {
public:
    mpipcSENSETimeSeg(Col<T1> kx, Col<T1> ky, Col<T1> kz, uword nx, uword ny, uword nz, uword nc,
                      Col<T1> t, uword l, uword type_interp, Col<complex<T1>> SENSEmap,
                      Col<T1> FieldMap, Col<T1> ShotPhaseMap, uword rank, uword nranks,
                      float kernelWidth = 4.0f)
    {
        detail::require_float<T1>();
        Ni = nx * ny * nz;
        Nc = nc;
        Ns = Ni ? ShotPhaseMap.n_rows / Ni : 0;
        Nd = Ns ? kx.n_rows / Ns : 0;
        if (Ns * Ni != ShotPhaseMap.n_rows || Nd * Ns != kx.n_rows)
            throw Error(PG_ERR_ARG, "mpipcSENSETimeSeg: array lengths do not give whole shots");
        pg_op h = nullptr;
        check(pg_pcsense_rank_create(&h, kx.memptr(), ky.memptr(),
                                     kz.n_rows ? kz.memptr() : nullptr, nx, ny, nz, nc, t.memptr(),
                                     l, type_interp,
                                     reinterpret_cast<const float*>(SENSEmap.memptr()),
                                     FieldMap.memptr(), ShotPhaseMap.memptr(), Ns, Nd, kernelWidth,
                                     rank, nranks));
        this->h_ = detail::OpHandle(h);
    }
    // (shot, coil) of each column of operator*'s output
    std::vector<std::pair<uword, uword>> tasks() const
    {
        size_t n = 0;
        check(pg_pcsense_tasks(this->handle(), &n, nullptr, nullptr, nullptr));
        std::vector<uint32_t> s(n), c(n);
        check(pg_pcsense_tasks(this->handle(), &n, s.data(), c.data(), nullptr));
        std::vector<std::pair<uword, uword>> out(n);
        for (size_t k = 0; k < n; ++k) out[k] = {s[k], c[k]};
        return out;
    }
    uword Ni = 0, Nc = 0, Ns = 0, Nd = 0;
};

// ------------------------------------------------------------ penalties ---
namespace detail {
struct PenHandle {
    explicit PenHandle(pg_penalty p) : h(p) {}
    PenHandle(const PenHandle&) = delete;
    PenHandle& operator=(const PenHandle&) = delete;
    ~PenHandle()
    {
        if (h) pg_penalty_destroy(h);
    }
    pg_penalty h = nullptr;
};
inline std::shared_ptr<PenHandle> make_penalty(int kind, uword nx, uword ny, uword nz, float beta,
                                               float delta, uword dims2penalize)
{
    pg_penalty h = nullptr;
    check(pg_penalty_create(&h, kind, nx, ny, nz, beta, delta, dims2penalize));
    return std::make_shared<PenHandle>(h);
}
}  // namespace detail

// Penalties are cheap to copy (the reference passes them by value,
// reconSolve.h:63-66); copies share one device handle.
template <typename T1>
class QuadPenalty {
public:
    // QuadPenalty.h:43
    QuadPenalty(uword nx, uword ny, uword nz, T1 beta, uword dims2penalize = 3)
        : Nx(nx), Ny(ny), Nz(nz), Beta(beta), Dims2Penalize(dims2penalize),
          h_(detail::make_penalty(PG_PENALTY_QUAD, nx, ny, nz, beta, 0.0f, dims2penalize))
    {
        detail::require_float<T1>();
    }
    pg_penalty handle() const { return h_->h; }
    uword Nx, Ny, Nz;
    T1 Beta;
    uword Dims2Penalize;

private:
    std::shared_ptr<detail::PenHandle> h_;
};

template <typename T1>
class TVPenalty {
public:
    // TVPenalty.h:44
    TVPenalty(uword nx, uword ny, uword nz, T1 beta, T1 delta, uword dims2penalize = 3)
        : Nx(nx), Ny(ny), Nz(nz), Beta(beta), Delta(delta), Dims2Penalize(dims2penalize),
          h_(detail::make_penalty(PG_PENALTY_TV, nx, ny, nz, beta, delta, dims2penalize))
    {
        detail::require_float<T1>();
    }
    pg_penalty handle() const { return h_->h; }
    uword Nx, Ny, Nz;
    T1 Beta, Delta;
    uword Dims2Penalize;

private:
    std::shared_ptr<detail::PenHandle> h_;
};

// --------------------------------------------------------------- solver ---
// solve_pwls_pcg.hpp:48-167, run on the device.  info (optional) receives
// the completed iterations and the exit reason.
template <typename T1, typename Tobj, typename Robj>
Col<complex<T1>> solve_pwls_pcg(const Col<complex<T1>>& xInitial, Tobj const& A,
                                Col<T1> const& W, Col<complex<T1>> const& yi, Robj const& R,
                                uword niter, pg_cg_info* info = nullptr, int mode = PG_CG_GRAPH)
{
    detail::require_float<T1>();
    const uword N = A.n_in(), M = A.n_out();
    if (xInitial.size() != N || yi.size() != M || (W.size() && W.size() != M))
        throw Error(PG_ERR_ARG, "solve_pwls_pcg: vector lengths do not match the operator");
    detail::DevBuf x0(xInitial.memptr(), sizeof(complex<T1>) * N);
    detail::DevBuf y(yi.memptr(), sizeof(complex<T1>) * M);
    detail::DevBuf xo(sizeof(complex<T1>) * N);
    std::vector<double> resid(niter ? niter : 1);
    pg_cg_info inf{};
    if (W.size()) {
        detail::DevBuf w(W.memptr(), sizeof(T1) * M);
        check(pg_solve_pwls_pcg(A.handle(), R.handle(), x0.p, (const float*)w.p, y.p, niter, mode,
                                xo.p, resid.data(), &inf, nullptr));
    } else {
        check(pg_solve_pwls_pcg(A.handle(), R.handle(), x0.p, nullptr, y.p, niter, mode, xo.p,
                                resid.data(), &inf, nullptr));
    }
    Col<complex<T1>> x(N);
    check(pg_memcpy_dtoh(x.memptr(), xo.p, sizeof(complex<T1>) * N, nullptr));
    if (info) *info = inf;
    return x;
}

// reconSolve.h:63-82: W = ones, x0 = 0
template <typename T1, typename TObj, typename RObj>
Col<complex<T1>> reconSolve(Col<complex<T1>> data, TObj& Sg, RObj R, Col<T1> kx, Col<T1> ky,
                            Col<T1> kz, uword Nx, uword Ny, uword Nz, Col<T1> tvec, uword niter)
{
    (void)kx;
    (void)ky;
    (void)kz;
    (void)tvec;
    Col<complex<T1>> xinit(Nx * Ny * Nz, complex<T1>(0, 0));
    Col<T1> W;  // empty = ones (reconSolve.h:72-73)
    return solve_pwls_pcg<T1, TObj, RObj>(xinit, Sg, W, data, R, niter);
}

// ------------------------------------------------- direct reconstruction ---
// directRecon.cpp:28-60 (Pipe-Menon density compensation on a 3x grid)
template <typename T>
Col<T> calc3DDensityCompensation(Col<T> kx, Col<T> ky, Col<T> kz, uword Ninplane, uword Nz)
{
    detail::require_float<T>();
    const uword M = kx.size();
    if (ky.size() != M || (Nz > 1 && kz.size() != M))
        throw Error(PG_ERR_ARG, "calc3DDensityCompensation: trajectory lengths differ");
    detail::DevBuf d(sizeof(T) * M);
    check(pg_density_compensation(kx.memptr(), ky.memptr(), kz.size() ? kz.memptr() : nullptr, M,
                                  Ninplane, Nz, (float*)d.p, nullptr));
    Col<T> out(M);
    check(pg_memcpy_dtoh(out.memptr(), d.p, sizeof(T) * M, nullptr));
    return out;
}

// directRecon.cpp:63-73: sqrt of the coil sum of magnitudes over
// Nx*Nx*Nz*NSliceMax voxels
template <typename T>
Col<T> calcSumOfSquaresImage(Col<complex<T>> coilImages, uword Nx, uword Nz, uword NSliceMax,
                             uword Ncoils)
{
    detail::require_float<T>();
    const uword nvox = Nx * Nx * Nz * NSliceMax;
    if (coilImages.size() != nvox * Ncoils)
        throw Error(PG_ERR_ARG, "calcSumOfSquaresImage: wrong input length");
    detail::DevBuf x(coilImages.memptr(), sizeof(complex<T>) * nvox * Ncoils);
    detail::DevBuf o(sizeof(T) * nvox);
    check(pg_sum_of_squares(x.p, nvox, Ncoils, (float*)o.p, nullptr));
    Col<T> out(nvox);
    check(pg_memcpy_dtoh(out.memptr(), o.p, sizeof(T) * nvox, nullptr));
    return out;
}

// the coil loop of gridCoilImages (directRecon.cpp:118-127): image c =
// G / (dcf % data_c), coil-major
template <typename T>
Col<complex<T>> gridCoilImages(const Gnufft<T>& G, const Col<T>& dcf, const Col<complex<T>>& data,
                               uword Ncoils)
{
    const uword M = G.n2, N = G.n1;
    if (dcf.size() != M || data.size() != M * Ncoils)
        throw Error(PG_ERR_ARG, "gridCoilImages: wrong input lengths");
    detail::DevBuf w(dcf.memptr(), sizeof(T) * M);
    detail::DevBuf y(data.memptr(), sizeof(complex<T>) * M * Ncoils);
    detail::DevBuf im(sizeof(complex<T>) * N * Ncoils);
    check(pg_grid_coil_images(G.handle(), (const float*)w.p, y.p, Ncoils, im.p, nullptr));
    Col<complex<T>> out(N * Ncoils);
    check(pg_memcpy_dtoh(out.memptr(), im.p, sizeof(complex<T>) * N * Ncoils, nullptr));
    return out;
}

// ------------------------------------------------------------ image output ---
// processNIFTI.hpp:41-101: <filename>.nii, float32
template <typename T1>
void writeNiftiRealImage(std::string filename, Col<T1> imageData, uword Nx, uword Ny, uword Nz = 1,
                         uword Nslab = 1, uword N_TR = 1)
{
    detail::require_float<T1>();
    if (imageData.size() != Nx * Ny * Nz * Nslab * N_TR)
        throw Error(PG_ERR_ARG, "writeNiftiRealImage: image size does not match");
    check(pg_write_nifti_real(filename.c_str(), imageData.memptr(), Nx, Ny, Nz, Nslab, N_TR));
}

// processNIFTI.hpp:104-119: <filename>_mag.nii and <filename>_phs.nii
template <typename T1>
void writeNiftiMagPhsImage(std::string filename, Col<complex<T1>> imageData, uword Nx, uword Ny,
                           uword Nz, uword Nslab = 1, uword N_TR = 1)
{
    detail::require_float<T1>();
    const uword n = Nx * Ny * Nz * Nslab * N_TR;
    if (imageData.size() != n) throw Error(PG_ERR_ARG, "writeNiftiMagPhsImage: size mismatch");
    detail::DevBuf d(imageData.memptr(), sizeof(complex<T1>) * n);
    check(pg_write_nifti_mag_phase(filename.c_str(), d.p, Nx, Ny, Nz, Nslab, N_TR, nullptr));
}

// reconSolve.cpp:30-54: y-fastest image-space coordinates in [-1/2, 1/2)
template <typename T1>
void initImageSpaceCoords(Col<T1>& ix, Col<T1>& iy, Col<T1>& iz, uword Nx, uword Ny, uword Nz)
{
    ix.set_size(Nx * Ny * Nz);
    iy.set_size(Nx * Ny * Nz);
    iz.set_size(Nx * Ny * Nz);
    for (uword kk = 0; kk < Nz; ++kk)
        for (uword jj = 0; jj < Nx; ++jj)
            for (uword ii = 0; ii < Ny; ++ii) {
                const uword idx = ii + jj * Ny + kk * Nx * Ny;
                ix(idx) = ((T1)jj - (T1)Nx / 2.0) / ((T1)Nx);
                iy(idx) = ((T1)ii - (T1)Ny / 2.0) / ((T1)Ny);
                iz(idx) = ((T1)kk - (T1)Nz / 2.0) / ((T1)Nz);
            }
}

}  // namespace powergrid

#endif
