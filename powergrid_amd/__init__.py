"""powergrid_amd -- MI355X (gfx950) implementation of PowerGrid's
field-corrected non-Cartesian SENSE path, behind the reference's operator API.

The classes mirror the reference (PowerGrid/Gnufft.h:51-53, Gdft.h:47-49,
TimeSegmentation.h:70-71, SENSE.h:64, QuadPenalty.h:43, TVPenalty.h:44,
solve_pwls_pcg.hpp:48-51, reconSolve.h:63-66): `A * x` is the forward
operator (operator*) and `A / y` the adjoint (operator/).  Every call goes
through the C ABI of libpowergrid_amd.so (include/powergrid_amd.h); numpy
arrays take the host-buffer entry points, DeviceArrays the
device-buffer ones.  Device memory and streams come from the library
itself (DeviceArray, Stream): no other GPU framework is loaded.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as _L
from ._lib import CgInfo, PgError, check, lib

MODE_BATCHED = 0
MODE_PLAIN = 1
CG_GRAPH = 0
CG_HOST = 1
EXIT_REASONS = {0: "niter", 1: "zero_gradient", 2: "not_descent", 3: "bad_denominator",
                4: "tolerance"}


def _f32(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Stream:
    """A HIP stream owned by the library (pg_stream_create)."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().pg_stream_create(C.byref(h)))
        self.handle = h

    def synchronize(self):
        check(lib().pg_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().pg_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class DeviceArray:
    """A 1-D device buffer allocated through the C ABI (pg_malloc).

    Device memory is owned by the library's own HIP runtime; no other GPU
    framework is involved."""

    def __init__(self, n, dtype=np.complex64):
        self.size = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.size * self.dtype.itemsize
        h = C.c_void_p()
        check(lib().pg_malloc(C.byref(h), max(self.nbytes, 1)))
        self.ptr = h

    @classmethod
    def from_host(cls, a, dtype=None, stream=None):
        a = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel())
        d = cls(a.size, a.dtype)
        if a.nbytes:
            check(lib().pg_memcpy_htod(d.ptr, _p(a), a.nbytes, _sh(stream)))
        return d

    def upload(self, a, stream=None):
        """copy host data of the same size into this buffer"""
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype).ravel())
        assert a.nbytes == self.nbytes
        if a.nbytes:
            check(lib().pg_memcpy_htod(self.ptr, _p(a), a.nbytes, _sh(stream)))
        return self

    def to_host(self, stream=None):
        out = np.empty(self.size, self.dtype)
        if self.nbytes:
            check(lib().pg_memcpy_dtoh(_p(out), self.ptr, self.nbytes, _sh(stream)))
        return out

    def copy_from(self, other, stream=None):
        assert other.nbytes == self.nbytes
        check(lib().pg_memcpy_dtod(self.ptr, other.ptr, self.nbytes, _sh(stream)))
        return self

    def __len__(self):
        return self.size

    def __del__(self):
        try:
            if self.ptr:
                lib().pg_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def _sh(stream):
    return None if stream is None else stream.handle


def _as_device(a, dtype):
    if isinstance(a, DeviceArray):
        if a.dtype != np.dtype(dtype):
            raise ValueError(f"expected a {np.dtype(dtype)} device array")
        return a
    return DeviceArray.from_host(a, dtype)


class _Op:
    """Common operator behaviour (operator* / operator/)."""

    def __init__(self, handle, deps=()):
        self._h = handle
        self._deps = deps  # inner operators must outlive composites
        n_in, n_out = C.c_size_t(), C.c_size_t()
        check(lib().pg_op_shape(handle, C.byref(n_in), C.byref(n_out)))
        self.n_in, self.n_out = n_in.value, n_out.value

    @property
    def handle(self):
        return self._h

    def set_mode(self, mode):
        """MODE_BATCHED (default) or MODE_PLAIN (one NUFFT at a time)."""
        check(lib().pg_op_set_mode(self._h, int(mode)))
        return self

    def forward(self, x, out=None, stream=None):
        return self._apply(x, out, True, stream)

    def adjoint(self, y, out=None, stream=None):
        return self._apply(y, out, False, stream)

    def apply_static(self, x, out, adjoint=False, stream=None):
        """pg_dft_apply_static: the forward (or adjoint) of a Gdft / GdftR2
        operator, or of a SENSE operator over one, with its partial sums in
        the library's static device scratch: no device allocation.  x and out
        are DeviceArrays the caller allocated."""
        n_in, n_out = (self.n_out, self.n_in) if adjoint else (self.n_in, self.n_out)
        for a, n in ((x, n_in), (out, n_out)):
            if not isinstance(a, DeviceArray) or a.size != n or a.dtype != np.complex64:
                raise ValueError(f"expected a complex64 device array of {n} elements")
        check(lib().pg_dft_apply_static(self._h, 1 if adjoint else 0, x.ptr, out.ptr,
                                        _sh(stream)))
        return out

    def _apply(self, x, out, fwd, stream):
        n_in, n_out = (self.n_in, self.n_out) if fwd else (self.n_out, self.n_in)
        if isinstance(x, DeviceArray):
            if x.size != n_in or x.dtype != np.complex64:
                raise ValueError(f"expected a complex64 device array of {n_in} elements")
            if out is None:
                out = DeviceArray(n_out, np.complex64)
            fn = lib().pg_op_forward if fwd else lib().pg_op_adjoint
            check(fn(self._h, x.ptr, out.ptr, _sh(stream)))
            return out
        xin = np.ascontiguousarray(np.asarray(x, dtype=np.complex64).ravel())
        if xin.size != n_in:
            raise ValueError(f"expected {n_in} values, got {xin.size}")
        res = np.empty(n_out, dtype=np.complex64)
        fn = lib().pg_op_forward_host if fwd else lib().pg_op_adjoint_host
        check(fn(self._h, _p(xin), _p(res)))
        return res

    def __mul__(self, x):
        return self.forward(x)

    def __truediv__(self, y):
        return self.adjoint(y)

    def __del__(self):
        try:
            if self._h:
                lib().pg_op_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _unary(fn, h, x, n_in, dt_in, n_out, dt_out, out, stream):
    """fn(h, in_dev, out_dev, stream): DeviceArrays in -> DeviceArray out,
    host arrays in -> numpy out (synchronised)."""
    dev = isinstance(x, DeviceArray)
    xd = _as_device(x, dt_in)
    if xd.size != n_in:
        raise ValueError(f"expected {n_in} values, got {xd.size}")
    od = out if out is not None else DeviceArray(n_out, dt_out)
    if od.size != n_out or od.dtype != np.dtype(dt_out):
        raise ValueError(f"expected a {np.dtype(dt_out)} output of {n_out} elements")
    check(fn(h, xd.ptr, od.ptr, _sh(stream)))
    return od if dev else od.to_host(stream)


def _create(fn, *args):
    h = C.c_void_p()
    check(fn(C.byref(h), *args))
    return h


class Gnufft(_Op):
    """Gnufft<float>(dataLength, gridos, nx, ny, nz, kx, ky, kz, ix, iy, iz)
    (Gnufft.cpp:34-122); kernel_width defaults to the reference's 4.0."""

    def __init__(self, M, gridos, Nx, Ny, Nz, kx, ky, kz, ix=None, iy=None, iz=None,
                 kernel_width=4.0):
        kx, ky = _f32(kx, M), _f32(ky, M)
        kz = _f32(kz, M) if kz is not None else np.zeros(M, np.float32)
        h = _create(lib().pg_gnufft_create, M, float(gridos), Nx, Ny, Nz, _p(kx), _p(ky), _p(kz),
                    None, None, None, float(kernel_width))
        super().__init__(h)
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.kernel_width = kernel_width
        g = [C.c_size_t() for _ in range(3)]
        check(lib().pg_gnufft_grid_size(h, *[C.byref(v) for v in g]))
        self.grid_shape = tuple(v.value for v in g)  # (Gx, Gy, Gz); y fastest in memory
        self.grid_size = int(np.prod(self.grid_shape))

    def forwardSpatialInterp(self, grid, out=None, stream=None):
        """Gnufft::forwardSpatialInterp (Gnufft.cpp:214-263): the oversampled
        grid (Gx*Gy*Gz, y fastest, unshifted) interpolated at the samples."""
        return _unary(lib().pg_gnufft_forward_spatial, self._h, grid, self.grid_size,
                      np.complex64, self.n_out, np.complex64, out, stream)

    def adjointSpatialInterp(self, samples, out=None, stream=None):
        """Gnufft::adjointSpatialInterp (Gnufft.cpp:266-337): the samples
        spread onto the zeroed oversampled grid."""
        return _unary(lib().pg_gnufft_adjoint_spatial, self._h, samples, self.n_out,
                      np.complex64, self.grid_size, np.complex64, out, stream)


class Gdft(_Op):
    """Gdft<float>(a, b, kx, ky, kz, ix, iy, iz, fm, t) (Gdft.cpp:32-49).
    precision 0: FP32 phase and trig (the reference's class); 1: FP64."""

    def __init__(self, M, N, kx, ky, kz, ix, iy, iz, fm, t, precision=0):
        arrs = [_f32(kx, M), _f32(ky, M), _f32(kz, M), _f32(ix, N), _f32(iy, N), _f32(iz, N),
                _f32(fm, N), _f32(t, M)]
        h = _create(lib().pg_gdft_create, M, N, *[_p(a) for a in arrs], int(precision))
        super().__init__(h)


class GdftR2(_Op):
    """GdftR2<float>(a, b, kx, ky, kz, ix, iy, iz, fm, t, numX, numY, numZ)
    (GdftR2.cpp:32-57): Gdft with the per-pair field-gradient sinc factors
    (ftCpuWithGrads.cpp:98-212)."""

    def __init__(self, M, N, kx, ky, kz, ix, iy, iz, fm, t, Nx, Ny, Nz):
        arrs = [_f32(kx, M), _f32(ky, M), _f32(kz, M), _f32(ix, N), _f32(iy, N), _f32(iz, N),
                _f32(fm, N), _f32(t, M)]
        h = _create(lib().pg_gdftr2_create, M, N, *[_p(a) for a in arrs], Nx, Ny, Nz)
        super().__init__(h)


class TimeSegmentation(_Op):
    """TimeSegmentation<float, G>(G, fieldmap, tvec, M, N, L, interptype=1,
    shots=1) (TimeSegmentation.cpp:36-221); interptype 1 Hanning, 2 the exact
    least-squares min-max interpolator, 3 its histogram approximation."""

    def __init__(self, G, fieldmap, tvec, M, N, L, interptype=1, shots=1):
        fm, t = _f32(fieldmap, N), _f32(tvec, M)
        h = _create(lib().pg_ts_create, G.handle, _p(fm), _p(t), M, N, L, interptype, shots)
        super().__init__(h, (G,))
        self.L = max(1, int(L))


class PcSenseTimeSeg(_Op):
    """pcSenseTimeSeg<float>(kx, ky, kz, nx, ny, nz, nc, t, L, intType,
    SENSEmap, FieldMap, ShotPhaseMap) (pcSenseTimeSeg.cpp:40-124): kx, ky, kz
    and t hold the shots' samples shot-major; ShotPhaseMap holds one real
    phase map (rad) per shot; the data is Nd x (Ns*nc), column s + c*Ns."""

    def __init__(self, kx, ky, kz, nx, ny, nz, nc, t, L, interptype, sens, fieldmap,
                 shot_phase, kernel_width=4.0):
        Ni = nx * ny * nz
        ph = np.ascontiguousarray(np.ravel(shot_phase), dtype=np.float32)
        Ns = ph.size // Ni
        assert Ns >= 1 and Ns * Ni == ph.size
        Mt = len(np.ravel(kx))
        assert Mt % Ns == 0
        Nd = Mt // Ns
        arrs = [_f32(kx, Mt), _f32(ky, Mt), None if kz is None else _f32(kz, Mt), _f32(t, Mt)]
        s = np.ascontiguousarray(np.ravel(sens), dtype=np.complex64)
        assert s.size == Ni * nc
        fm = _f32(fieldmap, Ni)
        h = _create(lib().pg_pcsense_create, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), nx, ny, nz,
                    nc, _p(arrs[3]), L, interptype, _p(s), _p(fm), _p(ph), Ns, Nd,
                    float(kernel_width))
        super().__init__(h)
        self.Ns, self.Nd, self.nc = Ns, Nd, nc


class PcSenseRank(PcSenseTimeSeg):
    """One MPI rank's share of mpipcSENSETimeSeg<float>
    (MPI/mpipcSENSETimeSeg.cpp:39-262): the (shot, coil) tasks t = rank,
    rank + nranks, ... of the Ns*nc (shot t // nc, coil t % nc), each with a
    TimeSegmentation built with shots = Ns as in the reference.  Forward: the
    rank's Nd x ntask columns in task order; adjoint: those columns to the
    rank's partial image.  `tasks` holds (shot, coil, full-data column
    shot + coil*Ns) per column; the gather and the cross-rank sum are the
    caller's collectives."""

    def __init__(self, kx, ky, kz, nx, ny, nz, nc, t, L, interptype, sens, fieldmap, shot_phase,
                 rank, nranks, kernel_width=4.0):
        Ni = nx * ny * nz
        ph = np.ascontiguousarray(np.ravel(shot_phase), dtype=np.float32)
        Ns = ph.size // Ni
        assert Ns >= 1 and Ns * Ni == ph.size
        Mt = len(np.ravel(kx))
        assert Mt % Ns == 0
        Nd = Mt // Ns
        arrs = [_f32(kx, Mt), _f32(ky, Mt), None if kz is None else _f32(kz, Mt), _f32(t, Mt)]
        s = np.ascontiguousarray(np.ravel(sens), dtype=np.complex64)
        assert s.size == Ni * nc
        fm = _f32(fieldmap, Ni)
        h = _create(lib().pg_pcsense_rank_create, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), nx, ny,
                    nz, nc, _p(arrs[3]), L, interptype, _p(s), _p(fm), _p(ph), Ns, Nd,
                    float(kernel_width), rank, nranks)
        _Op.__init__(self, h)
        self.Ns, self.Nd, self.nc = Ns, Nd, nc
        n = C.c_size_t()
        check(lib().pg_pcsense_tasks(h, C.byref(n), None, None, None))
        sh, co = np.zeros(n.value, np.uint32), np.zeros(n.value, np.uint32)
        check(lib().pg_pcsense_tasks(h, C.byref(n), _p(sh), _p(co), None))
        self.tasks = [(int(a), int(b), int(a) + int(b) * Ns) for a, b in zip(sh, co)]


class PcSENSE(_Op):
    """pcSENSE<float>(kx, ky, kz, nx, ny, nz, nc, t, SENSEmap, FieldMap,
    ShotPhaseMap) (pcSENSE.cpp:39-102): an exact DFT per shot.  The forward
    applies S_c exp(+i P_s) and the adjoint conj(S_c) exp(+i P_s), as in the
    reference (pcSENSE.cpp:97-98)."""

    def __init__(self, kx, ky, kz, nx, ny, nz, nc, t, sens, fieldmap, shot_phase):
        Ni = nx * ny * nz
        ph = np.ascontiguousarray(np.ravel(shot_phase), dtype=np.float32)
        Ns = ph.size // Ni
        assert Ns >= 1 and Ns * Ni == ph.size
        Mt = len(np.ravel(kx))
        assert Mt % Ns == 0
        Nd = Mt // Ns
        arrs = [_f32(kx, Mt), _f32(ky, Mt), None if kz is None else _f32(kz, Mt), _f32(t, Mt)]
        s = np.ascontiguousarray(np.ravel(sens), dtype=np.complex64)
        assert s.size == Ni * nc
        fm = _f32(fieldmap, Ni)
        h = _create(lib().pg_pcsense_gdft_create, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), nx, ny,
                    nz, nc, _p(arrs[3]), _p(s), _p(fm), _p(ph), Ns, Nd)
        super().__init__(h)
        self.Ns, self.Nd, self.nc = Ns, Nd, nc


class SENSE(_Op):
    """SENSE<float, G>(G, sens, M, N, nc) (SENSE.cpp:32-44); sens holds nc
    coil maps of N values, coil-major."""

    def __init__(self, G, sens, M, N, nc):
        s = np.ascontiguousarray(np.asarray(sens, dtype=np.complex64).ravel())
        if s.size != N * nc:
            raise ValueError("sens must hold N*nc values")
        h = _create(lib().pg_sense_create, G.handle, _p(s), M, N, nc)
        super().__init__(h, (G,))
        self.nc = nc


def op_tuning(op, B):
    """The kernel variants the autotuner chose for the Gnufft plan under op at
    B grids: (interp_lds, spread_mfma), each 1/0, or (-1, -1) when untuned."""
    a, b = C.c_int(), C.c_int()
    check(lib().pg_op_tuning(op.handle, int(B), C.byref(a), C.byref(b)))
    return a.value, b.value


def set_tuning(key, value):
    """pg_set_tuning: "spread_mfma" 1 runs the 2-D batched spread on the
    matrix cores (grid count a multiple of 4), 0 (default) on the vector
    units; "autotune" 1 lets each plan time its interpolation and spread
    variants when first prepared for a grid count and keep the faster;
    "zclass" 1 (default) builds 3-D plans of stacked trajectories for z-class
    planes, 0 for the full 3-D grid (plans created afterwards); "spread_rows"
    1 (default) spreads time-segmented batches from compact sample rows (coil
    values and segment weights), 0 from rows of nc*L values; "trace" 1
    pushes roctx tracing ranges (trace()); "zfuse" 1 (default) runs the
    adjoint y IFFT and z stage of z-class planes as one kernel;
    "spread_stage" 1 (default) stages the compact rows of each group of
    patch entries in LDS, 0 reads them per entry; "zstage_mfma" 1 (default)
    forms the forward z stage's class sums on the matrix cores; "zadj_mfma"
    1 (default) the fused adjoint's, with its y IFFTs on four wavefronts;
    "zfwd_fuse" 1 (default) runs the matrix-core forward z stage and the
    class planes' y FFT as one kernel (bit-identical to the two kernels)."""
    check(lib().pg_set_tuning(key.encode(), int(value)))


def trace(on=True):
    """Tracing ranges (roctx, rocprofv3 --marker-trace) around operator
    applications, their launch stages, solves and set-up: the reference's
    RANGE markers (Tracer.hpp)."""
    set_tuning("trace", 1 if on else 0)


def trace_ranges():
    """(libroctx64 available, ranges pushed so far)"""
    a, n = C.c_int(), C.c_ulonglong()
    check(lib().pg_trace_ranges(C.byref(a), C.byref(n)))
    return bool(a.value), int(n.value)


def zclasses(op):
    """The number of z-class planes of the Gnufft plan under op (0: the plan
    grids on the full 3-D grid)."""
    n = C.c_size_t()
    check(lib().pg_gnufft_zclasses(op.handle, C.byref(n)))
    return n.value


def calc3DDensityCompensation(kx, ky, kz, Ninplane, Nz, stream=None):
    """calc3DDensityCompensation(kx, ky, kz, Ninplane, Nz)
    (directRecon.cpp:28-60): Pipe-Menon density compensation, two
    iterations of spread and interpolation on a 3x oversampled grid,
    normalised to unit 2-norm.  Returns M float32 weights (numpy)."""
    M = len(np.ravel(kx))
    kx, ky = _f32(kx, M), _f32(ky, M)
    kz = _f32(kz, M) if kz is not None else np.zeros(M, np.float32)
    out = DeviceArray(M, np.float32)
    check(lib().pg_density_compensation(_p(kx), _p(ky), _p(kz), M, Ninplane, Nz, out.ptr,
                                        _sh(stream)))
    return out.to_host(stream)


def grid_coil_images(G, dcf, data, ncoils, stream=None):
    """The coil loop of gridCoilImages (directRecon.cpp:118-127): image c =
    G / (dcf % data_c).  data: ncoils x M (coil-major); returns ncoils x N
    (coil-major), a DeviceArray when data is one."""
    dev = isinstance(data, DeviceArray)
    dd = _as_device(dcf, np.float32)
    xd = _as_device(data, np.complex64)
    if dd.size != G.n_out or xd.size != G.n_out * ncoils:
        raise ValueError("dcf must hold M values and data M*ncoils")
    out = DeviceArray(G.n_in * ncoils, np.complex64)
    check(lib().pg_grid_coil_images(G.handle, dd.ptr, xd.ptr, ncoils, out.ptr, _sh(stream)))
    return out if dev else out.to_host(stream)


def calcSumOfSquaresImage(coilImages, Nx, Nz, NSliceMax, Ncoils, stream=None):
    """calcSumOfSquaresImage(coilImages, Nx, Nz, NSliceMax, Ncoils)
    (directRecon.cpp:63-73): sqrt of the sum over coils of the magnitudes
    (the reference sums |x|, not |x|^2) over Nx*Nx*Nz*NSliceMax voxels."""
    nvox = Nx * Nx * Nz * NSliceMax
    dev = isinstance(coilImages, DeviceArray)
    xd = _as_device(coilImages, np.complex64)
    if xd.size != nvox * Ncoils:
        raise ValueError(f"expected {nvox * Ncoils} values, got {xd.size}")
    out = DeviceArray(nvox, np.float32)
    check(lib().pg_sum_of_squares(xd.ptr, nvox, Ncoils, out.ptr, _sh(stream)))
    return out if dev else out.to_host(stream)


class _Penalty:
    def __init__(self, kind, Nx, Ny, Nz, beta, delta, dims2penalize):
        self._h = _create(lib().pg_penalty_create, kind, Nx, Ny, Nz, float(beta), float(delta),
                          dims2penalize)
        self.N = Nx * Ny * Nz

    @property
    def handle(self):
        return self._h

    def Gradient(self, x, stream=None):
        xd = _as_device(x, np.complex64)
        g = DeviceArray(self.N, np.complex64)
        check(lib().pg_penalty_gradient(self._h, xd.ptr, g.ptr, _sh(stream)))
        return g if isinstance(x, DeviceArray) else g.to_host()

    def Denom(self, ddir, x, stream=None):
        out = (C.c_double * 2)()
        # keep the device copies referenced until the call has returned
        dd, xd = _as_device(ddir, np.complex64), _as_device(x, np.complex64)
        check(lib().pg_penalty_denom(self._h, dd.ptr, xd.ptr, out, _sh(stream)))
        return complex(out[0], out[1])

    def Penalty(self, x, stream=None):
        out = C.c_double()
        xd = _as_device(x, np.complex64)
        check(lib().pg_penalty_value(self._h, xd.ptr, C.byref(out), _sh(stream)))
        return out.value

    def __del__(self):
        try:
            if self._h:
                lib().pg_penalty_destroy(self._h)
                self._h = None
        except Exception:
            pass


class QuadPenalty(_Penalty):
    """QuadPenalty<float>(nx, ny, nz, beta, dims2penalize=3) (QuadPenalty.cpp:32-42)."""

    def __init__(self, Nx, Ny, Nz, beta, dims2penalize=3):
        super().__init__(0, Nx, Ny, Nz, beta, 0.0, dims2penalize)


class TVPenalty(_Penalty):
    """TVPenalty<float>(nx, ny, nz, beta, delta, dims2penalize=3) (TVPenalty.cpp:32-41)."""

    def __init__(self, Nx, Ny, Nz, beta, delta, dims2penalize=3):
        super().__init__(1, Nx, Ny, Nz, beta, delta, dims2penalize)


def solve_pwls_pcg(x0, A, W, yi, R, niter, mode=CG_GRAPH, stream=None, tol=0.0):
    """solve_pwls_pcg(xInitial, A, W, yi, R, niter) (solve_pwls_pcg.hpp:48-167),
    run on the device.  Arguments are numpy arrays or DeviceArrays; the
    result has the type of yi.  Returns (x, info dict).  tol > 0 (no
    counterpart in the reference) also stops before the update of the first
    iteration k >= 1 whose gradient norm is <= tol times the first one's
    (exit_reason "tolerance")."""
    y = _as_device(yi, np.complex64)
    x0d = None if x0 is None else _as_device(x0, np.complex64)
    Wd = None if W is None else _as_device(W, np.float32)
    out = DeviceArray(A.n_in, np.complex64)
    resid = (C.c_double * max(int(niter), 1))()
    info = CgInfo()
    check(lib().pg_solve_pwls_pcg_tol(A.handle, R.handle if R is not None else None,
                                      None if x0d is None else x0d.ptr,
                                      None if Wd is None else Wd.ptr, y.ptr, int(niter),
                                      float(tol), int(mode), out.ptr, resid, C.byref(info),
                                      _sh(stream)))
    res = dict(iterations=info.iterations, exit_reason=EXIT_REASONS.get(info.exit_reason, "?"),
               resid=np.array(resid[:info.iterations]))
    if not isinstance(yi, DeviceArray):
        out = out.to_host()
    return out, res


def reconSolve(data, Sg, R, kx=None, ky=None, kz=None, Nx=None, Ny=None, Nz=None, tvec=None,
               niter=10, mode=CG_GRAPH, stream=None, tol=0.0):
    """reconSolve(data, Sg, R, kx, ky, kz, Nx, Ny, Nz, tvec, niter)
    (reconSolve.h:63-82): W = ones, x0 = 0; tol as solve_pwls_pcg."""
    return solve_pwls_pcg(None, Sg, None, data, R, niter, mode, stream, tol)


def recon_solve_batch(items, niter, tol=0.0, stream=None):
    """K independent reconstructions (reconSolve each) as one graph launch
    (pg_recon_solve_batch): items is a list of (A, y, R) with R possibly None
    and y a numpy array or DeviceArray; the operators must not share inner
    plans.  Returns a list of (x, info dict), x of y's type."""
    K = len(items)
    own = stream is None
    st = Stream() if own else stream
    ys, xs, keep = [], [], []
    for A, y, R in items:
        yd = _as_device(y, np.complex64)
        keep.append(yd)
        ys.append(yd)
        xs.append(DeviceArray(A.n_in, np.complex64))
    Ah = (C.c_void_p * K)(*[A.handle for A, _, _ in items])
    Rh = (C.c_void_p * K)(*[(R.handle if R is not None else None) for _, _, R in items])
    yh = (C.c_void_p * K)(*[d.ptr for d in ys])
    xh = (C.c_void_p * K)(*[d.ptr for d in xs])
    infos = (CgInfo * K)()
    check(lib().pg_recon_solve_batch(K, Ah, Rh, yh, int(niter), float(tol), xh, infos,
                                     _sh(st)))
    out = []
    for k, (A, y, R) in enumerate(items):
        res = dict(iterations=infos[k].iterations,
                   exit_reason=EXIT_REASONS.get(infos[k].exit_reason, "?"))
        x = xs[k] if isinstance(y, DeviceArray) else xs[k].to_host(st)
        out.append((x, res))
    st.synchronize()
    return out


class profile:
    """Per-kernel timing from HIP events recorded by the library."""

    @staticmethod
    def enable(on=True):
        check(lib().pg_profile_enable(1 if on else 0))

    @staticmethod
    def reset():
        check(lib().pg_profile_reset())

    @staticmethod
    def query(name):
        n, ms = C.c_int(), C.c_double()
        check(lib().pg_profile_query(name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value


__all__ = ["DeviceArray", "Stream", "Gnufft", "Gdft", "GdftR2", "TimeSegmentation", "SENSE",
           "PcSenseTimeSeg", "PcSenseRank", "PcSENSE", "QuadPenalty", "TVPenalty", "calc3DDensityCompensation",
           "grid_coil_images", "calcSumOfSquaresImage", "set_tuning", "op_tuning", "trace",
           "trace_ranges", "zclasses",
           "solve_pwls_pcg", "recon_solve_batch",
           "reconSolve", "profile", "PgError", "MODE_BATCHED", "MODE_PLAIN", "CG_GRAPH",
           "CG_HOST"]
