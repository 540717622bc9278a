"""ctypes wrapper of the CPU restatement in oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker.  The product package
(powergrid_amd) never imports it.

Arrays are numpy; complex data is complex64 (float variant) or complex128
(double variant).  The classes mirror the reference operators
(Gnufft.cpp, TimeSegmentation.cpp, SENSE.cpp, Gdft.cpp) with `*` = forward
and `/` = adjoint, like the reference's operator* / operator/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_vp = C.c_void_p


def _declare(L):
    for suf in ("f", "d"):
        R = C.c_float if suf == "f" else C.c_double
        def f(name, res, *args):
            fn = getattr(L, name + "_" + suf)
            fn.restype = res
            fn.argtypes = list(args)
        f("orc_lut_size", C.c_size_t, R)
        f("orc_lut", None, R, R, _vp)
        f("orc_kb_beta", R, R, R)
        f("orc_fft", None, _vp, C.c_int, C.c_int, C.c_int, C.c_int)
        f("orc_fft1d", None, _vp, C.c_int, C.c_int)
        f("orc_deapod", None, _vp, _vp, C.c_int, C.c_int, C.c_int, R, R, R)
        f("orc_zero_pad", None, _vp, _vp, C.c_int, C.c_int, C.c_int, R)
        f("orc_crop", None, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int)
        f("orc_fftshift", None, _vp, _vp, C.c_int, C.c_int, C.c_int)
        f("orc_ifftshift", None, _vp, _vp, C.c_int, C.c_int, C.c_int)
        f("orc_grid_forward", None, C.c_int, C.c_int, C.c_int, C.c_int, R, R,
          _vp, _vp, _vp, _vp, C.c_size_t, _vp, _vp)
        f("orc_grid_adjoint", None, C.c_int, C.c_int, C.c_int, C.c_int, R, R,
          _vp, _vp, _vp, _vp, C.c_size_t, _vp, _vp)
        f("orc_ft", None, *([_vp] * 12), C.c_int, C.c_int)
        f("orc_ift", None, *([_vp] * 12), C.c_int, C.c_int)
        f("orc_gnufft_create", _vp, C.c_int, R, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, R)
        f("orc_gdft_create", _vp, C.c_int, C.c_int, *([_vp] * 8))
        f("orc_ts_create", _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int)
        f("orc_sense_create", _vp, _vp, _vp, C.c_int, C.c_int, C.c_int)
        f("orc_dense_create", _vp, _vp, C.c_int, C.c_int)
        f("orc_pcsense_create", _vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
          C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, _vp, R)
        f("orc_gdftr2_create", _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
          C.c_int, C.c_int, C.c_int)
        f("orc_gdftr2_grad", None, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp)
        f("orc_pcsense_rank_create", _vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int,
          C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, _vp, R, C.c_int, C.c_int)
        f("orc_pcsense_rank_tasks", C.c_int, _vp, _vp, _vp)
        f("orc_pcsense_dft_create", _vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int,
          C.c_int, C.c_int, _vp, _vp, _vp, _vp)
        f("orc_op_forward", None, _vp, _vp, _vp)
        f("orc_op_adjoint", None, _vp, _vp, _vp)
        f("orc_op_n_in", C.c_int, _vp)
        f("orc_op_n_out", C.c_int, _vp)
        f("orc_op_destroy", None, _vp)
        f("orc_ts_tables", C.c_int, _vp, _vp, _vp, _vp)
        f("orc_ts_set_aa_complex", C.c_int, _vp, _vp)
        f("orc_penalty_create", _vp, C.c_int, C.c_int, C.c_int, C.c_int, R, R, C.c_int)
        f("orc_penalty_destroy", None, _vp)
        f("orc_penalty_cd", None, _vp, _vp, C.c_int, _vp)
        f("orc_penalty_ctd", None, _vp, _vp, C.c_int, _vp)
        f("orc_penalty_gradient", None, _vp, _vp, _vp)
        f("orc_penalty_denom", None, _vp, _vp, _vp, _vp)
        f("orc_penalty_value", R, _vp, _vp)
        f("orc_pcg", C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp)
    for name in ("orc_nudft_exact_fwd", "orc_nudft_exact_adj"):
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = [_vp] * 10 + [C.c_int, C.c_int]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = None
    L.orc_get_threads.restype = C.c_int


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def _suf(dtype):
    return "f" if np.dtype(dtype) in (np.float32, np.complex64) else "d"


def _rt(suf):
    return np.float32 if suf == "f" else np.float64


def _ct(suf):
    return np.complex64 if suf == "f" else np.complex128


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _fn(name, suf):
    return getattr(lib(), name + "_" + suf)


def _r(a, suf):
    return np.ascontiguousarray(a, dtype=_rt(suf))


def _c(a, suf):
    return np.ascontiguousarray(a, dtype=_ct(suf))


# ----------------------------------------------------------- primitives --

def kb_beta(gridos=2.0, width=4.0, prec="f"):
    return float(_fn("orc_kb_beta", prec)(gridos, width))


def lut(beta, width, prec="f"):
    n = int(_fn("orc_lut_size", prec)(width))
    out = np.zeros(n, dtype=_rt(prec))
    _fn("orc_lut", prec)(beta, width, _p(out))
    return out


def fft(data, nx, ny=1, nz=1, sign=-1):
    """In-place-style unnormalised C2C over a (nx fastest, ny, nz) buffer."""
    suf = _suf(data.dtype)
    a = _c(data, suf).copy()
    _fn("orc_fft", suf)(_p(a), nx, ny, nz, sign)
    return a


def deapod(img, Nx, Ny, Nz, W, beta, os_=2.0):
    suf = _suf(img.dtype)
    src = _c(img, suf)
    dst = np.zeros_like(src)
    _fn("orc_deapod", suf)(_p(dst), _p(src), Nx, Ny, Nz, W, beta, os_)
    return dst


def zero_pad(img, Nx, Ny, Nz, os_=2.0):
    suf = _suf(img.dtype)
    src = _c(img, suf)
    G = int(round(Nx * os_ * Ny * os_ * (Nz * os_ if Nz > 1 else 1)))
    dst = np.zeros(G, dtype=_ct(suf))
    _fn("orc_zero_pad", suf)(_p(dst), _p(src), Nx, Ny, Nz, os_)
    return dst


def crop(grid, Nx, Ny, Nz, Gx, Gy, Gz):
    suf = _suf(grid.dtype)
    src = _c(grid, suf)
    dst = np.zeros(Nx * Ny * Nz, dtype=_ct(suf))
    _fn("orc_crop", suf)(_p(dst), _p(src), Nx, Ny, Nz, Gx, Gy, Gz)
    return dst


def fftshift(grid, Gx, Gy, Gz=1, inverse=False):
    suf = _suf(grid.dtype)
    src = _c(grid, suf)
    dst = np.zeros_like(src)
    _fn("orc_ifftshift" if inverse else "orc_fftshift", suf)(_p(dst), _p(src), Gx, Gy, Gz)
    return dst


def grid_forward(grid, kx, ky, kz, Nx, Ny, Nz, os_=2.0, W=4.0, prec=None):
    suf = prec or _suf(grid.dtype)
    beta = kb_beta(os_, W, suf)
    L = lut(beta, W, suf)
    M = len(kx)
    out = np.zeros(M, dtype=_ct(suf))
    g = _c(grid, suf)
    _fn("orc_grid_forward", suf)(M, Nx, Ny, Nz, os_, W, _p(_r(kx, suf)), _p(_r(ky, suf)),
                                 _p(_r(kz, suf)), _p(L), len(L), _p(g), _p(out))
    return out


def grid_adjoint(samples, kx, ky, kz, Nx, Ny, Nz, Gsize, os_=2.0, W=4.0, prec=None):
    suf = prec or _suf(samples.dtype)
    beta = kb_beta(os_, W, suf)
    L = lut(beta, W, suf)
    M = len(kx)
    out = np.zeros(Gsize, dtype=_ct(suf))
    s = _c(samples, suf)
    _fn("orc_grid_adjoint", suf)(M, Nx, Ny, Nz, os_, W, _p(_r(kx, suf)), _p(_r(ky, suf)),
                                 _p(_r(kz, suf)), _p(L), len(L), _p(s), _p(out))
    return out


def nudft_exact(x, kx, ky, kz, ix, iy, iz, fm, t, adjoint=False):
    """FP64 exact field-corrected NUDFT (double sin/cos)."""
    d = np.float64
    args = [np.ascontiguousarray(a, dtype=d) for a in (kx, ky, kz, ix, iy, iz, fm, t)]
    xin = np.ascontiguousarray(x, dtype=np.complex128)
    M, N = len(kx), len(ix)
    out = np.zeros(N if adjoint else M, dtype=np.complex128)
    fn = lib().orc_nudft_exact_adj if adjoint else lib().orc_nudft_exact_fwd
    fn(_p(out), _p(xin), *[_p(a) for a in args], M, N)
    return out


# ------------------------------------------------------------ operators --

class _Op:
    def __init__(self, handle, suf, keep=()):
        self.h = handle
        self.suf = suf
        self._keep = keep
        self.n_in = _fn("orc_op_n_in", suf)(handle)
        self.n_out = _fn("orc_op_n_out", suf)(handle)

    def __mul__(self, x):
        x = _c(x, self.suf)
        assert x.size == self.n_in
        out = np.zeros(self.n_out, dtype=_ct(self.suf))
        _fn("orc_op_forward", self.suf)(self.h, _p(x), _p(out))
        return out

    def __truediv__(self, y):
        y = _c(y, self.suf)
        assert y.size == self.n_out
        out = np.zeros(self.n_in, dtype=_ct(self.suf))
        _fn("orc_op_adjoint", self.suf)(self.h, _p(y), _p(out))
        return out

    def __del__(self):
        try:
            if self.h and not self._keep:
                _fn("orc_op_destroy", self.suf)(self.h)
        except Exception:
            pass


class Gnufft(_Op):
    """Gnufft.cpp:34-122 (kernelWidth defaults to the reference's 4.0)."""

    def __init__(self, M, gridos, Nx, Ny, Nz, kx, ky, kz, kernel_width=4.0, prec="f"):
        h = _fn("orc_gnufft_create", prec)(M, gridos, Nx, Ny, Nz, _p(_r(kx, prec)),
                                           _p(_r(ky, prec)), _p(_r(kz, prec)), kernel_width)
        super().__init__(h, prec)
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz


class Gdft(_Op):
    """Gdft.cpp:32-49."""

    def __init__(self, M, N, kx, ky, kz, ix, iy, iz, fm, t, prec="f"):
        arrs = [_r(a, prec) for a in (kx, ky, kz, ix, iy, iz, fm, t)]
        h = _fn("orc_gdft_create", prec)(M, N, *[_p(a) for a in arrs])
        super().__init__(h, prec)


class GdftR2(_Op):
    """GdftR2.cpp:32-207: the field-corrected DFT with the per-pair factor
    sinc(kx/Nx + Gx t) sinc(ky/Ny + Gy t) sinc(kz/Nz + Gz t), G = Cd(FM, dim)."""

    def __init__(self, M, N, kx, ky, kz, ix, iy, iz, fm, t, nx, ny, nz, prec="f"):
        args = [_p(_r(a, prec)) for a in (kx, ky, kz, ix, iy, iz, fm, t)]
        h = _fn("orc_gdftr2_create", prec)(M, N, *args, nx, ny, nz)
        super().__init__(h, prec)


def gdftr2_grad(fm, nx, ny, nz, dim, prec="f"):
    """GdftR2::Cd (GdftR2.cpp:67-107) of the field map along dim."""
    out = np.zeros(nx * ny * nz, dtype=_rt(prec))
    _fn("orc_gdftr2_grad", prec)(_p(_r(fm, prec)), nx, ny, nz, dim, _p(out))
    return out


def arma_pinv_rcond(rows, cols, prec):
    """pinv's default tolerance in Armadillo 9.200 (not vendored):
    max(rows, cols) * max singular value * eps of the element type, i.e. a
    relative cutoff of max(rows, cols) * eps."""
    eps = np.finfo(np.float32 if prec == "f" else np.float64).eps
    return max(rows, cols) * eps


def minmax_ls_weights(fm, tv, tau, L, shots, prec="f"):
    """TimeSegmentation.cpp:94-150, the exact least-squares min-max
    interpolator, in the element precision of `prec`:
      g = exp(i fm tau); G = [1, g, ..., g^(L-1)] (N x L);
      GTG = N on the diagonal, glsum(k) = sum_j conj(g_j^(k+1)) on the k+1-th
      sub-diagonal and its conjugate above (G^H G, since |g| = 1);
      iGTGGT = inv(GTG) G^H, or pinv(GTG) G^H when 1/cond(GTG) <= 2e-15;
      AA[i, :] = conj(iGTGGT exp(i fm tv_i)) for the first shot's samples,
      repeated over the shots.  tv: sample times minus their minimum."""
    rt, ct = _rt(prec), _ct(prec)
    fm = np.asarray(fm, dtype=rt).ravel()
    tv = np.asarray(tv, dtype=rt).ravel()
    n2 = fm.size
    gg = np.exp(ct(1j) * (fm * rt(tau)).astype(rt)).astype(ct)
    gl = np.stack([gg ** (k + 1) for k in range(L)], axis=1).astype(ct)  # n2 x L
    G = np.concatenate([np.ones((n2, 1), ct), gl[:, :L - 1]], axis=1)
    glsum = np.sum(gl.conj().T, axis=1).astype(ct)
    GTG = np.zeros((L, L), ct)
    GTG[np.diag_indices(L)] += rt(n2)
    for k in range(L - 1):
        d = k + 1
        for r in range(d, L):
            GTG[r, r - d] += glsum[k]
            GTG[r - d, r] += np.conj(glsum[k])
    rcn = 1.0 / np.linalg.cond(GTG)
    if rcn > 10 * 2e-16:
        inv = np.linalg.inv(GTG)
    else:
        inv = np.linalg.pinv(GTG, rcond=arma_pinv_rcond(L, L, prec))
    iGTGGT = (inv.astype(ct) @ G.conj().T).astype(ct)  # L x n2
    none = tv.size // shots
    F = np.exp(ct(1j) * np.outer(fm, tv[:none]).astype(rt)).astype(ct)  # n2 x none
    R = (iGTGGT @ F).astype(ct)  # L x none
    AA = np.tile(R.conj().T, (shots, 1))  # M x L
    return AA, float(rcn)


def arma_hist(v, nbins, prec="f"):
    """Armadillo's hist(X, n_bins) (op_hist, Armadillo 9.x, not vendored):
    n_bins centres min + (k + 1/2)/n_bins * (max - min), formed in the
    element type; each value is counted at its nearest centre, the lower
    one on a tie."""
    rt = _rt(prec)
    v = np.asarray(v, dtype=rt).ravel()
    mn, mx = v.min(), v.max()
    c = ((0.5 + np.arange(nbins, dtype=np.float64)) / nbins).astype(rt)
    c = (rt(mx - mn) * c + mn).astype(rt)
    # first minimum of |c - v| over the (sorted) centres
    idx = np.argmin(np.abs(c[None, :] - v[:, None]), axis=1)
    return np.bincount(idx, minlength=nbins)[:nbins]


def histogram_ls_weights(fm, tv, tau, L, shots, prec="f"):
    """TimeSegmentation.cpp:157-207 (interptype 3), the approximate min-max
    interpolator, in the element precision of `prec`:
      h = hist(fm, 256); KK = floor(2 pi / ((range/256) tau));
      dwn = 2 pi / (KK tau);
      ftwe = fft_KK(h exp(i dwn n tau L)) * exp(-i (min + dwn/2) tau (k - L));
      GTGap: ftwe(d + L) on the d-th diagonal, d = -L..L, (L+1) x (L+1);
      iGTGap = inv(GTGap.st()), or pinv when rcond <= 10 * double eps;
      per sample: ftc = fft_KK(h exp(i n dwn t)) * exp(i (min + dwn/2)(t - tau k)),
      row = (iGTGap ftc[0..L])^H, of which the L-column row keeps the first
      L entries (an (L+1)-vector assigned to an L-column row under
      ARMA_NO_DEBUG).  fft_KK zero-pads or trims to KK (calcFFT1D,
      TimeSegmentation.cpp:224-245), forward sign.  tv: sample times minus
      their minimum.  Returns (AA M x L, rcond, KK)."""
    rt, ct = _rt(prec), _ct(prec)
    fm = np.asarray(fm, dtype=rt).ravel()
    tv = np.asarray(tv, dtype=rt).ravel()
    nbins = 256
    h = arma_hist(fm, nbins, prec).astype(rt)
    mn, mx = fm.min(), fm.max()
    rng = rt(mx - mn)
    KK = int(np.floor(2.0 * np.pi / float(rt(rng / rt(nbins)) * rt(tau))))
    dwn = rt(2 * np.pi / float(rt(KK * rt(tau))))
    n = np.arange(nbins, dtype=rt)
    temp = (h * np.exp(ct(1j) * (((dwn * n).astype(rt) * rt(tau)).astype(rt) * rt(L)))).astype(ct)
    ftwe = np.fft.fft(temp, n=KK).astype(ct)
    k = np.arange(KK, dtype=rt)
    ftwe = (np.exp(ct(-1j) * (rt(mn + dwn / rt(2)) * rt(tau) * (k - rt(L))).astype(rt)) * ftwe)
    ftwe = ftwe.astype(ct)
    GTGap = np.zeros((L + 1, L + 1), ct)
    for d in range(-L, L + 1):
        for r in range(L + 1):
            c = r + d
            if 0 <= c <= L:
                GTGap[r, c] = ftwe[d + L]
    Mst = np.ascontiguousarray(GTGap.T)
    rcond = 1.0 / float(np.real(np.linalg.cond(Mst, 1)))
    if rcond > 10 * np.finfo(np.float64).eps:
        iG = np.linalg.inv(Mst).astype(ct)
    else:
        iG = np.linalg.pinv(Mst, rcond=arma_pinv_rcond(L + 1, L + 1, prec)).astype(ct)
    none = tv.size // shots
    t1 = tv[:none]
    temp = (h[None, :] * np.exp(ct(1j) * (n[None, :] * dwn * t1[:, None]).astype(rt))).astype(ct)
    ftc = np.fft.fft(temp, n=KK, axis=1).astype(ct)
    ph = (rt(mn + dwn / rt(2)) * (t1[:, None] - rt(tau) * k[None, :])).astype(rt)
    ftc = (np.exp(ct(1j) * ph) * ftc).astype(ct)
    res = (ftc[:, :L + 1] @ iG.T).astype(ct)  # none x (L+1): rows (iG GTc)^T
    AA = np.tile(res.conj()[:, :L], (shots, 1))
    return AA, rcond, KK


class TimeSegmentation(_Op):
    """TimeSegmentation.cpp:36-221: interptype 1 Hanning (in C), 2 the exact
    least-squares min-max interpolator and 3 its histogram approximation
    (complex weights computed here with numpy in the element precision and
    handed to the C operator)."""

    def __init__(self, G, fm, t, M, N, L, interptype=1, shots=1):
        prec = G.suf
        h = _fn("orc_ts_create", prec)(G.h, _p(_r(fm, prec)), _p(_r(t, prec)), M, N, L,
                                       interptype, shots)
        super().__init__(h, prec)
        self.G = G
        # the composite does not own the inner op (TimeSegmentation.h:55);
        # the C side frees only its own context.
        self.rcn = None
        Lc = max(1, int(L)) if int(L) > 1 else 1
        if interptype in (2, 3) and Lc > 1:
            _, _, tau, tmin = self.tables()
            rt = _rt(prec)
            tv = (np.asarray(t, dtype=rt) - rt(tmin)).astype(rt)
            if interptype == 2:
                AA, self.rcn = minmax_ls_weights(fm, tv, tau, Lc, shots, prec)
            else:
                AA, self.rcn, self.KK = histogram_ls_weights(fm, tv, tau, Lc, shots, prec)
            AAc = np.ascontiguousarray(AA.T).astype(_ct(prec))  # L x M = column-major M x L
            _fn("orc_ts_set_aa_complex", prec)(self.h, _p(AAc))

    def tables(self):
        L = _fn("orc_ts_tables", self.suf)(self.h, None, None, None)
        AA = np.zeros((L, self.n_out), dtype=_rt(self.suf))
        Wo = np.zeros((L, self.n_in), dtype=_ct(self.suf))
        tt = np.zeros(2, dtype=_rt(self.suf))
        _fn("orc_ts_tables", self.suf)(self.h, _p(AA), _p(Wo), _p(tt))
        return AA, Wo, float(tt[0]), float(tt[1])


class SENSE(_Op):
    """SENSE.cpp:32-44."""

    def __init__(self, G, sens, M, N, nc):
        prec = G.suf
        h = _fn("orc_sense_create", prec)(G.h, _p(_c(sens, prec)), M, N, nc)
        super().__init__(h, prec)
        self.G = G


class PcSenseTimeSeg(_Op):
    """pcSenseTimeSeg.cpp:40-171: shot-phase-corrected, field-corrected SENSE.
    kx, ky, kz, t hold Ns shots of Nd samples, shot-major; sens is Nc maps
    (coil-major); shot_phase is Ns maps of Ni real phases (rad); the data is
    Nd x (Ns*Nc), column s + c*Ns."""

    def __init__(self, kx, ky, kz, Nx, Ny, Nz, nc, t, L, interptype, sens, fieldmap, shot_phase,
                 kernel_width=4.0, prec="f"):
        Ni = Nx * Ny * Nz
        Ns = len(np.ravel(shot_phase)) // Ni
        Nd = len(np.ravel(kx)) // Ns
        h = _fn("orc_pcsense_create", prec)(
            Nd, Ns, _p(_r(kx, prec)), _p(_r(ky, prec)), _p(_r(kz, prec)), Nx, Ny, Nz, nc,
            _p(_r(t, prec)), L, interptype, _p(_c(sens, prec)), _p(_r(fieldmap, prec)),
            _p(_r(shot_phase, prec)), kernel_width)
        super().__init__(h, prec)
        self.Nd, self.Ns, self.Nc = Nd, Ns, nc


class PcSenseRank(_Op):
    """MPI/mpipcSENSETimeSeg.cpp:39-262, the share of rank `rank` of `nranks`:
    tasks t = rank, rank + nranks, ... of the Ns*Nc (shot t // Nc, coil
    t % Nc), a TimeSegmentation with shots = Ns per task.  Forward: the
    rank's Nd x ntask columns in task order; adjoint: those columns to the
    rank's partial image.  `tasks` lists (shot, coil) per column."""

    def __init__(self, kx, ky, kz, Nx, Ny, Nz, nc, t, L, interptype, sens, fieldmap, shot_phase,
                 rank, nranks, kernel_width=4.0, prec="f"):
        Ni = Nx * Ny * Nz
        Ns = len(np.ravel(shot_phase)) // Ni
        Nd = len(np.ravel(kx)) // Ns
        h = _fn("orc_pcsense_rank_create", prec)(
            Nd, Ns, _p(_r(kx, prec)), _p(_r(ky, prec)), _p(_r(kz, prec)), Nx, Ny, Nz, nc,
            _p(_r(t, prec)), L, interptype, _p(_c(sens, prec)), _p(_r(fieldmap, prec)),
            _p(_r(shot_phase, prec)), kernel_width, rank, nranks)
        super().__init__(h, prec)
        n = _fn("orc_pcsense_rank_tasks", prec)(h, None, None)
        sh = np.zeros(max(n, 1), np.int32)
        co = np.zeros(max(n, 1), np.int32)
        _fn("orc_pcsense_rank_tasks", prec)(h, _p(sh), _p(co))
        self.tasks = list(zip(sh[:n].tolist(), co[:n].tolist()))
        self.Nd, self.Ns, self.Nc = Nd, Ns, nc


class PcSENSE(_Op):
    """pcSENSE.cpp:39-153: shot-phase-corrected SENSE over an exact DFT per
    shot (Gdft).  As in the reference, the forward applies S_c exp(+i P_s)
    and the adjoint conj(S_c) exp(+i P_s)."""

    def __init__(self, kx, ky, kz, Nx, Ny, Nz, nc, t, sens, fieldmap, shot_phase, prec="f"):
        Ni = Nx * Ny * Nz
        Ns = len(np.ravel(shot_phase)) // Ni
        Nd = len(np.ravel(kx)) // Ns
        h = _fn("orc_pcsense_dft_create", prec)(
            Nd, Ns, _p(_r(kx, prec)), _p(_r(ky, prec)), _p(_r(kz, prec)), Nx, Ny, Nz, nc,
            _p(_r(t, prec)), _p(_c(sens, prec)), _p(_r(fieldmap, prec)), _p(_r(shot_phase, prec)))
        super().__init__(h, prec)
        self.Nd, self.Ns, self.Nc = Nd, Ns, nc


class Dense(_Op):
    """A dense complex matrix as an operator (test problems only)."""

    def __init__(self, A, prec="d"):
        A = np.asfortranarray(A, dtype=_ct(prec))
        h = _fn("orc_dense_create", prec)(_p(A.ravel(order="F")), A.shape[0], A.shape[1])
        super().__init__(h, prec)


class Penalty:
    """Robject.cpp / QuadPenalty.cpp (kind 0) / TVPenalty.cpp (kind 1)."""

    def __init__(self, kind, Nx, Ny, Nz, beta, delta=0.0, dims2penalize=3, prec="f"):
        self.suf = prec
        self.N = Nx * Ny * Nz
        self.h = _fn("orc_penalty_create", prec)(kind, Nx, Ny, Nz, beta, delta, dims2penalize)

    def Cd(self, d, dim):
        out = np.zeros(self.N, dtype=_ct(self.suf))
        _fn("orc_penalty_cd", self.suf)(self.h, _p(_c(d, self.suf)), dim, _p(out))
        return out

    def Ctd(self, d, dim):
        out = np.zeros(self.N, dtype=_ct(self.suf))
        _fn("orc_penalty_ctd", self.suf)(self.h, _p(_c(d, self.suf)), dim, _p(out))
        return out

    def Gradient(self, x):
        out = np.zeros(self.N, dtype=_ct(self.suf))
        _fn("orc_penalty_gradient", self.suf)(self.h, _p(_c(x, self.suf)), _p(out))
        return out

    def Denom(self, ddir, x):
        out = np.zeros(2, dtype=_rt(self.suf))
        _fn("orc_penalty_denom", self.suf)(self.h, _p(_c(ddir, self.suf)),
                                           _p(_c(x, self.suf)), _p(out))
        return complex(out[0], out[1])

    def Penalty(self, x):
        return float(_fn("orc_penalty_value", self.suf)(self.h, _p(_c(x, self.suf))))

    def __del__(self):
        try:
            _fn("orc_penalty_destroy", self.suf)(self.h)
        except Exception:
            pass


def QuadPenalty(Nx, Ny, Nz, beta, dims2penalize=3, prec="f"):
    return Penalty(0, Nx, Ny, Nz, beta, 0.0, dims2penalize, prec)


def TVPenalty(Nx, Ny, Nz, beta, delta, dims2penalize=3, prec="f"):
    return Penalty(1, Nx, Ny, Nz, beta, delta, dims2penalize, prec)


def solve_pwls_pcg(x0, A, W, yi, R, niter):
    """solve_pwls_pcg.hpp:48-167.  Returns (x, completed_iterations, resid)."""
    suf = A.suf
    x0 = _c(x0, suf)
    y = _c(yi, suf)
    Wr = None if W is None else _r(W, suf)
    out = np.zeros(A.n_in, dtype=_ct(suf))
    res = np.zeros(max(niter, 1), dtype=_rt(suf))
    n = _fn("orc_pcg", suf)(A.h, R.h if R is not None else None, _p(x0), _p(Wr), _p(y),
                            niter, _p(out), _p(res))
    return out, int(n), res[:n]


def reconSolve(data, Sg, R, Nx, Ny, Nz, niter):
    """reconSolve.h:63-82: W = ones, x0 = 0."""
    x0 = np.zeros(Nx * Ny * Nz, dtype=_ct(Sg.suf))
    return solve_pwls_pcg(x0, Sg, None, data, R, niter)


# ------------------------------------------------ direct reconstruction --

def grid_shape(Nx, Ny, Nz, os_):
    """Oversampled grid (Gx, Gy, Gz) of Gnufft.cpp:228-234."""
    import math
    Gx = math.ceil(os_ * Nx)
    Gx += Gx % 2
    Gy = math.ceil(os_ * Ny)
    Gy += Gy % 2
    Gz = 1 if Nz == 1 else math.ceil(os_ * Nz)
    return Gx, Gy, Gz


def density_compensation(kx, ky, kz, Ninplane, Nz, prec="f"):
    """calc3DDensityCompensation (directRecon.cpp:28-60): Gnufft on a 3x
    oversampled grid (:31-35) with the reference Gnufft's kernel width 4.0
    (Gnufft.cpp:84); dcf = ones / (M*M) (:38-39); twice
    dcf = dcf / real(forwardSpatialInterp(adjointSpatialInterp(dcf)))
    (:46-52); then abs(normalise(dcf, 2)) (:54-56).  Arithmetic in the
    precision `prec` throughout."""
    os_, W = 3.0, 4.0
    rt, ct = _rt(prec), _ct(prec)
    M = len(kx)
    kz = np.zeros(M) if kz is None else kz
    Gx, Gy, Gz = grid_shape(Ninplane, Ninplane, Nz, os_)
    dcf = np.ones(M, dtype=ct)
    dcf /= rt(M) * rt(M)
    for _ in range(2):
        c1 = grid_adjoint(dcf, kx, ky, kz, Ninplane, Ninplane, Nz, Gx * Gy * Gz, os_, W, prec)
        c2 = grid_forward(c1, kx, ky, kz, Ninplane, Ninplane, Nz, os_, W, prec)
        dcf = (dcf / c2.real).astype(ct)
    n = arma_norm2(dcf)
    dcf = dcf / (n if n != 0 else rt(1))  # normalise(X, 2) divides by 1 when the norm is 0
    return np.abs(dcf).astype(rt)


def arma_norm2(x):
    """norm(X, 2) of Armadillo 9.200 (a third-party dependency of the
    reference, not vendored): the direct sqrt(sum |x_i|^2) in the element
    precision, and, when that underflows to 0 or overflows, the scaled sum
    max|x| * sqrt(sum (|x_i| / max|x|)^2).  The density-compensation weights
    before normalisation are ~1e-30 in FP32, where the direct sum underflows."""
    a = np.abs(x)
    rt = a.dtype.type
    acc = rt(0)
    with np.errstate(over="ignore", under="ignore"):
        for v in np.square(a):  # sequential, in element precision
            acc = rt(acc + v)
    r = rt(np.sqrt(acc))
    if r != 0 and np.isfinite(r):
        return r
    mx = a.max() if a.size else rt(0)
    if mx == 0:
        return rt(0)
    acc = rt(0)
    for v in np.square(a / mx):
        acc = rt(acc + v)
    return rt(np.sqrt(acc) * mx)


def grid_coil_images(G, dcf, data, ncoils):
    """The coil loop of gridCoilImages (directRecon.cpp:118-127): image c =
    G / (dcf % data_c), coil-major."""
    M = G.n_out
    d = np.asarray(data).reshape(ncoils, M)
    w = np.asarray(dcf, dtype=_rt(G.suf))
    return np.concatenate([G / (w * d[c]).astype(_ct(G.suf)) for c in range(ncoils)])


def sum_of_squares(coil_images, nvox, ncoils):
    """calcSumOfSquaresImage (directRecon.cpp:63-73): sqrt(sum_c |x_c|), in
    the input's precision, coils summed in order."""
    x = np.asarray(coil_images).reshape(ncoils, nvox)
    rt = np.float32 if x.dtype == np.complex64 else np.float64
    acc = np.zeros(nvox, dtype=rt)
    for c in range(ncoils):
        acc += np.abs(x[c]).astype(rt)
    return np.sqrt(acc)
