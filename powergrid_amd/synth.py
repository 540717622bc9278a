"""Seeded synthetic MRI inputs for the field-corrected SENSE path.

Everything here is host-side numpy and deterministic for a given seed
(numpy PCG64).  Units follow the reference: k in cycles/FOV within
[-N/2, N/2], image coordinates in [-1/2, 1/2), field map in rad/s, time in s
(TimeSegmentation.h:56, Gnufft.cpp:67-81, reconSolve.cpp:30-54).

The image vector is y-fastest: idx = y + x*Ny + z*Nx*Ny.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def image_coords(Nx, Ny, Nz=1, dtype=np.float32):
    """reconSolve.cpp:30-54 (initImageSpaceCoords), y-fastest order."""
    y = (np.arange(Ny, dtype=np.float64) - Ny / 2.0) / Ny
    x = (np.arange(Nx, dtype=np.float64) - Nx / 2.0) / Nx
    z = (np.arange(Nz, dtype=np.float64) - Nz / 2.0) / Nz
    Z, X, Y = np.meshgrid(z, x, y, indexing="ij")
    return (X.ravel().astype(dtype), Y.ravel().astype(dtype), Z.ravel().astype(dtype))


def ellipse_phantom(N, Nz=1, n_ellipses=10, seed=0):
    """Random ellipses (ellipsoids in 3-D): centres U(-0.3,0.3), semi-axes
    U(0.05,0.4), rotation U(0,pi), amplitude U(0.1,1)*exp(i U(-pi,pi));
    overlapping shapes add.  Returns complex128, y-fastest."""
    rng = np.random.default_rng(seed)
    ix, iy, iz = image_coords(N, N, Nz, np.float64)
    img = np.zeros(ix.shape, dtype=np.complex128)
    for _ in range(n_ellipses):
        cx, cy = rng.uniform(-0.3, 0.3, 2)
        a, b = rng.uniform(0.05, 0.4, 2)
        th = rng.uniform(0, np.pi)
        amp = rng.uniform(0.1, 1.0) * np.exp(1j * rng.uniform(-np.pi, np.pi))
        if Nz > 1:
            cz = rng.uniform(-0.3, 0.3)
            c = rng.uniform(0.05, 0.4)
        xr = (ix - cx) * np.cos(th) + (iy - cy) * np.sin(th)
        yr = -(ix - cx) * np.sin(th) + (iy - cy) * np.cos(th)
        r2 = (xr / a) ** 2 + (yr / b) ** 2
        if Nz > 1:
            r2 = r2 + ((iz - cz) / c) ** 2
        img[r2 <= 1.0] += amp
    return img


def spiral(N, n_interleaves, n_samples, Nz=1, dtype=np.float32):
    """Archimedean spiral k(tau) = kmax*tau*exp(i(2 pi n_turns tau + 2 pi s/Ns)),
    tau in [0,1), kmax = N/2, n_turns = ceil(kmax / Ns) so the radial step is
    <= 1/FOV.  3-D: a stack of Nz partitions at kz = p - Nz/2."""
    kmax = N / 2.0
    n_turns = max(1, math.ceil(kmax / n_interleaves))
    tau = np.arange(n_samples, dtype=np.float64) / n_samples
    kxs, kys = [], []
    for s in range(n_interleaves):
        ph = 2 * np.pi * n_turns * tau + 2 * np.pi * s / n_interleaves
        kxs.append(kmax * tau * np.cos(ph))
        kys.append(kmax * tau * np.sin(ph))
    kx = np.concatenate(kxs)
    ky = np.concatenate(kys)
    kz = np.zeros_like(kx)
    if Nz > 1:
        kx = np.tile(kx, Nz)
        ky = np.tile(ky, Nz)
        kz = np.repeat(np.arange(Nz, dtype=np.float64) - Nz / 2.0, n_interleaves * n_samples)
    return kx.astype(dtype), ky.astype(dtype), kz.astype(dtype)


def readout_times(n_shots, n_samples, dwell=12.5e-6, te=0.0, dtype=np.float32):
    """t = TE + n*dwell for each readout (same for every shot/partition)."""
    t1 = te + dwell * np.arange(n_samples, dtype=np.float64)
    return np.tile(t1, n_shots).astype(dtype)


def coil_maps(N, nc, Nz=1, sigma=0.4, radius=0.5):
    """Gaussian |S_c| = exp(-|r - r_c|^2 / (2 sigma^2)), r_c on a ring (a
    cylinder in 3-D) at angle 2 pi c/nc, phase exp(i 2 pi c/nc).
    Returns complex128 of shape (nc, N*N*Nz)."""
    ix, iy, _ = image_coords(N, N, Nz, np.float64)
    out = np.zeros((nc, ix.size), dtype=np.complex128)
    for c in range(nc):
        ang = 2 * np.pi * c / nc
        rx, ry = radius * np.cos(ang), radius * np.sin(ang)
        mag = np.exp(-((ix - rx) ** 2 + (iy - ry) ** 2) / (2 * sigma ** 2))
        out[c] = mag * np.exp(1j * ang)
    return out


def field_map(N, Nz=1, peak_hz=120.0, n_bumps=4, seed=1):
    """Sum of Gaussian bumps (centres U(-0.3,0.3), widths U(0.1,0.25), random
    sign), scaled so max |fm| = peak_hz.  Returned in rad/s."""
    rng = np.random.default_rng(seed)
    ix, iy, iz = image_coords(N, N, Nz, np.float64)
    fm = np.zeros(ix.shape)
    for _ in range(n_bumps):
        c = rng.uniform(-0.3, 0.3, 3)
        w = rng.uniform(0.1, 0.25)
        sgn = rng.choice([-1.0, 1.0])
        r2 = (ix - c[0]) ** 2 + (iy - c[1]) ** 2
        if Nz > 1:
            r2 = r2 + (iz - c[2]) ** 2
        fm += sgn * np.exp(-r2 / (2 * w * w))
    fm *= peak_hz / max(np.abs(fm).max(), 1e-12)
    return 2 * np.pi * fm


def add_noise(y, snr_db=40.0, seed=2):
    rng = np.random.default_rng(seed)
    rms = np.sqrt(np.mean(np.abs(y) ** 2))
    sd = rms * 10 ** (-snr_db / 20.0) / np.sqrt(2.0)
    return y + sd * (rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape))


@dataclass
class Problem:
    """One reconstruction's inputs (host arrays, float32/complex64 unless noted)."""
    Nx: int
    Ny: int
    Nz: int
    nc: int
    L: int
    kx: np.ndarray
    ky: np.ndarray
    kz: np.ndarray
    t: np.ndarray
    ix: np.ndarray
    iy: np.ndarray
    iz: np.ndarray
    fm: np.ndarray
    sens: np.ndarray          # (nc, N) complex64
    x_true: np.ndarray        # (N,) complex128
    data: np.ndarray = field(default=None)  # (nc*M,) complex64, coil-major
    niter: int = 20
    penalty: str = "quad"
    beta: float = 1e-3
    delta: float = 1e-2
    kernel_width: float = 4.0
    meta: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.Nx * self.Ny * self.Nz

    @property
    def M(self):
        return len(self.kx)


def make_problem(N, n_interleaves, n_samples, nc, L, Nz=1, peak_hz=120.0,
                 field=True, niter=20, penalty="quad", beta=1e-3, delta=1e-2,
                 kernel_width=4.0, seed=0):
    """Inputs for one recon, without the k-space data (which the caller
    synthesises with an exact NUDFT of its choice)."""
    kx, ky, kz = spiral(N, n_interleaves, n_samples, Nz)
    t = readout_times(n_interleaves * Nz, n_samples)
    ix, iy, iz = image_coords(N, N, Nz)
    fm = field_map(N, Nz, peak_hz, seed=seed + 1) if field else np.zeros(N * N * Nz)
    sens = coil_maps(N, nc, Nz).astype(np.complex64)
    x_true = ellipse_phantom(N, Nz, seed=seed)
    return Problem(N, N, Nz, nc, L, kx, ky, kz, t, ix, iy, iz, fm.astype(np.float32), sens,
                   x_true, None, niter, penalty, beta, delta, kernel_width,
                   dict(n_interleaves=n_interleaves, n_samples=n_samples, peak_hz=peak_hz,
                        seed=seed))


# The workloads this path is measured on (spiral MRI configurations).
def config_plain2d(seed=0):
    """64x64, spiral 8x512, 1 coil, no field map, 10 PCG iterations,
    QuadPenalty beta 1e-3."""
    return make_problem(64, 8, 512, 1, 1, field=False, niter=10, beta=1e-3, seed=seed)


def config_sense2d(seed=0, kernel_width=4.0, beta=1e4):
    """128x128, spiral 16x2048 (25.6 ms readout, 12.5 us dwell), 8 coils,
    field map peak +/-120 Hz, L=8 Hanning, 20 PCG iterations, QuadPenalty.

    beta = 1e4 regularises the unsampled k-space corners at this data scale
    (|y| ~ 1e5) without changing the image error against the phantom; with
    beta <= 1e3 the 20-iteration PCG runs in its semi-convergent regime,
    where any FP32 operator (the reference's CPU path included) lands
    1e-3..5e-3 away from the FP64 solution."""
    return make_problem(128, 16, 2048, 8, 8, peak_hz=120.0, niter=20, beta=beta,
                        kernel_width=kernel_width, seed=seed)


def config_sos3d(seed=0):
    """128x128x32 stack of spirals: 32 partitions x 8 interleaves x 2048,
    32 coils, +/-200 Hz, L=12, 15 PCG iterations, TVPenalty delta 1e-2."""
    return make_problem(128, 8, 2048, 32, 12, Nz=32, peak_hz=200.0, niter=15,
                        penalty="tv", beta=1e-3, delta=1e-2, seed=seed)
