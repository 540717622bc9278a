"""Pins the oracle's direct-reconstruction support (density compensation,
gridded coil images, coil-combined magnitude image; directRecon.cpp) against
a dense-matrix evaluation written from the specification with scipy's I0,
closed forms and the Armadillo norm's documented fallback.  CPU only."""
import numpy as np
import pytest
import scipy.special as sp

from powergrid_amd import synth as S


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def dense_interp(k, N, Nz, os_=3.0, W=4.0):
    """Dense interpolation matrix B (M x grid), y fastest then x then z:
    B[i, g] = prod_d LUT[floor(d_d^2 * 4/W^2 * S)] over the taps inside the
    clamped window [max(0, ceil(s - W os/2)), min(os N - 1, floor(s + W os/2))],
    s = os (k + N/2), d = |s - n|/os, LUT[j] = I0(beta sqrt(1 - j/S))/W with
    S = 10000 W and beta of Gnufft.cpp:85-88 (written here from that
    definition, scipy's I0)."""
    beta = np.pi * np.sqrt((os_ - 0.5) ** 2 * (W * W * 4.0) / (os_ * os_) - 0.8)
    Sl = int(10000 * W)
    dims = [N, N] + ([Nz] if Nz > 1 else [])
    Gd = [int(os_ * n) for n in dims]
    M = k.shape[1]
    per_dim = []
    for d, (n, g) in enumerate(zip(dims, Gd)):
        s = os_ * (k[d] + n / 2.0)
        grid = np.arange(g)[None, :]
        lo = np.maximum(0, np.ceil(s - W * os_ / 2))[:, None]
        hi = np.minimum(os_ * n - 1, np.floor(s + W * os_ / 2))[:, None]
        dist = np.abs(s[:, None] - grid) / os_
        j = np.floor(dist * dist * 4.0 / (W * W) * Sl).astype(np.int64)
        w = np.where(j < Sl, sp.i0(beta * np.sqrt(np.clip(1.0 - j / Sl, 0, None))) / W, 0.0)
        w = np.where((grid >= lo) & (grid <= hi), w, 0.0)
        per_dim.append(w)
    wx, wy = per_dim[0], per_dim[1]
    B = (wx[:, :, None] * wy[:, None, :]).reshape(M, -1)  # [i, x, y] -> y fastest
    if Nz > 1:
        B = (per_dim[2][:, :, None] * B[:, None, :]).reshape(M, -1)  # z slowest
    return B


@pytest.mark.parametrize("N,Nz,M", [(8, 1, 60), (8, 4, 40)])
def test_density_compensation_matches_dense_pipe_iteration(orc, N, Nz, M):
    rng = np.random.default_rng(N + Nz)
    k = rng.uniform(-N / 2, N / 2, (3, M))
    if Nz > 1:
        k[2] = rng.uniform(-Nz / 2, Nz / 2, M)
    else:
        k[2] = 0.0
    B = dense_interp(k, N, Nz)
    d = np.full(M, 1.0 / (M * M), np.complex128)
    for _ in range(2):
        d = d / (B @ (B.T @ d)).real
    ref = np.abs(d / np.linalg.norm(d))
    got = orc.density_compensation(k[0], k[1], k[2], N, Nz, prec="d")
    # the oracle's I0 is the reference's polynomial (~2e-7 relative); the
    # ratio cancels most of it
    assert rel(got, ref) < 1e-7
    assert abs(np.linalg.norm(got) - 1.0) < 1e-12
    # one iteration more or less is far outside that
    d3 = d / (B @ (B.T @ d)).real
    assert rel(np.abs(d3 / np.linalg.norm(d3)), ref) > 1e-3


def test_spatial_interp_pair_matches_dense_matrix(orc):
    # forward/adjointSpatialInterp (Gnufft.cpp:214-337) at os 3: the same B
    N = 8
    rng = np.random.default_rng(5)
    k = rng.uniform(-N / 2, N / 2, (3, 30))
    k[2] = 0
    B = dense_interp(k, N, 1)
    g = rng.standard_normal(B.shape[1]) + 1j * rng.standard_normal(B.shape[1])
    s = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    assert rel(orc.grid_forward(g, k[0], k[1], k[2], N, N, 1, 3.0, 4.0, prec="d"), B @ g) < 1e-6
    assert rel(orc.grid_adjoint(s, k[0], k[1], k[2], N, N, 1, B.shape[1], 3.0, 4.0, prec="d"),
               B.T @ s) < 1e-6
    # the grid transposed (x fastest) is far off
    Bt = B.reshape(30, 24, 24).transpose(0, 2, 1).reshape(30, -1)
    assert rel(orc.grid_forward(g, k[0], k[1], k[2], N, N, 1, 3.0, 4.0, prec="d"), Bt @ g) > 0.1


def test_density_compensation_uniform_on_cartesian_interior(orc):
    # a fully sampled Cartesian grid: by symmetry every sample whose
    # neighbourhood of influence (two iterations of spread + interpolation,
    # +-8 samples) lies inside the grid gets the same weight
    N = 32
    ky, kx = np.meshgrid(np.arange(N) - N / 2, np.arange(N) - N / 2)
    kx, ky = kx.ravel(), ky.ravel()
    d = orc.density_compensation(kx, ky, np.zeros_like(kx), N, 1, prec="d")
    inner = (kx >= -N / 2 + 9) & (kx <= N / 2 - 10) & (ky >= -N / 2 + 9) & (ky <= N / 2 - 10)
    assert inner.sum() >= 16
    v = d[inner]
    assert (v.max() - v.min()) / v.max() < 1e-12
    # the edge samples see fewer neighbours: larger weight
    assert d[(kx == -N / 2) & (ky == -N / 2)][0] > v.max()


def test_dcf_fp32_normalisation_does_not_underflow(orc):
    # the weights before normalisation are ~1e-30 in FP32: the direct
    # sum of squares underflows and Armadillo's scaled fallback is needed
    kx, ky, kz = S.spiral(32, 4, 256)
    d32 = orc.density_compensation(kx, ky, kz, 32, 1, prec="f")
    d64 = orc.density_compensation(kx, ky, kz, 32, 1, prec="d")
    assert np.all(np.isfinite(d32))
    assert rel(d32, d64) < 2e-4


def test_arma_norm2_direct_and_fallback(orc):
    x = np.full(100, 1e-30 + 0j, np.complex64)
    assert np.float32(np.sqrt(np.sum(np.abs(x) ** 2))) == 0  # direct sum underflows
    assert abs(orc.arma_norm2(x) - 1e-29) < 1e-6 * 1e-29
    y = np.full(4, 3e20 + 4e20j, np.complex64)  # |y|^2 overflows FP32
    assert abs(orc.arma_norm2(y) - 1e21) < 1e-6 * 1e21
    z = np.array([3 + 4j, 0, 12j], np.complex128)
    assert orc.arma_norm2(z) == 13.0
    assert orc.arma_norm2(np.zeros(3, np.complex64)) == 0


def test_sum_of_squares_closed_forms(orc):
    rng = np.random.default_rng(2)
    nvox = 50
    x = (rng.standard_normal(nvox) + 1j * rng.standard_normal(nvox)).astype(np.complex64)
    # one coil: sqrt(|x|) (the reference sums magnitudes, not their squares)
    assert rel(orc.sum_of_squares(x, nvox, 1), np.sqrt(np.abs(x.astype(np.complex128)))) < 1e-6
    # coils x, -x, i x: sqrt(3 |x|)
    three = np.concatenate([x, -x, 1j * x])
    assert rel(orc.sum_of_squares(three, nvox, 3),
               np.sqrt(3 * np.abs(x.astype(np.complex128)))) < 1e-6


def test_grid_coil_images_is_weighted_adjoint(orc):
    N, nc = 16, 3
    kx, ky, kz = S.spiral(N, 2, 128)
    M = len(kx)
    G = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    rng = np.random.default_rng(9)
    data = rng.standard_normal(nc * M) + 1j * rng.standard_normal(nc * M)
    dcf = rng.uniform(0.5, 2.0, M)
    im = orc.grid_coil_images(G, dcf, data, nc).reshape(nc, -1)
    for c in range(nc):
        # linear in the data: G/(dcf % d) = sum_i dcf_i d_i (G / e_i) checked
        # through the adjoint identity <G x, dcf % d> = <x, image>
        x = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
        lhs = np.vdot(G * x, dcf * data[c * M:(c + 1) * M])
        rhs = np.vdot(x, im[c])
        assert abs(lhs - rhs) < 1e-10 * abs(lhs)
