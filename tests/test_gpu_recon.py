"""GPU parity of the gridding-only entry points and the direct
reconstruction support (directRecon.cpp) against the oracle, through the C
ABI.

* forward/adjointSpatialInterp (Gnufft.cpp:214-337) at 8 taps (gridos 2)
  and 12 taps (gridos 3), 2-D and 3-D: rel-L2 5e-6 of the FP64 and 1e-4 of
  the FP32 restatement, and adjointness of the pair.
* a Gnufft at gridos 3 (12 taps, 3N-point FFTs) as an operator.
* calc3DDensityCompensation (directRecon.cpp:28-60): 2-D at the measured
  spiral (128^2, 16 x 2048) against both restatements (FP32: 1e-4, the
  FP32 restatement is itself 4e-5 from FP64; FP64: 5e-6); 3-D against the
  FP64 restatement (the reference's float instantiation overflows there and
  returns NaN, which the FP32 restatement reproduces).
* gridCoilImages' coil loop (directRecon.cpp:118-127), batched and plain.
* calcSumOfSquaresImage (directRecon.cpp:63-73): 2e-6 elementwise.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import powergrid_amd as P
from powergrid_amd import synth as S


def rel(a, b):
    # in FP64: 3-D gridding sums reach 1e22, whose squares overflow FP32
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("os_,N,Nz", [(2.0, 64, 1), (3.0, 64, 1), (2.0, 32, 8), (3.0, 32, 8)])
def test_spatial_interp_matches_oracle(orc, os_, N, Nz):
    print(f"[spatial os={os_} N={N} Nz={Nz}]", flush=True)
    kx, ky, kz = S.spiral(N, 4, 512, Nz=Nz)
    M = len(kx)
    G = P.Gnufft(M, os_, N, N, Nz, kx, ky, kz)
    Gx, Gy, Gz = orc.grid_shape(N, N, Nz, os_)
    assert G.grid_shape == (Gx, Gy, Gz)
    rng = np.random.default_rng(int(os_ * 10) + Nz)
    g = (rng.standard_normal(G.grid_size) + 1j * rng.standard_normal(G.grid_size)).astype(np.complex64)
    s = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    fwd = G.forwardSpatialInterp(g)
    adj = G.adjointSpatialInterp(s)
    for prec, tol in (("d", 5e-6), ("f", 1e-4)):
        gp = g.astype(np.complex128) if prec == "d" else g
        sp_ = s.astype(np.complex128) if prec == "d" else s
        assert rel(fwd, orc.grid_forward(gp, kx, ky, kz, N, N, Nz, os_, 4.0, prec=prec)) < tol
        assert rel(adj, orc.grid_adjoint(sp_, kx, ky, kz, N, N, Nz, G.grid_size, os_, 4.0,
                                         prec=prec)) < tol
    lhs = np.vdot(s.astype(np.complex128), fwd.astype(np.complex128))
    rhs = np.vdot(adj.astype(np.complex128), g.astype(np.complex128))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    # the adjoint overwrites every grid point (no accumulation across calls)
    adj2 = G.adjointSpatialInterp(s)
    if Nz == 1:  # the 2-D patch spread has a fixed summation order
        np.testing.assert_array_equal(adj, adj2)
    else:
        assert rel(adj2, adj) < 1e-6


def test_gnufft_gridos3_operator(orc):
    print("[gnufft gridos 3 operator]", flush=True)
    N = 64
    kx, ky, kz = S.spiral(N, 8, 512)
    M = len(kx)
    G = P.Gnufft(M, 3.0, N, N, 1, kx, ky, kz)
    Od = orc.Gnufft(M, 3.0, N, N, 1, kx, ky, kz, prec="d")
    x = S.ellipse_phantom(N, seed=6).astype(np.complex64)
    y = (Od * x).astype(np.complex64)
    assert rel(G * x, Od * x) < 5e-6
    assert rel(G / y, Od / y) < 5e-6


def test_gnufft_spatial_only_grid(orc):
    # 3 x 68 = 204 = 2^2 3 17 points: a prime factor above 13, so only the
    # spatial pair exists
    print("[gnufft spatial-only grid 204]", flush=True)
    N = 68
    kx, ky, kz = S.spiral(N, 4, 256)
    M = len(kx)
    G = P.Gnufft(M, 3.0, N, N, 1, kx, ky, kz)
    s = np.ones(M, np.complex64)
    adj = G.adjointSpatialInterp(s)
    assert rel(adj, orc.grid_adjoint(s.astype(np.complex128), kx, ky, kz, N, N, 1, G.grid_size,
                                     3.0, 4.0, prec="d")) < 5e-6
    with pytest.raises(P.PgError) as e:
        G * np.zeros(N * N, np.complex64)
    assert e.value.status == 2


def test_density_compensation_2d_spiral(orc):
    print("[dcf 2-D 128^2 16x2048]", flush=True)
    kx, ky, kz = S.spiral(128, 16, 2048)
    d = P.calc3DDensityCompensation(kx, ky, kz, 128, 1)
    d64 = orc.density_compensation(kx, ky, kz, 128, 1, prec="d")
    d32 = orc.density_compensation(kx, ky, kz, 128, 1, prec="f")
    print(f"  gpu vs fp64 {rel(d, d64):.2e}, vs fp32 {rel(d, d32):.2e}, "
          f"fp32 vs fp64 {rel(d32, d64):.2e}")
    assert d.dtype == np.float32 and np.all(d > 0)
    assert rel(d, d64) < 5e-6
    assert rel(d, d32) < 1e-4
    assert abs(np.linalg.norm(d.astype(np.float64)) - 1.0) < 1e-5


def test_density_compensation_3d_stack(orc):
    print("[dcf 3-D 32x32x8]", flush=True)
    kx, ky, kz = S.spiral(32, 4, 256, Nz=8)
    d = P.calc3DDensityCompensation(kx, ky, kz, 32, 8)
    d64 = orc.density_compensation(kx, ky, kz, 32, 8, prec="d")
    with np.errstate(all="ignore"):
        d32 = orc.density_compensation(kx, ky, kz, 32, 8, prec="f")
    assert not np.all(np.isfinite(d32))  # the float instantiation overflows
    print(f"  gpu vs fp64 {rel(d, d64):.2e}")
    assert np.all(np.isfinite(d))
    assert rel(d, d64) < 5e-6


@pytest.mark.parametrize("mode", [P.MODE_BATCHED, P.MODE_PLAIN])
def test_grid_coil_images(orc, mode):
    print(f"[grid coil images mode={mode}]", flush=True)
    N, nc = 64, 8
    kx, ky, kz = S.spiral(N, 8, 1024)
    M = len(kx)
    G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz).set_mode(mode)
    Od = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    rng = np.random.default_rng(4)
    data = (rng.standard_normal(nc * M) + 1j * rng.standard_normal(nc * M)).astype(np.complex64)
    dcf = P.calc3DDensityCompensation(kx, ky, kz, N, 1)
    im = P.grid_coil_images(G, dcf, data, nc)
    ref = orc.grid_coil_images(Od, dcf.astype(np.float64), data.astype(np.complex128), nc)
    assert im.shape == (nc * N * N,)
    assert rel(im, ref) < 5e-6
    sos = P.calcSumOfSquaresImage(im, N, 1, 1, nc)
    sref = orc.sum_of_squares(ref, N * N, nc)
    assert rel(sos, sref) < 5e-6


@pytest.mark.parametrize("nvox,nc", [(128 * 128 * 4, 8), (1001, 3)])
def test_sum_of_squares(orc, nvox, nc):
    print(f"[sum of squares nvox={nvox} nc={nc}]", flush=True)
    rng = np.random.default_rng(nvox)
    x = (rng.standard_normal(nvox * nc) + 1j * rng.standard_normal(nvox * nc)).astype(np.complex64)
    Nx = int(round(np.sqrt(nvox / 4))) if nvox % 4 == 0 else None
    if Nx:
        got = P.calcSumOfSquaresImage(x, Nx, 4, 1, nc)
    else:
        got = P.calcSumOfSquaresImage(x, 1, nvox, 1, nc)
    ref32 = orc.sum_of_squares(x, nvox, nc)
    ref64 = np.sqrt(np.abs(x.astype(np.complex128)).reshape(nc, nvox).sum(0))
    assert got.dtype == np.float32
    assert np.max(np.abs(got - ref64) / ref64) < 2e-6
    assert np.max(np.abs(got - ref32) / ref32) < 2e-6


def test_nifti_mag_phase_from_device(tmp_path):
    # processNIFTI.hpp:104-119: |x| and arg(x) formed on the device
    print("[nifti mag/phase]", flush=True)
    from powergrid_amd import io as PIO
    N, Nz = 32, 3
    rng = np.random.default_rng(12)
    x = (rng.standard_normal(N * N * Nz) + 1j * rng.standard_normal(N * N * Nz)).astype(np.complex64)
    xd = P.DeviceArray.from_host(x)
    PIO.write_nifti_mag_phase(tmp_path / "img", xd, N, N, Nz)
    hm, mag = PIO.read_nifti(tmp_path / "img_mag.nii")
    hp, phs = PIO.read_nifti(tmp_path / "img_phs.nii")
    assert hm["dim"][:4] == [3, N, N, Nz] and hp["magic"] == b"n+1\0"
    x64 = x.astype(np.complex128)
    assert np.max(np.abs(mag - np.abs(x64)) / np.abs(x64)) < 2e-7
    assert np.max(np.abs(phs - np.angle(x64))) < 1e-6
