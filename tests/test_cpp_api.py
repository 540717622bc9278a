"""The reference-shaped C++ API (include/powergrid/powergrid.hpp) driven by a
C++ program in the reference's own call pattern (PowerGridIsmrmrd.cpp:242-251):
Gnufft -> TimeSegmentation -> SENSE -> QuadPenalty -> reconSolve, then the
direct-reconstruction calls (density compensation, gridded coil images,
coil-combined image, the spatial interpolation pair), checked against the
FP64 restatement."""
import os
import subprocess

import numpy as np
import pytest

from powergrid_amd import synth as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "operator_api_test")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_cpp_operator_api_matches_oracle(orc, tmp_path):
    assert os.path.exists(EXE), "build() compiles tests/cpp/operator_api_test"
    N, nshot, ns, nc, L, niter, beta = 32, 4, 512, 3, 4, 6, 1e-2
    kx, ky, kz = S.spiral(N, nshot, ns)
    M = len(kx)
    t = S.readout_times(nshot, ns, dwell=2.5e-5, te=1e-3)
    fm = S.field_map(N, peak_hz=100.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    x = S.ellipse_phantom(N, seed=9).astype(np.complex64)
    Gd = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    Td = orc.TimeSegmentation(Gd, fm, t, M, N * N, L)
    Sd = orc.SENSE(Td, sens.ravel(), M, N * N, nc)
    y = (Sd * x).astype(np.complex64)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([N, M, nc, L, niter], np.int32).tofile(f)
        np.array([beta], np.float32).tofile(f)
        for a in (kx, ky, t, fm):
            np.asarray(a, np.float32).tofile(f)
        for a in (sens.ravel(), x, y):
            np.asarray(a, np.complex64).tofile(f)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = np.fromfile(fout, np.complex64, count=M * nc + 2 * N * N)
    it = int(np.fromfile(fout, np.int32, offset=8 * (M * nc + 2 * N * N))[0])
    fwd, adj, img = out[:M * nc], out[M * nc:M * nc + N * N], out[M * nc + N * N:]
    assert rel(fwd, y) < 5e-6
    assert rel(adj, Sd / y) < 5e-6
    R = orc.QuadPenalty(N, N, 1, beta, dims2penalize=2, prec="d")
    xo, n, _ = orc.reconSolve(y, Sd, R, N, N, 1, niter)
    assert it == n == niter
    assert rel(img, xo) < 1e-3
    # direct reconstruction (directRecon.cpp) through the same API
    Gn = (2 * N) ** 2
    off = 8 * (M * nc + 2 * N * N) + 4
    dcf = np.fromfile(fout, np.float32, count=M, offset=off)
    off += 4 * M
    coil = np.fromfile(fout, np.complex64, count=N * N * nc, offset=off)
    off += 8 * N * N * nc
    sos = np.fromfile(fout, np.float32, count=N * N, offset=off)
    off += 4 * N * N
    grid = np.fromfile(fout, np.complex64, count=Gn, offset=off)
    off += 8 * Gn
    samp = np.fromfile(fout, np.complex64, count=M, offset=off)
    assert samp.size == M
    dcf64 = orc.density_compensation(kx, ky, kz, N, 1, prec="d")
    assert rel(dcf, dcf64) < 5e-6
    coil64 = orc.grid_coil_images(Gd, dcf.astype(np.float64), y.astype(np.complex128), nc)
    assert rel(coil, coil64) < 5e-6
    assert rel(sos, orc.sum_of_squares(coil64, N * N, nc)) < 5e-6
    g64 = orc.grid_adjoint(y[:M].astype(np.complex128), kx, ky, kz, N, N, 1, Gn, prec="d")
    assert rel(grid, g64) < 5e-6
    assert rel(samp, orc.grid_forward(g64, kx, ky, kz, N, N, 1, prec="d")) < 5e-6
    from powergrid_amd import io as PIO
    _, mag = PIO.read_nifti(str(fout) + "_img_mag.nii")
    _, phs = PIO.read_nifti(str(fout) + "_img_phs.nii")
    np.testing.assert_allclose(mag, np.abs(img), rtol=1e-6)
    assert np.max(np.abs(np.exp(1j * phs) - np.exp(1j * np.angle(img)))) < 1e-5


def test_header_templates_compile():
    # every class template of include/powergrid/powergrid.hpp instantiated
    # explicitly (tests/cpp/header_instantiation.cpp), with the host compiler
    src = os.path.join(ROOT, "tests", "cpp", "header_instantiation.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall",
                        "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
