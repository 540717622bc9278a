"""The reference-shaped C++ API (include/powergrid/powergrid.hpp) driven by a
C++ program in the reference's own call pattern (PowerGridIsmrmrd.cpp:242-251):
Gnufft -> TimeSegmentation -> SENSE -> QuadPenalty -> reconSolve, checked
against the FP64 restatement."""
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
