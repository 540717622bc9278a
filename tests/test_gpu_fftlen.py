"""GPU parity on oversampled grids whose lengths have the prime factors 5,
7, 11 and 13 (the mixed-radix LDS Stockham FFT's odd radices): SENSE over
TimeSegmentation over Gnufft against the FP32 and FP64 restatements, whose
FFT is a plain mixed-radix DFT of any length.  Tolerances as in
test_gpu_ops.py (1e-4 of FP32, 5e-6 of FP64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import powergrid_amd as P
from powergrid_amd import synth as S


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# image N -> grid 2N: 160 = 2^5 5, 176 = 2^4 11, 208 = 2^4 13, 224 = 2^5 7,
# 120 = 2^3 3 5; 3-D: 20 = 2^2 5 and 28 = 2^2 7 along z
@pytest.mark.parametrize("N,Nz", [(80, 1), (88, 1), (104, 1), (112, 1), (60, 1), (16, 10),
                                  (12, 14)])
def test_sense_ts_odd_radix_grids(orc, N, Nz):
    print(f"[odd-radix grid N={N} Nz={Nz}]", flush=True)
    nc, L = 3, 4
    pr = S.make_problem(N, 4 if Nz > 1 else 8, 256 if Nz > 1 else 1024, nc, L, Nz=Nz,
                        peak_hz=150.0, seed=N + Nz)
    x = pr.x_true.astype(np.complex64)
    Gg = P.Gnufft(pr.M, 2.0, N, N, Nz, pr.kx, pr.ky, pr.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, pr.fm, pr.t, pr.M, pr.N, L), pr.sens, pr.M, pr.N, nc)
    ref = {}
    for prec in "fd":
        Go = orc.Gnufft(pr.M, 2.0, N, N, Nz, pr.kx, pr.ky, pr.kz, prec=prec)
        To = orc.TimeSegmentation(Go, pr.fm, pr.t, pr.M, pr.N, L)
        ref[prec] = orc.SENSE(To, pr.sens.ravel(), pr.M, pr.N, nc)
    yg = Sg * x
    y64 = ref["d"] * x
    e64, e32 = rel(yg, y64), rel(yg, ref["f"] * x)
    print(f"  forward {e64:.1e} from FP64, {e32:.1e} from FP32", flush=True)
    assert e64 < 5e-6 and e32 < 1e-4
    xa = Sg / y64
    a64, a32 = rel(xa, ref["d"] / y64), rel(xa, ref["f"] / y64)
    print(f"  adjoint {a64:.1e} from FP64, {a32:.1e} from FP32", flush=True)
    assert a64 < 5e-6 and a32 < 1e-4


def test_gridos3_grid_300_operator(orc):
    # 12 taps on a 300 = 2^2 3 5^2 point grid (gridos 3, N = 100)
    print("[odd-radix gridos 3, grid 300]", flush=True)
    N = 100
    kx, ky, kz = S.spiral(N, 4, 512)
    M = len(kx)
    G = P.Gnufft(M, 3.0, N, N, 1, kx, ky, kz)
    Od = orc.Gnufft(M, 3.0, N, N, 1, kx, ky, kz, prec="d")
    x = S.ellipse_phantom(N, seed=8).astype(np.complex64)
    y = (Od * x).astype(np.complex64)
    assert rel(G * x, Od * x) < 5e-6
    assert rel(G / y, Od / y) < 5e-6
