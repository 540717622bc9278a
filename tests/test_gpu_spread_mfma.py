"""The 2-D batched spread on the matrix cores (pg_set_tuning("spread_mfma", 1),
k_spread_patch_mfma: 16x16x4 FP32 products of the entries' patch weights and
their rows of grid values) against the FP64 restatement at 5e-6, on the
shapes that exercise it: 64 grids (the 2-D SENSE workload), 72 grids (two
lane chunks) and a dense k-space centre (patches summed in parts)."""
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


@pytest.fixture
def mfma():
    P.set_tuning("spread_mfma", 1)
    try:
        yield
    finally:
        P.set_tuning("spread_mfma", 0)


@pytest.mark.parametrize("N,nshot,ns,nc,L", [(128, 16, 2048, 8, 8), (32, 8, 512, 9, 8),
                                             (32, 16, 2048, 4, 4)])
def test_spread_matrix_cores_matches_oracle(orc, mfma, N, nshot, ns, nc, L):
    print(f"[spread mfma N={N} grids={nc * L}]", flush=True)
    kx, ky, kz = S.spiral(N, nshot, ns)
    M = len(kx)
    t = S.readout_times(nshot, ns, dwell=5e-6, te=5e-4)
    fm = S.field_map(N, peak_hz=120.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, fm, t, M, N * N, L), sens, M, N * N, nc)
    Gd = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    Sd = orc.SENSE(orc.TimeSegmentation(Gd, fm, t, M, N * N, L), sens.ravel(), M, N * N, nc)
    rng = np.random.default_rng(N + nc)
    y = (rng.standard_normal(nc * M) + 1j * rng.standard_normal(nc * M)).astype(np.complex64)
    if N >= 128:
        orc.set_threads(8)
    try:
        ref = Sd / y
    finally:
        orc.set_threads(1)
    e = rel(Sg / y, ref)
    print(f"  adjoint vs fp64 oracle {e:.2e}")
    assert e < 5e-6
