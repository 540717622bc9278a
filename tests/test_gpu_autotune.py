"""The autotuner (pg_set_tuning("autotune", 1)): a plan first used for a grid
count times its interpolation (LDS-staged or gather) and spread (vector units
or matrix cores) and keeps the faster.  The tuned operator matches the FP64
restatement at the usual 5e-6, and a tuned PCG matches the default one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import powergrid_amd as P
from powergrid_amd import synth as S


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture
def autotune():
    P.set_tuning("autotune", 1)
    try:
        yield
    finally:
        P.set_tuning("autotune", 0)


def ops(N, nshot, ns, nc, L):
    kx, ky, kz = S.spiral(N, nshot, ns)
    M = len(kx)
    t = S.readout_times(nshot, ns, dwell=5e-6, te=5e-4)
    fm = S.field_map(N, peak_hz=120.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(P.TimeSegmentation(G, fm, t, M, N * N, L), sens, M, N * N, nc)
    return (kx, ky, kz, t, fm, sens, M), G, Sg


@pytest.mark.parametrize("N,nshot,ns,nc,L", [(128, 16, 2048, 8, 8), (32, 8, 512, 9, 8)])
def test_autotuned_operator_matches_oracle(orc, autotune, N, nshot, ns, nc, L):
    print(f"[autotune N={N} grids={nc * L}]", flush=True)
    (kx, ky, kz, t, fm, sens, M), G, Sg = ops(N, nshot, ns, nc, L)
    assert P.op_tuning(Sg, nc * L) == (-1, -1)
    x = S.ellipse_phantom(N, seed=3).astype(np.complex64)
    y = Sg * x
    lds, mfma = P.op_tuning(Sg, nc * L)
    print(f"  chose interp {['gather', 'lds bins 128', 'lds 4-segment bins 64'][lds]}, "
          f"spread {'mfma' if mfma else 'valu'}")
    assert lds in (0, 1, 2) and mfma in (0, 1)
    Gd = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    Sd = orc.SENSE(orc.TimeSegmentation(Gd, fm, t, M, N * N, L), sens.ravel(), M, N * N, nc)
    if N >= 128:
        orc.set_threads(8)
    try:
        yd = Sd * x
        xa = Sd / yd
    finally:
        orc.set_threads(1)
    assert rel(y, yd) < 5e-6
    assert rel(Sg / yd.astype(np.complex64), xa) < 5e-6


def test_autotuned_pcg_matches_default(autotune):
    print("[autotune pcg]", flush=True)
    N, nc, L = 64, 8, 6
    _, G, Sg = ops(N, 8, 1024, nc, L)
    x = S.ellipse_phantom(N, seed=4).astype(np.complex64)
    y = S.add_noise(Sg * x, 40.0, seed=5).astype(np.complex64)
    R = P.QuadPenalty(N, N, 1, 1e2)
    xt, it = P.reconSolve(y, Sg, R, niter=8)
    assert P.op_tuning(Sg, nc * L) != (-1, -1)
    P.set_tuning("autotune", 0)
    _, _, S0 = ops(N, 8, 1024, nc, L)
    x0, i0 = P.reconSolve(y, S0, R, niter=8)
    assert P.op_tuning(S0, nc * L) == (-1, -1)
    assert it["iterations"] == i0["iterations"] == 8
    assert rel(xt, x0) < 1e-5


def test_autotune_untunable_shape_matches_default():
    # 3 coils without segments: no LDS interpolation and a grid count that
    # is not a multiple of 4, so nothing is timed; the zeroed timing
    # workspaces must not race the first forward/adjoint
    print("[autotune untunable nc=3 L=1]", flush=True)
    N, nc = 64, 3
    kx, ky, kz = S.spiral(N, 8, 512)
    M = len(kx)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    x = S.ellipse_phantom(N, seed=6).astype(np.complex64)
    res = []
    for tune in (1, 0):
        P.set_tuning("autotune", tune)
        try:
            G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
            Sg = P.SENSE(G, sens, M, N * N, nc)
            y = Sg * x
            res.append((y, Sg / y))
        finally:
            P.set_tuning("autotune", 0)
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


def test_cached_solve_follows_mode_and_tuning_changes():
    # a solve graph cached on a batched operator is recaptured after
    # set_mode (on the composite and on its inner operators) and after a
    # tuning change: each result equals that of a fresh operator in the new
    # configuration, bit for bit
    print("[solve cache vs mode/tuning]", flush=True)
    N, nc, L = 64, 4, 4
    (kx, ky, kz, t, fm, sens, M), _, _ = ops(N, 8, 1024, nc, L)

    def build(mode):
        G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
        T = P.TimeSegmentation(G, fm, t, M, N * N, L)
        Sg = P.SENSE(T, sens, M, N * N, nc)
        for op in (G, T, Sg):
            op.set_mode(mode)
        return G, T, Sg

    ops_b = build(P.MODE_BATCHED)
    x = S.ellipse_phantom(N, seed=7).astype(np.complex64)
    y = S.add_noise(ops_b[2] * x, 40.0, seed=8).astype(np.complex64)
    R = P.QuadPenalty(N, N, 1, 1e2)
    xb, _ = P.reconSolve(y, ops_b[2], R, niter=5)
    # the inner operators only: the composite's cached graph must follow
    for op in ops_b[:2]:
        op.set_mode(P.MODE_PLAIN)
    ops_b[2].set_mode(P.MODE_PLAIN)
    xp, _ = P.reconSolve(y, ops_b[2], R, niter=5)
    xp_fresh, _ = P.reconSolve(y, build(P.MODE_PLAIN)[2], R, niter=5)
    assert np.array_equal(xp, xp_fresh)
    for op in ops_b:
        op.set_mode(P.MODE_BATCHED)
    xb2, _ = P.reconSolve(y, ops_b[2], R, niter=5)
    assert np.array_equal(xb2, xb)
    # force the four-segment interpolation: the cached graph is recaptured
    P.set_tuning("interp_lds4", 2)
    try:
        xi, _ = P.reconSolve(y, ops_b[2], R, niter=5)
        xi_fresh, _ = P.reconSolve(y, build(P.MODE_BATCHED)[2], R, niter=5)
    finally:
        P.set_tuning("interp_lds4", 1)
    assert np.array_equal(xi, xi_fresh)
    assert rel(xi, xb) < 1e-5
    xb3, _ = P.reconSolve(y, ops_b[2], R, niter=5)
    assert np.array_equal(xb3, xb)
