"""GPU parity of the shot-phase-corrected SENSE operators (pcSenseTimeSeg,
and pcSENSE over an exact DFT per shot) against the oracle's restatement,
through the C ABI.

Tolerances as for the other operators: 1e-4 rel-L2 of the FP32 and 5e-6 of
the FP64 restatement; PCG within 1e-3 NRMSE of the FP64 same-operator PCG.
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


def case(Ns=3, Nd=1024, N=32, nc=4, seed=0):
    rng = np.random.default_rng(seed)
    kx, ky, kz = S.spiral(N, Ns, Nd)
    t = S.readout_times(Ns, Nd, dwell=2.5e-5, te=5e-4)
    fm = S.field_map(N, peak_hz=120.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    # smooth per-shot phase maps (motion-induced phase of multi-shot data)
    ix, iy, _ = S.image_coords(N, N, 1, np.float64)
    phase = np.stack([rng.uniform(-1, 1) * np.pi * (ix * rng.uniform(-1, 1) + iy * rng.uniform(-1, 1))
                      + rng.uniform(-np.pi, np.pi) for _ in range(Ns)]).astype(np.float32)
    return kx, ky, kz, t, fm, sens, phase


@pytest.mark.parametrize("mode", [P.MODE_BATCHED, P.MODE_PLAIN])
def test_pcsense_matches_oracle(orc, mode):
    print(f"[pcsense mode={mode}]", flush=True)
    Ns, Nd, N, nc, L = 3, 1024, 32, 4, 4
    kx, ky, kz, t, fm, sens, phase = case(Ns, Nd, N, nc)
    A = P.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens, fm, phase)
    A.set_mode(mode)
    assert A.n_in == N * N and A.n_out == Nd * Ns * nc
    Of = orc.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens.ravel(), fm, phase.ravel(),
                            prec="f")
    Od = orc.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens.ravel(), fm, phase.ravel(),
                            prec="d")
    x = S.ellipse_phantom(N, seed=1).astype(np.complex64)
    y = A * x
    assert rel(y, Of * x) < 1e-4
    assert rel(y, Od * x) < 5e-6
    yo = (Od * x).astype(np.complex64)
    xa = A / yo
    assert rel(xa, Of / yo) < 1e-4
    assert rel(xa, Od / yo) < 5e-6


def test_pcsense_one_shot_is_sense_ts():
    print("[pcsense one shot]", flush=True)
    Nd, N, nc, L = 2048, 32, 3, 5
    kx, ky, kz, t, fm, sens, _ = case(1, Nd, N, nc)
    A = P.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens, fm, np.zeros(N * N, np.float32))
    G = P.Gnufft(Nd, 2.0, N, N, 1, kx, ky, kz)
    B = P.SENSE(P.TimeSegmentation(G, fm, t, Nd, N * N, L), sens, Nd, N * N, nc)
    x = S.ellipse_phantom(N, seed=2).astype(np.complex64)
    np.testing.assert_array_equal(A * x, B * x)
    y = B * x
    np.testing.assert_array_equal(A / y, B / y)


def test_pcsense_pcg_matches_oracle(orc):
    print("[pcsense pcg]", flush=True)
    Ns, Nd, N, nc, L, niter = 2, 1024, 32, 4, 3, 6
    kx, ky, kz, t, fm, sens, phase = case(Ns, Nd, N, nc, seed=4)
    A = P.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens, fm, phase)
    Od = orc.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens.ravel(), fm, phase.ravel(),
                            prec="d")
    x = S.ellipse_phantom(N, seed=3)
    y = S.add_noise(Od * x, 40.0, seed=5).astype(np.complex64)
    R = P.QuadPenalty(N, N, 1, 1e2)
    Ro = orc.QuadPenalty(N, N, 1, 1e2, prec="d")
    x64, n64, _ = orc.reconSolve(y.astype(np.complex128), Od, Ro, N, N, 1, niter)
    for mode in (P.CG_GRAPH, P.CG_HOST):
        xg, info = P.reconSolve(y, A, R, niter=niter, mode=mode)
        print(f"  mode={mode} nrmse vs fp64 oracle {rel(xg, x64):.2e}")
        assert info["iterations"] == n64 == niter
        assert rel(xg, x64) < 1e-3


@pytest.mark.parametrize("mode,Nd", [(P.MODE_BATCHED, 512), (P.MODE_PLAIN, 512),
                                     (P.MODE_BATCHED, 500)])
def test_pcsense_gdft_matches_oracle(orc, mode, Nd):
    # pcSENSE.cpp:39-153 over an exact DFT per shot; FP32 hardware trig, so
    # 1e-4 against both restatements (whose sin/cos are FP32 as well).  The
    # batched mode runs every shot in one launch; a shot of 500 samples ends
    # inside a 64-sample forward tile, and its adjoint slices of 250 end
    # inside a 64-sample step.
    print(f"[pcsense gdft mode={mode} Nd={Nd}]", flush=True)
    Ns, N, nc = 3, 16, 4
    kx, ky, kz, t, fm, sens, phase = case(Ns, Nd, N, nc, seed=2)
    A = P.PcSENSE(kx, ky, kz, N, N, 1, nc, t, sens, fm, phase)
    A.set_mode(mode)
    assert A.n_in == N * N and A.n_out == Nd * Ns * nc
    x = S.ellipse_phantom(N, seed=1).astype(np.complex64)
    for prec in "fd":
        O = orc.PcSENSE(kx, ky, kz, N, N, 1, nc, t, sens.ravel(), fm, phase.ravel(), prec=prec)
        y = (O * x).astype(np.complex64)
        assert rel(A * x, O * x) < 1e-4
        assert rel(A / y, O / y) < 1e-4


def test_pcsense_gdft_pcg_matches_oracle(orc):
    print("[pcsense gdft pcg]", flush=True)
    Ns, Nd, N, nc, niter = 2, 512, 16, 3, 5
    kx, ky, kz, t, fm, sens, phase = case(Ns, Nd, N, nc, seed=7)
    A = P.PcSENSE(kx, ky, kz, N, N, 1, nc, t, sens, fm, phase)
    Od = orc.PcSENSE(kx, ky, kz, N, N, 1, nc, t, sens.ravel(), fm, phase.ravel(), prec="d")
    x = S.ellipse_phantom(N, seed=3)
    y = S.add_noise(Od * x, 40.0, seed=5).astype(np.complex64)
    R = P.QuadPenalty(N, N, 1, 1e2)
    Ro = orc.QuadPenalty(N, N, 1, 1e2, prec="d")
    x64, n64, _ = orc.reconSolve(y.astype(np.complex128), Od, Ro, N, N, 1, niter)
    for mode in (P.CG_GRAPH, P.CG_HOST):
        xg, info = P.reconSolve(y, A, R, niter=niter, mode=mode)
        print(f"  mode={mode} nrmse vs fp64 oracle {rel(xg, x64):.2e}")
        assert info["iterations"] == n64
        assert rel(xg, x64) < 1e-3


@pytest.mark.parametrize("mode", [P.MODE_BATCHED, P.MODE_PLAIN])
def test_pcsense_rank_shares_match_oracle(orc, mode):
    # MPI/mpipcSENSETimeSeg.cpp:101-262: 4 shots x 6 coils dealt to 5 ranks;
    # each rank's share against the oracle's, and the shares gathered to the
    # full data layout are a whole, shots = Ns segmentation included
    print(f"[pcsense rank shares mode={mode}]", flush=True)
    Ns, Nd, N, nc, L, nr = 4, 1024, 32, 6, 4, 5
    kx, ky, kz, t, fm, sens, phase = case(Ns, Nd, N, nc, seed=6)
    x = S.ellipse_phantom(N, seed=7).astype(np.complex64)
    Y = np.zeros((Ns * nc, Nd), np.complex128)
    img = np.zeros(N * N, np.complex128)
    yfull = orc.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, 1, 1, sens.ravel(), fm,
                               phase.ravel(), prec="d") * x
    for r in range(nr):
        A = P.PcSenseRank(kx, ky, kz, N, N, 1, nc, t, L, 1, sens, fm, phase, r, nr)
        A.set_mode(mode)
        Of = orc.PcSenseRank(kx, ky, kz, N, N, 1, nc, t, L, 1, sens.ravel(), fm, phase.ravel(),
                             r, nr, prec="f")
        Od = orc.PcSenseRank(kx, ky, kz, N, N, 1, nc, t, L, 1, sens.ravel(), fm, phase.ravel(),
                             r, nr, prec="d")
        assert [(s, c) for s, c, _ in A.tasks] == Of.tasks
        y = A * x
        assert rel(y, Of * x) < 1e-4
        assert rel(y, Od * x) < 5e-6
        yo = (Od * x).astype(np.complex64)
        xa = A / yo
        assert rel(xa, Of / yo) < 1e-4
        assert rel(xa, Od / yo) < 5e-6
        # at L = 1 the shares gather to pcSenseTimeSeg's data, and their
        # partial images sum to its adjoint
        A1 = P.PcSenseRank(kx, ky, kz, N, N, 1, nc, t, 1, 1, sens, fm, phase, r, nr)
        A1.set_mode(mode)
        cols = [col for _, _, col in A1.tasks]
        Y[cols] = (A1 * x).reshape(len(cols), Nd)
        img += A1 / yfull.reshape(Ns * nc, Nd)[cols].astype(np.complex64).ravel()
    assert rel(Y.ravel(), yfull) < 5e-6
    Of1 = orc.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, 1, 1, sens.ravel(), fm, phase.ravel(),
                             prec="d")
    assert rel(img, Of1 / yfull.astype(np.complex64)) < 5e-6
