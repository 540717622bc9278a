"""GPU parity on the shapes of the measured workloads.

* plain 2-D (64x64, spiral 8x512, 1 coil, no field map, 10 PCG iterations,
  QuadPenalty 1e-3) at full size: operator and reconstruction against the
  oracle.  Its kernel width is 6 oversampled points (kernelWidth 3.0); the
  reference's hard-coded 4.0 is checked too.
* 2-D field-corrected SENSE (128x128, 16x2048, 8 coils, L=8) at full size:
  operator against the FP64 and FP32 restatements.
* 3-D stack-of-spirals SENSE with TV: a reduced size against the oracle's
  PCG, and at full size (128x128x32, 32 coils, L=12) through size-independent
  properties (adjointness, linearity) and a residual that decreases (data in
  scanner units; at the operator's own scale the reference's line search
  stops at its 1e25 denominator bound, which is checked too).
* the uneven-batch shapes (96^2 and 192^2 images: radix-3 grids of 192 and
  384 points, 16 coils / L=12): operator against the FP64 restatement.

Tolerances as in test_gpu_ops.py: operators 1e-4 rel-L2 of the FP32 and
5e-6 of the FP64 restatement.  PCG results are gated against the FP64
same-operator PCG: at 1e-3 where the problem is well conditioned, and, for
the nearly unregularised plain 2-D recon (beta 1e-3), at the distance of the
reference's own FP32 arithmetic (the FP32 restatement) from FP64, which
amplifies rounding by ~1e4 over 10 iterations.
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


def plain2d(W):
    p = S.make_problem(64, 8, 512, 1, 1, field=False, niter=10, beta=1e-3, kernel_width=W, seed=0)
    return p


@pytest.mark.parametrize("W", [3.0, 4.0])
def test_plain2d_full_size(orc, W):
    print(f"[plain2d 64x64 W={W}]", flush=True)
    p = plain2d(W)
    from oracle import oracle as O
    y = O.nudft_exact(p.x_true, p.kx, p.ky, p.kz, p.ix, p.iy, p.iz, np.zeros(p.N), p.t)
    y = S.add_noise(y, 40.0, seed=11).astype(np.complex64)
    ones = np.ones(p.N, np.complex64)
    Gg = P.Gnufft(p.M, 2.0, 64, 64, 1, p.kx, p.ky, p.kz, kernel_width=W)
    Sg = P.SENSE(Gg, ones, p.M, p.N, 1)
    ops = {}
    for prec in "fd":
        Go = orc.Gnufft(p.M, 2.0, 64, 64, 1, p.kx, p.ky, p.kz, kernel_width=W, prec=prec)
        ops[prec] = (Go, orc.SENSE(Go, ones, p.M, p.N, 1))
    x = p.x_true.astype(np.complex64)
    assert rel(Gg * x, ops["d"][0] * x) < 5e-6
    assert rel(Gg / y, ops["d"][0] / y) < 5e-6
    assert rel(Gg * x, ops["f"][0] * x) < 1e-4
    Rg = P.QuadPenalty(64, 64, 1, p.beta)
    xg, info = P.reconSolve(y, Sg, Rg, niter=p.niter)
    ref = {}
    for prec in "fd":
        Ro = orc.QuadPenalty(64, 64, 1, p.beta, prec=prec)
        yy = y if prec == "f" else y.astype(np.complex128)
        ref[prec] = orc.reconSolve(yy, ops[prec][1], Ro, 64, 64, 1, p.niter)
    e64 = rel(xg, ref["d"][0])
    e32 = rel(ref["f"][0], ref["d"][0])
    print(f"  gpu vs fp64 oracle {e64:.2e}; fp32 oracle vs fp64 oracle {e32:.2e}; "
          f"fp64 vs phantom {rel(ref['d'][0], p.x_true):.3f}")
    assert info["iterations"] == ref["d"][1] == p.niter
    assert e64 < max(1e-3, e32)
    # the same recon, well regularised: gated at 1e-3
    Rg2 = P.QuadPenalty(64, 64, 1, 1e4)
    xg2, _ = P.reconSolve(y, Sg, Rg2, niter=p.niter)
    Ro2 = orc.QuadPenalty(64, 64, 1, 1e4, prec="d")
    x64, _, _ = orc.reconSolve(y.astype(np.complex128), ops["d"][1], Ro2, 64, 64, 1, p.niter)
    assert rel(xg2, x64) < 1e-3


def test_sense2d_full_size_operator(orc):
    print("[sense2d 128x128 nc=8 L=8 operator]", flush=True)
    p = S.config_sense2d(seed=0)
    x = p.x_true.astype(np.complex64)
    Gg = P.Gnufft(p.M, 2.0, 128, 128, 1, p.kx, p.ky, p.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    orc.set_threads(8)
    try:
        refs = {}
        for prec in "fd":
            Go = orc.Gnufft(p.M, 2.0, 128, 128, 1, p.kx, p.ky, p.kz, prec=prec)
            To = orc.TimeSegmentation(Go, p.fm, p.t, p.M, p.N, p.L)
            refs[prec] = orc.SENSE(To, p.sens.ravel(), p.M, p.N, p.nc)
        yf = Sg * x
        assert rel(yf, refs["d"] * x) < 5e-6
        assert rel(yf, refs["f"] * x) < 1e-4
        yo = refs["d"] * x
        xa = Sg / yo
        assert rel(xa, refs["d"] / yo) < 5e-6
        assert rel(xa, refs["f"] / yo) < 1e-4
    finally:
        orc.set_threads(1)


def test_sos3d_reduced_pcg(orc):
    print("[sos3d reduced 32x32x8 nc=4 L=12 TV]", flush=True)
    p = S.make_problem(32, 4, 256, 4, 12, Nz=8, peak_hz=200.0, niter=5, penalty="tv",
                       beta=1e-3, delta=1e-2, seed=3)
    x = p.x_true.astype(np.complex64)
    Gg = P.Gnufft(p.M, 2.0, 32, 32, 8, p.kx, p.ky, p.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    Go = orc.Gnufft(p.M, 2.0, 32, 32, 8, p.kx, p.ky, p.kz, prec="d")
    So = orc.SENSE(orc.TimeSegmentation(Go, p.fm, p.t, p.M, p.N, p.L), p.sens.ravel(), p.M, p.N,
                   p.nc)
    y = (So * x).astype(np.complex64)
    y = S.add_noise(y.astype(np.complex128), 40.0, seed=4).astype(np.complex64)
    assert rel(Sg * x, So * x) < 5e-6
    assert rel(Sg / y, So / y) < 5e-6
    Rg = P.TVPenalty(32, 32, 8, p.beta, p.delta)
    Ro = orc.TVPenalty(32, 32, 8, p.beta, p.delta, prec="d")
    xg, info = P.reconSolve(y, Sg, Rg, niter=p.niter)
    orc.set_threads(8)
    try:
        x64, n64, _ = orc.reconSolve(y.astype(np.complex128), So, Ro, 32, 32, 8, p.niter)
    finally:
        orc.set_threads(1)
    print(f"  gpu vs fp64 oracle {rel(xg, x64):.2e}")
    assert info["iterations"] == n64
    assert rel(xg, x64) < 1e-3


def test_sos3d_full_size_properties():
    print("[sos3d full 128x128x32 nc=32 L=12]", flush=True)
    p = S.config_sos3d(seed=0)
    Gg = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    rng = np.random.default_rng(7)
    x1 = (rng.standard_normal(p.N) + 1j * rng.standard_normal(p.N)).astype(np.complex64)
    x2 = p.x_true.astype(np.complex64)
    y1 = (rng.standard_normal(p.M * p.nc) + 1j * rng.standard_normal(p.M * p.nc)).astype(np.complex64)
    Ax1 = Sg * x1
    Ay1 = Sg / y1
    lhs = np.vdot(y1.astype(np.complex128), Ax1.astype(np.complex128))
    rhs = np.vdot(Ay1.astype(np.complex128), x1.astype(np.complex128))
    print(f"  adjointness {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    lin = Sg * (a * x1 + b * x2).astype(np.complex64)
    assert rel(lin, a * Ax1 + b * (Sg * x2)) < 1e-5
    y = (Sg * x2).astype(np.complex64)
    R = P.TVPenalty(p.Nx, p.Ny, p.Nz, p.beta, p.delta)
    # at the operator's own scale the first line-search denominator exceeds
    # 1e25 and the PCG stops before its first update, as the reference's
    # (solve_pwls_pcg.hpp:137); in scanner units it runs
    xr, info = P.reconSolve(y, Sg, R, niter=3)
    assert info["iterations"] == 0 and info["exit_reason"] == "bad_denominator"
    xr, info = P.reconSolve((1e-3 * y).astype(np.complex64), Sg, R, niter=3)
    assert info["iterations"] == 3
    assert np.all(np.diff(info["resid"]) < 0)


def _sos3d_full_grid(orc, precs):
    """The 3-D workload's full grid (128x128x32, 256x256x64 oversampled,
    L=12, TV) with 2 coils and 32 partitions x 2 interleaves x 2048 samples,
    so that the oracle's FP64 restatement finishes in seconds; 24 grids per
    launch run on the z-class planes."""
    p = S.make_problem(128, 2, 2048, 2, 12, Nz=32, peak_hz=200.0, niter=2, penalty="tv",
                       beta=1e-3, delta=1e-2, seed=0)
    Gg = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
    assert P.zclasses(Gg) == 32
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    refs = {}
    for prec in precs:
        Go = orc.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz, prec=prec)
        To = orc.TimeSegmentation(Go, p.fm, p.t, p.M, p.N, p.L)
        refs[prec] = (orc.SENSE(To, p.sens.ravel(), p.M, p.N, p.nc), To, Go)
    return p, Sg, refs


def test_sos3d_full_grid_operator(orc):
    print("[sos3d full grid 128x128x32 nc=2 L=12 operator]", flush=True)
    orc.set_threads(16)
    try:
        p, Sg, refs = _sos3d_full_grid(orc, "fd")
        x = p.x_true.astype(np.complex64)
        yf = Sg * x
        y64 = refs["d"][0] * x
        e64, e32 = rel(yf, y64), rel(yf, refs["f"][0] * x)
        print(f"  forward: {e64:.2e} from FP64, {e32:.2e} from FP32", flush=True)
        assert e64 < 5e-6 and e32 < 1e-4
        xa = Sg / y64
        a64, a32 = rel(xa, refs["d"][0] / y64), rel(xa, refs["f"][0] / y64)
        print(f"  adjoint: {a64:.2e} from FP64, {a32:.2e} from FP32", flush=True)
        assert a64 < 5e-6 and a32 < 1e-4
    finally:
        orc.set_threads(1)


def test_sos3d_full_grid_pcg(orc):
    print("[sos3d full grid 128x128x32 nc=2 L=12 TV, 2 PCG iterations]", flush=True)
    orc.set_threads(16)
    try:
        p, Sg, refs = _sos3d_full_grid(orc, "d")
        So = refs["d"][0]
        # exact field-corrected data in scanner units (x 1e-3), 40 dB noise
        y = 1e-3 * S.add_noise(So * p.x_true.astype(np.complex64), 40.0, seed=4)
        y = y.astype(np.complex64)
        Rg = P.TVPenalty(p.Nx, p.Ny, p.Nz, p.beta, p.delta)
        Ro = orc.TVPenalty(p.Nx, p.Ny, p.Nz, p.beta, p.delta, prec="d")
        xg, info = P.reconSolve(y, Sg, Rg, niter=2)
        x64, n64, _ = orc.reconSolve(y.astype(np.complex128), So, Ro, p.Nx, p.Ny, p.Nz, 2)
    finally:
        orc.set_threads(1)
    e = rel(xg, x64)
    print(f"  gpu vs fp64 oracle PCG {e:.2e}", flush=True)
    assert info["iterations"] == n64 == 2
    assert e < 1e-3


@pytest.mark.parametrize("N,nc,L", [(96, 16, 4), (192, 8, 12)])
def test_uneven_batch_shapes_operator(orc, N, nc, L):
    print(f"[shape N={N} nc={nc} L={L}]", flush=True)
    p = S.make_problem(N, 8, 1024, nc, L, peak_hz=180.0, seed=5)
    x = p.x_true.astype(np.complex64)
    Gg = P.Gnufft(p.M, 2.0, N, N, 1, p.kx, p.ky, p.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, L), p.sens, p.M, p.N, nc)
    Go = orc.Gnufft(p.M, 2.0, N, N, 1, p.kx, p.ky, p.kz, prec="d")
    So = orc.SENSE(orc.TimeSegmentation(Go, p.fm, p.t, p.M, p.N, L), p.sens.ravel(), p.M, p.N, nc)
    yo = So * x
    assert rel(Sg * x, yo) < 5e-6
    assert rel(Sg / yo, So / yo) < 5e-6


def test_sos3d_fused_z_stage_matches_two_kernel_path():
    # the adjoint y IFFT fused with the adjoint z stage, against the two as
    # separate kernels with the class planes in HBM between them
    # (pg_set_tuning("zfuse", 0)): the same products and sums; the forward
    # is the same path either way
    print("[sos3d zfuse vs two kernels, 128x128x32 nc=2 L=12]", flush=True)
    p = S.make_problem(128, 2, 2048, 2, 12, Nz=32, peak_hz=200.0, seed=1)
    Gg = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    x = p.x_true.astype(np.complex64)
    rng = np.random.default_rng(3)
    y = (rng.standard_normal(p.M * p.nc) + 1j * rng.standard_normal(p.M * p.nc)).astype(np.complex64)
    out = {}
    for fuse in (1, 0):
        P.set_tuning("zfuse", fuse)
        try:
            out[fuse] = (Sg * x, Sg / y)
        finally:
            P.set_tuning("zfuse", 1)
    ef, ea = rel(out[1][0], out[0][0]), rel(out[1][1], out[0][1])
    print(f"  forward {ef:.1e}, adjoint {ea:.1e}", flush=True)
    assert ef < 1e-6 and ea < 1e-6


@pytest.mark.parametrize("N,Nz,nc,L,itype", [(128, 32, 2, 12, 1), (128, 1, 8, 8, 1),
                                             (64, 1, 4, 6, 2), (64, 1, 32, 4, 1),
                                             (64, 1, 8, 14, 2)])
def test_staged_row_spread_matches_patch_spread(N, Nz, nc, L, itype):
    # (the last two stage 36 and 38 floats per entry: the 48-float form; the
    # others 32 or fewer: the small-LDS form)
    # the patch spread with each stage's row slices staged in LDS
    # (pg_set_tuning("spread_stage", 1)) against the one that reads them per
    # entry from global memory: the same values, products and summation
    # order, so the same bits (real weights, complex min-max weights; 2-D and
    # z-class planes)
    print(f"[staged spread N={N} Nz={Nz} nc={nc} L={L} type={itype}]", flush=True)
    p = S.make_problem(N, 2 if Nz > 1 else 8, 2048, nc, L, Nz=Nz, peak_hz=200.0, seed=2)
    Gg = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
    T = P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L, interptype=itype)
    Sg = P.SENSE(T, p.sens, p.M, p.N, p.nc)
    rng = np.random.default_rng(5)
    y = (rng.standard_normal(p.M * p.nc) + 1j * rng.standard_normal(p.M * p.nc)).astype(np.complex64)
    out = {}
    for st in (1, 0):
        P.set_tuning("spread_stage", st)
        try:
            out[st] = Sg / y
        finally:
            P.set_tuning("spread_stage", 1)
    np.testing.assert_array_equal(out[1], out[0])


@pytest.mark.parametrize("N,Nz,nc,L,itype", [(128, 32, 2, 12, 1), (128, 1, 8, 8, 1),
                                             (64, 1, 4, 6, 2)])
def test_staged_row_mfma_spread_matches_value_row_mfma(N, Nz, nc, L, itype):
    # the matrix-core patch spread from compact rows staged in LDS against
    # the one that reads rows of nc*L values (spread_stage 0): the same
    # operands and chunks of 4 entries, so the same bits
    print(f"[staged mfma spread N={N} Nz={Nz} nc={nc} L={L} type={itype}]", flush=True)
    p = S.make_problem(N, 2 if Nz > 1 else 8, 2048, nc, L, Nz=Nz, peak_hz=200.0, seed=2)
    Gg = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
    T = P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L, interptype=itype)
    Sg = P.SENSE(T, p.sens, p.M, p.N, p.nc)
    rng = np.random.default_rng(6)
    y = (rng.standard_normal(p.M * p.nc) + 1j * rng.standard_normal(p.M * p.nc)).astype(np.complex64)
    out = {}
    P.set_tuning("spread_mfma", 1)
    try:
        for st in (1, 0):
            P.set_tuning("spread_stage", st)
            try:
                out[st] = Sg / y
            finally:
                P.set_tuning("spread_stage", 1)
    finally:
        P.set_tuning("spread_mfma", 0)
    np.testing.assert_array_equal(out[1], out[0])


@pytest.mark.parametrize("N,Nz,nc,L", [(64, 32, 2, 12), (64, 20, 2, 8), (128, 32, 2, 12)])
def test_sos3d_mfma_z_stage_matches_vector_z_stage(N, Nz, nc, L):
    # the z stages on the matrix cores (pg_set_tuning "zstage_mfma" 1: the
    # forward one; "zadj_mfma" 1: the adjoint's, fused with its y IFFT)
    # against the vector-unit ones: the same products, the class and z sums
    # in another order, so equal to FP32 rounding
    print(f"[sos3d z stage mfma vs vector, {N}x{N}x{Nz} nc={nc} L={L}]", flush=True)
    p = S.make_problem(N, 2, 1024 if N == 64 else 2048, nc, L, Nz=Nz, peak_hz=200.0, seed=4)
    Gg = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    rng = np.random.default_rng(8)
    x = (rng.standard_normal(p.N) + 1j * rng.standard_normal(p.N)).astype(np.complex64)
    y = (rng.standard_normal(p.M * p.nc) + 1j * rng.standard_normal(p.M * p.nc)).astype(np.complex64)
    out = {}
    for m in (1, 0):
        P.set_tuning("zstage_mfma", m)
        P.set_tuning("zadj_mfma", m)
        try:
            out[m] = (Sg * x, Sg / y)
        finally:
            P.set_tuning("zstage_mfma", 1)
            P.set_tuning("zadj_mfma", 1)
    ef, ea = rel(out[1][0], out[0][0]), rel(out[1][1], out[0][1])
    print(f"  forward {ef:.1e}, adjoint {ea:.1e}", flush=True)
    assert ef < 1e-6 and ea < 1e-6


@pytest.mark.parametrize("N,Nz,nc,L", [(64, 32, 2, 12), (64, 20, 2, 8), (128, 32, 3, 12)])
def test_sos3d_fused_forward_z_stage_matches_two_kernels(N, Nz, nc, L):
    # the forward z stage fused with the class planes' y FFT (pg_set_tuning
    # "zfwd_fuse" 1) against the two kernels with the class planes in HBM:
    # the same products in the same order, so bit-identical
    print(f"[sos3d fused forward z stage, {N}x{N}x{Nz} nc={nc} L={L}]", flush=True)
    p = S.make_problem(N, 2, 1024 if N == 64 else 2048, nc, L, Nz=Nz, peak_hz=200.0, seed=5)
    Gg = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    rng = np.random.default_rng(9)
    x = (rng.standard_normal(p.N) + 1j * rng.standard_normal(p.N)).astype(np.complex64)
    out = {}
    for m in (1, 0):
        P.set_tuning("zfwd_fuse", m)
        try:
            out[m] = Sg * x
        finally:
            P.set_tuning("zfwd_fuse", 1)
    print(f"  forward {rel(out[1], out[0]):.1e}", flush=True)
    np.testing.assert_array_equal(out[1], out[0])
