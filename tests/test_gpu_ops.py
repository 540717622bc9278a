"""GPU parity: every operator of the HIP path against the CPU restatement
(oracle/) on the same seeded inputs, through the C ABI.

Tolerances: every operator within 1e-4 relative L2 of the FP32
restatement (the stated bar; measured ~2e-5, the FP32 reference's own
distance from FP64, dominated by forming k + N/2 in float).  The HIP path
computes its tables and tap weights in FP64 and does FP32 arithmetic, so it
is also checked against the FP64 restatement of the same operator at 5e-6.
Gdft precision 0 uses hardware sin/cos on revolutions, bounded at 1e-4.
CG results within 1e-3 NRMSE of the FP64 restatement's PCG.
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


def pair(orc, make):
    """the same operator from the FP32 and the FP64 restatement"""
    return make("f"), make("d")


def check(gpu, ref32, ref64, tol64=5e-6):
    e32, e64 = rel(gpu, ref32), rel(gpu, ref64)
    assert e32 < 1e-4, (e32, e64)
    assert e64 < tol64, (e32, e64)
    return e32, e64


def traj(N, nshot, ns, Nz=1, edge=True):
    kx, ky, kz = S.spiral(N, nshot, ns, Nz=Nz)
    if edge:
        # exercise the clamped window and integer (grid-aligned) positions
        ex = np.array([-N / 2, N / 2 - 1e-3, 0.0, 3.0, -5.5, N / 2 - 0.25], np.float32)
        ey = np.array([0.0, -N / 2, -N / 2, 2.0, 7.25, N / 2 - 0.25], np.float32)
        ez = np.zeros_like(ex) if Nz == 1 else np.array([-Nz / 2, 0, 1, Nz / 2 - 1, -1, 2], np.float32)
        kx, ky, kz = np.concatenate([kx, ex]), np.concatenate([ky, ey]), np.concatenate([kz, ez])
    return kx.astype(np.float32), ky.astype(np.float32), kz.astype(np.float32)


@pytest.mark.parametrize("N,W", [(32, 4.0), (96, 4.0), (128, 4.0), (64, 3.0), (256, 4.0),
                                 (512, 4.0)])
def test_gnufft2d_matches_oracle(orc, N, W):
    print(f"[gnufft2d N={N} W={W}]", flush=True)
    kx, ky, kz = traj(N, 8, 1024)
    M = len(kx)
    x = S.ellipse_phantom(N, seed=1).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, kernel_width=W)
    Gf, Gd = pair(orc, lambda p: orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, kernel_width=W, prec=p))
    check(Gg * x, Gf * x, Gd * x)
    yo = Gd * x
    check(Gg / yo, Gf / yo, Gd / yo)


@pytest.mark.parametrize("zclass", [1, 0])
@pytest.mark.parametrize("N,Nz", [(16, 8), (16, 32)])
def test_gnufft3d_matches_oracle(orc, N, Nz, zclass):
    # a stack of spirals: kz on the partition raster, so the plan grids 16+
    # grids on one x-y plane per partition (z taps in the z FFT pass), or,
    # with zclass 0, on the full 3-D grid with the tiled spread
    print(f"[gnufft3d {N}x{N}x{Nz} zclass={zclass}]", flush=True)
    kx, ky, kz = traj(N, 4, 256, Nz=Nz)
    M = len(kx)
    x = S.ellipse_phantom(N, Nz, seed=2).astype(np.complex64)
    P.set_tuning("zclass", zclass)
    try:
        Gg = P.Gnufft(M, 2.0, N, N, Nz, kx, ky, kz)
    finally:
        P.set_tuning("zclass", 1)
    assert P.zclasses(Gg) == (Nz if zclass else 0)
    Gf, Gd = pair(orc, lambda p: orc.Gnufft(M, 2.0, N, N, Nz, kx, ky, kz, prec=p))
    check(Gg * x, Gf * x, Gd * x)
    yo = Gd * x
    check(Gg / yo, Gf / yo, Gd / yo)
    # 16 grids at once take the tiled spread
    sens = S.coil_maps(N, 16, Nz).astype(np.complex64)
    Sg = P.SENSE(Gg, sens, M, x.size, 16)
    Sf, Sd = orc.SENSE(Gf, sens.ravel(), M, x.size, 16), orc.SENSE(Gd, sens.ravel(), M, x.size, 16)
    check(Sg * x, Sf * x, Sd * x)
    ys = Sd * x
    check(Sg / ys, Sf / ys, Sd / ys)


def test_gnufft3d_general_trajectory_full_grid(orc):
    # kz off the partition raster (every sample its own z window): no z
    # classes, the full 3-D grid with the tiled spread for 16 coils
    print("[gnufft3d general kz]", flush=True)
    N, Nz, nc = 16, 8, 16
    kx, ky, kz = traj(N, 4, 128, Nz=Nz)
    rng = np.random.default_rng(7)
    kz = np.clip(kz + rng.uniform(-0.5, 0.5, kz.shape), -Nz / 2, Nz / 2 - 1e-3).astype(np.float32)
    M = len(kx)
    x = S.ellipse_phantom(N, Nz, seed=2).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, Nz, kx, ky, kz)
    assert P.zclasses(Gg) == 0
    Gf, Gd = pair(orc, lambda p: orc.Gnufft(M, 2.0, N, N, Nz, kx, ky, kz, prec=p))
    sens = S.coil_maps(N, nc, Nz).astype(np.complex64)
    Sg = P.SENSE(Gg, sens, M, x.size, nc)
    Sf, Sd = orc.SENSE(Gf, sens.ravel(), M, x.size, nc), orc.SENSE(Gd, sens.ravel(), M, x.size, nc)
    check(Sg * x, Sf * x, Sd * x)
    ys = Sd * x
    check(Sg / ys, Sf / ys, Sd / ys)


def _ts_case(N=32, nshot=4, ns=512, L=4, peak=120.0):
    kx, ky, kz = traj(N, nshot, ns, edge=False)
    M = len(kx)
    t = S.readout_times(nshot, ns, dwell=2.5e-5, te=1e-3)
    fm = S.field_map(N, peak_hz=peak).astype(np.float32)
    x = S.ellipse_phantom(N, seed=3).astype(np.complex64)
    return kx, ky, kz, t, fm, x, M


@pytest.mark.parametrize("mode", [P.MODE_BATCHED, P.MODE_PLAIN])
def test_timesegmentation_matches_oracle(orc, mode):
    print(f"[ts mode={mode}]", flush=True)
    N, L = 32, 4
    kx, ky, kz, t, fm, x, M = _ts_case(N, L=L)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Tg = P.TimeSegmentation(Gg, fm, t, M, N * N, L).set_mode(mode)
    Tf, Td = pair(orc, lambda p: orc.TimeSegmentation(orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec=p),
                                                      fm, t, M, N * N, L))
    check(Tg * x, Tf * x, Td * x)
    yo = Td * x
    check(Tg / yo, Tf / yo, Td / yo)


@pytest.mark.parametrize("mode", [P.MODE_BATCHED, P.MODE_PLAIN])
@pytest.mark.parametrize("L", [1, 5, 8])
def test_sense_ts_matches_oracle(orc, mode, L):
    print(f"[sense-ts mode={mode} L={L}]", flush=True)
    N, nc = 32, 3
    kx, ky, kz, t, fm, x, M = _ts_case(N, L=L)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Tg = P.TimeSegmentation(Gg, fm, t, M, N * N, L)
    Sg = P.SENSE(Tg, sens, M, N * N, nc).set_mode(mode)
    Tg.set_mode(mode)

    def mk(p):
        Go = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec=p)
        To = orc.TimeSegmentation(Go, fm, t, M, N * N, L)
        return orc.SENSE(To, sens.ravel(), M, N * N, nc), To, Go
    (Sf, *kf), (Sd, *kd) = mk("f"), mk("d")
    check(Sg * x, Sf * x, Sd * x)
    yo = Sd * x
    check(Sg / yo, Sf / yo, Sd / yo)


@pytest.mark.parametrize("interp4", [1, 0])
@pytest.mark.parametrize("L", [8, 12])
def test_interp_staging_variants(orc, L, interp4):
    # the LDS-staged interpolation with four segments per stage (256-thread
    # blocks, items of <= 64 samples) and with all segments at once (512
    # threads, <= 128 samples): both against the FP64 restatement
    print(f"[interp variants L={L} interp_lds4={interp4}]", flush=True)
    N, nc = 64, 2
    kx, ky, kz, t, fm, x, M = _ts_case(N, nshot=8, ns=1024, L=L)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, fm, t, M, N * N, L), sens, M, N * N, nc)
    Gd = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    Sd = orc.SENSE(orc.TimeSegmentation(Gd, fm, t, M, N * N, L), sens.ravel(), M, N * N, nc)
    P.set_tuning("interp_lds4", interp4)
    try:
        yg = Sg * x
    finally:
        P.set_tuning("interp_lds4", 1)
    assert rel(yg, Sd * x) < 5e-6


def test_sense_gnufft_direct(orc):
    print("[sense-gnufft]", flush=True)
    N, nc = 64, 4
    kx, ky, kz = traj(N, 8, 512)
    M = len(kx)
    x = S.ellipse_phantom(N, seed=4).astype(np.complex64)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(Gg, sens, M, N * N, nc)

    def mk(p):
        Go = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec=p)
        return orc.SENSE(Go, sens.ravel(), M, N * N, nc), Go
    (Sf, gf), (Sd, gd) = mk("f"), mk("d")
    check(Sg * x, Sf * x, Sd * x)
    yo = Sd * x
    check(Sg / yo, Sf / yo, Sd / yo)


def test_spread_patch_split_and_many_grids(orc):
    # 64+ grids (two lane chunks), a dense k-space centre (split patches)
    print("[spread 72 grids]", flush=True)
    N, nc, L = 32, 9, 8
    kx, ky, kz, t, fm, x, M = _ts_case(N, nshot=8, ns=512, L=L)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, fm, t, M, N * N, L), sens, M, N * N, nc)
    Gd = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    Td = orc.TimeSegmentation(Gd, fm, t, M, N * N, L)
    Sd = orc.SENSE(Td, sens.ravel(), M, N * N, nc)
    yo = Sd * x
    assert rel(Sg / yo, Sd / yo) < 5e-6
    assert rel(Sg * x, yo) < 5e-6


@pytest.mark.parametrize("N", [32, 128])
def test_spread_dense_centre_parts(orc, N):
    # a 16 x 2048 spiral: at 32x32 the central 4x4 patches see up to ~6000
    # samples and are summed in up to 12 parts; the single-grid adjoint at
    # 128x128 (random data) covers work items whose last stage has fewer
    # than 64 entries after other kernels have used the LDS
    print(f"[spread dense centre N={N}]", flush=True)
    nc, L = 4, 4
    kx, ky, kz, t, fm, x, M = _ts_case(N, nshot=16, ns=2048, L=L)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(P.TimeSegmentation(Gg, fm, t, M, N * N, L), sens, M, N * N, nc)
    Gd = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    Sd = orc.SENSE(orc.TimeSegmentation(Gd, fm, t, M, N * N, L), sens.ravel(), M, N * N, nc)
    rng = np.random.default_rng(5)
    yo = (rng.standard_normal(nc * M) + 1j * rng.standard_normal(nc * M)).astype(np.complex64)
    orc.set_threads(8)
    try:
        assert rel(Sg / yo, Sd / yo) < 5e-6
        assert rel(Gg / yo[:M], Gd / yo[:M]) < 5e-6
    finally:
        orc.set_threads(1)


def test_gdft_matches_oracle(orc):
    print("[gdft]", flush=True)
    N, nc = 32, 3
    kx, ky, kz, t, fm, x, M = _ts_case(N, nshot=2, ns=400)
    ix, iy, iz = S.image_coords(N, N)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    Dg = P.Gdft(M, N * N, kx, ky, kz, ix, iy, iz, fm, t)
    Do = orc.Gdft(M, N * N, kx, ky, kz, ix, iy, iz, fm, t)
    yo = Do * x
    assert rel(Dg * x, yo) < 1e-4
    assert rel(Dg / yo, Do / yo) < 1e-4
    Sg = P.SENSE(Dg, sens, M, N * N, nc)
    So = orc.SENSE(Do, sens.ravel(), M, N * N, nc)
    ys = So * x
    assert rel(Sg * x, ys) < 1e-4
    assert rel(Sg / ys, So / ys) < 1e-4
    # FP64 mode against the exact NUDFT
    D64 = P.Gdft(M, N * N, kx, ky, kz, ix, iy, iz, fm, t, precision=1)
    ye = orc.nudft_exact(x, kx, ky, kz, ix, iy, iz, fm, t)
    assert rel(D64 * x, ye) < 1e-6
    xe = orc.nudft_exact(ye, kx, ky, kz, ix, iy, iz, fm, t, adjoint=True)
    assert rel(D64 / ye.astype(np.complex64), xe) < 1e-6


@pytest.mark.parametrize("kind", ["quad", "tv"])
def test_penalty_matches_oracle(orc, kind):
    print(f"[penalty {kind}]", flush=True)
    Nx, Ny, Nz = 16, 16, 4
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(Nx * Ny * Nz) + 1j * rng.standard_normal(Nx * Ny * Nz)).astype(np.complex64)
    d = (rng.standard_normal(Nx * Ny * Nz) + 1j * rng.standard_normal(Nx * Ny * Nz)).astype(np.complex64)
    for dims in (2, 3):
        if kind == "quad":
            Rg, Ro = P.QuadPenalty(Nx, Ny, Nz, 0.3, dims), orc.QuadPenalty(Nx, Ny, Nz, 0.3, dims, prec="f")
        else:
            Rg, Ro = P.TVPenalty(Nx, Ny, Nz, 0.3, 0.5, dims), orc.TVPenalty(Nx, Ny, Nz, 0.3, 0.5, dims, prec="f")
        assert rel(Rg.Gradient(x), Ro.Gradient(x)) < 1e-6
        dg, do = Rg.Denom(d, x), Ro.Denom(d, x)
        assert abs(dg - do) < 1e-5 * abs(do)
        assert abs(Rg.Penalty(x) - Ro.Penalty(x)) < 1e-5 * abs(Ro.Penalty(x))


def _recon_case(N=32, nc=2, L=3, niter=8):
    kx, ky, kz, t, fm, x, M = _ts_case(N, L=L)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    return kx, ky, kz, t, fm, x, M, sens


@pytest.mark.parametrize("penalty", ["quad", "tv"])
def test_pcg_matches_oracle(orc, penalty):
    # The FP32 restatement's own PCG drifts 4.3e-3 from its FP64 PCG on this
    # problem (CG amplifies FP32 rounding of the step/gamma scalars), so the
    # device PCG -- FP32 vectors, FP64 scalars -- is gated against the FP64
    # same-operator PCG at 1e-3 NRMSE and reported against the FP32 one.
    print(f"[pcg {penalty}]", flush=True)
    N, nc, L, niter = 32, 2, 3, 8
    kx, ky, kz, t, fm, x, M, sens = _recon_case(N, nc, L, niter)
    ref = {}
    for prec in ("f", "d"):
        Go = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec=prec)
        To = orc.TimeSegmentation(Go, fm, t, M, N * N, L)
        So = orc.SENSE(To, sens.ravel(), M, N * N, nc)
        if prec == "f":
            y = (So * x).astype(np.complex64)
        Ro = (orc.QuadPenalty(N, N, 1, 1e-2, prec=prec) if penalty == "quad"
              else orc.TVPenalty(N, N, 1, 1e-2, 1e-1, prec=prec))
        ref[prec] = orc.reconSolve(y, So, Ro, N, N, 1, niter)
    Rg = P.QuadPenalty(N, N, 1, 1e-2) if penalty == "quad" else P.TVPenalty(N, N, 1, 1e-2, 1e-1)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Tg = P.TimeSegmentation(Gg, fm, t, M, N * N, L)
    Sg = P.SENSE(Tg, sens, M, N * N, nc)
    for mode in (P.CG_GRAPH, P.CG_HOST):
        xg, info = P.reconSolve(y, Sg, Rg, niter=niter, mode=mode)
        e64, e32 = rel(xg, ref["d"][0]), rel(xg, ref["f"][0])
        print(f"  mode={mode} nrmse vs fp64 oracle {e64:.2e}, vs fp32 oracle {e32:.2e}")
        assert info["iterations"] == ref["d"][1] == niter
        assert e64 < 1e-3
        assert e32 < 1e-2
        assert rel(info["resid"], ref["d"][2]) < 1e-3


def test_pcg_plain_ops_match_batched():
    print("[pcg plain vs batched]", flush=True)
    N, nc, L, niter = 32, 2, 3, 6
    kx, ky, kz, t, fm, x, M, sens = _recon_case(N, nc, L, niter)
    Gg = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Tg = P.TimeSegmentation(Gg, fm, t, M, N * N, L)
    Sg = P.SENSE(Tg, sens, M, N * N, nc)
    y = Sg * x
    R = P.QuadPenalty(N, N, 1, 1e-2)
    xb, ib = P.reconSolve(y, Sg, R, niter=niter, mode=P.CG_GRAPH)
    Sg.set_mode(P.MODE_PLAIN)
    Tg.set_mode(P.MODE_PLAIN)
    xp, ip = P.reconSolve(y, Sg, R, niter=niter, mode=P.CG_HOST)
    assert ib["iterations"] == ip["iterations"] == niter
    assert rel(xp, xb) < 1e-5


def test_pcg_solve_cache_reuse_and_rebuild():
    # a solve keeps its work buffers and graph on the operator; the results of
    # reused, re-keyed and rebuilt solves equal those of fresh operators
    print("[pcg solve cache]", flush=True)
    N, nc, L, niter = 32, 2, 3, 6
    kx, ky, kz, t, fm, x, M, sens = _recon_case(N, nc, L, niter)

    def ops(nc_=nc):
        G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
        T = P.TimeSegmentation(G, fm, t, M, N * N, L)
        s_ = S.coil_maps(N, nc_).astype(np.complex64)
        return G, T, P.SENSE(T, s_, M, N * N, nc_)

    G, T, Sg = ops()
    R = P.QuadPenalty(N, N, 1, 1e-2)
    y = Sg * x
    yd = P.DeviceArray.from_host(y)
    x1, _ = P.reconSolve(yd, Sg, R, niter=niter)
    x2, _ = P.reconSolve(yd, Sg, R, niter=niter)  # cached graph
    np.testing.assert_array_equal(x1.to_host(), x2.to_host())

    def fresh(ydata, R_, n_):
        _, _, Sf = ops()
        xf, _ = P.reconSolve(ydata, Sf, R_, niter=n_)
        return xf.to_host() if hasattr(xf, "to_host") else xf

    y2 = (y * np.exp(0.3j)).astype(np.complex64)
    yd.upload(y2)  # same buffer, new data: the cached graph reads it
    x3, _ = P.reconSolve(yd, Sg, R, niter=niter)
    np.testing.assert_array_equal(x3.to_host(), fresh(y2, R, niter))
    x4, i4 = P.reconSolve(yd, Sg, R, niter=4)  # new key: iterations
    assert i4["iterations"] == 4
    np.testing.assert_array_equal(x4.to_host(), fresh(y2, R, 4))
    R2 = P.QuadPenalty(N, N, 1, 3e-2)  # new key: penalty
    x5, _ = P.reconSolve(yd, Sg, R2, niter=niter)
    np.testing.assert_array_equal(x5.to_host(), fresh(y2, R2, niter))
    # a wider composite over the same NUFFT grows its workspaces: the cached
    # graph of Sg is rebuilt on its next solve
    s4 = S.coil_maps(N, 4).astype(np.complex64)
    Sbig = P.SENSE(T, s4, M, N * N, 4)
    _ = Sbig * x
    x6, _ = P.reconSolve(yd, Sg, R, niter=niter)
    np.testing.assert_array_equal(x6.to_host(), x3.to_host())


def test_pcg_zero_data_exits_early():
    print("[pcg zero data]", flush=True)
    N, nc, L = 16, 1, 1
    kx, ky, kz = traj(N, 2, 128, edge=False)
    M = len(kx)
    G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(G, np.ones(N * N, np.complex64), M, N * N, 1)
    R = P.QuadPenalty(N, N, 1, 1e-2)
    for mode in (P.CG_GRAPH, P.CG_HOST):
        xg, info = P.reconSolve(np.zeros(M, np.complex64), Sg, R, niter=5, mode=mode)
        assert info["iterations"] == 0
        assert info["exit_reason"] in ("bad_denominator", "zero_gradient")
        assert np.all(xg == 0)


def test_device_arrays_and_stream():
    print("[device arrays]", flush=True)
    N = 32
    kx, ky, kz = traj(N, 4, 256)
    M = len(kx)
    G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    x = S.ellipse_phantom(N, seed=7).astype(np.complex64)
    s = P.Stream()
    xd = P.DeviceArray.from_host(x, stream=s)
    yd = G.forward(xd, stream=s)
    xa = G.adjoint(yd, stream=s)
    s.synchronize()
    assert rel(yd.to_host(s), G * x) < 1e-7
    # one grid at a time spreads with float atomics: run-to-run rounding only
    assert rel(xa.to_host(s), G / (G * x)) < 2e-6
    # a second call reuses the workspace and gives the same bits
    yd2 = G.forward(xd, stream=s)
    s.synchronize()
    assert np.array_equal(yd2.to_host(), yd.to_host())


def test_invalid_geometry_is_rejected():
    kx, ky, kz = traj(32, 2, 64, edge=False)
    with pytest.raises(P.PgError):
        P.Gnufft(len(kx), 2.0, 32, 16, 1, kx, ky, kz)  # Nx != Ny
    with pytest.raises(P.PgError):
        P.Gnufft(len(kx), 2.0, 32, 32, 1, kx, ky, kz, kernel_width=5.0)  # 10 taps


@pytest.mark.parametrize("case", ["2d", "3d_planes", "2d_minmax"])
def test_spread_compact_rows_match_value_rows(orc, case):
    # The spread of a time-segmented batch forms each grid's value from a
    # compact sample row (coil values and segment weights) by default; with
    # pg_set_tuning("spread_rows", 0) it reads a row of nc*L values.  The
    # products and the summation order are the same, so the adjoint and a
    # graph PCG agree bit for bit (complex min-max weights: to FP32 rounding)
    print(f"[spread rows vs value rows: {case}]", flush=True)
    if case == "3d_planes":
        p = S.make_problem(16, 4, 256, 4, 4, Nz=8, peak_hz=150.0, seed=8)
    else:
        p = S.make_problem(64, 8, 1024, 8, 8, peak_hz=150.0, seed=9)
    itype = 2 if case == "2d_minmax" else 1
    rng = np.random.default_rng(5)
    y = (rng.standard_normal(p.M * p.nc) + 1j * rng.standard_normal(p.M * p.nc)).astype(np.complex64)
    out = {}
    for rows in (1, 0):
        P.set_tuning("spread_rows", rows)
        try:
            G = P.Gnufft(p.M, 2.0, p.Nx, p.Ny, p.Nz, p.kx, p.ky, p.kz)
            T = P.TimeSegmentation(G, p.fm, p.t, p.M, p.N, p.L, interptype=itype)
            A = P.SENSE(T, p.sens, p.M, p.N, p.nc)
            R = P.QuadPenalty(p.Nx, p.Ny, p.Nz, 1e2)
            xs, _ = P.reconSolve(y, A, R, niter=3)
            out[rows] = (A / y, xs)
        finally:
            P.set_tuning("spread_rows", 1)
    if case == "2d_minmax":
        # the complex products may be contracted into FMAs differently
        assert rel(out[1][0], out[0][0]) < 1e-5
        assert rel(out[1][1], out[0][1]) < 1e-4
    else:
        np.testing.assert_array_equal(out[1][0], out[0][0])
        np.testing.assert_array_equal(out[1][1], out[0][1])


def test_trace_ranges_around_operators_and_solves():
    # pg_set_tuning("trace", 1): the operator applications, their stages and
    # the solve push roctx ranges (when libroctx64 is present); off again, none
    print("[trace ranges]", flush=True)
    p = S.make_problem(32, 4, 256, 2, 3, peak_hz=100.0, seed=12)
    G = P.Gnufft(p.M, 2.0, 32, 32, 1, p.kx, p.ky, p.kz)
    A = P.SENSE(P.TimeSegmentation(G, p.fm, p.t, p.M, p.N, p.L), p.sens, p.M, p.N, p.nc)
    x = p.x_true.astype(np.complex64)
    y0 = A * x
    avail, n0 = P.trace_ranges()
    P.trace(True)
    try:
        y1 = A * x
        A / y1
        P.reconSolve(y1, A, P.QuadPenalty(32, 32, 1, 1e2), niter=2)
    finally:
        P.trace(False)
    _, n1 = P.trace_ranges()
    A * x
    _, n2 = P.trace_ranges()
    print(f"  roctx available {avail}: {n1 - n0} ranges", flush=True)
    np.testing.assert_array_equal(y0, y1)
    if avail:
        assert n1 - n0 >= 2 * 3 + 1  # SENSE, its stages, the solve
    else:
        assert n1 == n0
    assert n2 == n1
