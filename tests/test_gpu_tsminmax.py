"""GPU parity of the min-max time-segmentation interpolators -- the exact
least-squares one (TimeSegmentation.cpp:94-150, interptype 2) and its
histogram approximation (TimeSegmentation.cpp:157-207, interptype 3) --
against the oracle's restatement, through the C ABI.

Their Gram matrix G^H G is ill-conditioned (1/cond ~1e-6 on these maps), so
the reference's float instantiation is itself 5e-4..6e-3 from its double
one.  The device forms the weights in FP64 (stored FP32), so the operator is
gated at 5e-6 against the FP64 restatement and, against the FP32
restatement, at the larger of 1e-4 and twice that restatement's own
distance from FP64.  PCG within 1e-3 of the FP64 same-operator PCG."""
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


# shots > 1: the interleaves share one readout's times, so the weights of the
# first shot's samples are repeated over the others (TimeSegmentation.cpp:71,150-156)
@pytest.mark.parametrize("itype", [2, 3])
@pytest.mark.parametrize("N,nshot,ns,L,nc,shots", [(32, 4, 512, 4, 2, 1), (32, 4, 512, 4, 2, 4),
                                                   (128, 16, 2048, 8, 4, 16)])
def test_minmax_ts_matches_oracle(orc, N, nshot, ns, L, nc, shots, itype):
    print(f"[ts minmax type {itype} N={N} L={L} nc={nc} shots={shots}]", flush=True)
    kx, ky, kz = S.spiral(N, nshot, ns)
    M = len(kx)
    t = S.readout_times(nshot, ns, dwell=12.5e-6)
    fm = S.field_map(N, peak_hz=120.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    x = S.ellipse_phantom(N, seed=1).astype(np.complex64)
    G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    T = P.TimeSegmentation(G, fm, t, M, N * N, L, interptype=itype, shots=shots)
    Sg = P.SENSE(T, sens, M, N * N, nc)
    if N >= 128:
        orc.set_threads(8)
    try:
        ops = {}
        for prec in "fd":
            Go = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec=prec)
            To = orc.TimeSegmentation(Go, fm, t, M, N * N, L, interptype=itype, shots=shots)
            ops[prec] = (To, orc.SENSE(To, sens.ravel(), M, N * N, nc), Go)
        y64 = ops["d"][0] * x
        d32 = rel(ops["f"][0] * x, y64)
        yo = y64.astype(np.complex64)
        xa64 = ops["d"][0] / y64
        ys = ops["d"][1] * x
        xs64 = ops["d"][1] / ys
        # the fused batched kernels (complex weights as real and imaginary
        # tables) and the per-segment path
        outs = {}
        for mode in (P.MODE_BATCHED, P.MODE_PLAIN):
            T.set_mode(mode)
            Sg.set_mode(mode)
            e64 = rel(T * x, y64)
            e32 = rel(T * x, ops["f"][0] * x)
            print(f"  mode {mode} fwd: gpu vs fp64 {e64:.2e}, vs fp32 {e32:.2e}; "
                  f"fp32 vs fp64 {d32:.2e}")
            assert e64 < 5e-6
            assert e32 < max(1e-4, 2 * d32)
            assert rel(T / yo, xa64) < 5e-6
            outs[mode] = Sg * x
            assert rel(outs[mode], ys) < 5e-6
            assert rel(Sg / ys.astype(np.complex64), xs64) < 5e-6
        assert rel(outs[P.MODE_BATCHED], outs[P.MODE_PLAIN]) < 2e-6
    finally:
        orc.set_threads(1)


@pytest.mark.parametrize("itype", [2, 3])
def test_minmax_ts_pcg(orc, itype):
    print(f"[ts minmax pcg type {itype}]", flush=True)
    N, nshot, ns, L, nc, niter = 32, 4, 512, 4, 3, 6
    kx, ky, kz = S.spiral(N, nshot, ns)
    M = len(kx)
    t = S.readout_times(nshot, ns, dwell=12.5e-6)
    fm = S.field_map(N, peak_hz=120.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    G = P.Gnufft(M, 2.0, N, N, 1, kx, ky, kz)
    Sg = P.SENSE(P.TimeSegmentation(G, fm, t, M, N * N, L, interptype=itype), sens, M, N * N,
                 nc)
    Go = orc.Gnufft(M, 2.0, N, N, 1, kx, ky, kz, prec="d")
    So = orc.SENSE(orc.TimeSegmentation(Go, fm, t, M, N * N, L, interptype=itype), sens.ravel(),
                   M, N * N, nc)
    xt = S.ellipse_phantom(N, seed=2)
    y = S.add_noise(So * xt, 40.0, seed=3).astype(np.complex64)
    R = P.QuadPenalty(N, N, 1, 1e2)
    Ro = orc.QuadPenalty(N, N, 1, 1e2, prec="d")
    x64, n64, _ = orc.reconSolve(y.astype(np.complex128), So, Ro, N, N, 1, niter)
    for mode in (P.CG_GRAPH, P.CG_HOST):
        xg, info = P.reconSolve(y, Sg, R, niter=niter, mode=mode)
        print(f"  mode={mode} nrmse vs fp64 oracle {rel(xg, x64):.2e}")
        assert info["iterations"] == n64
        assert rel(xg, x64) < 1e-3
