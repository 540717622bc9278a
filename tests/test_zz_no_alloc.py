"""The allocation-free entry point pg_dft_apply_static, called through the C
ABI on the GPU, against pg_op_forward / pg_op_adjoint of the same operator
(which keep their partial sums in the handle's own workspace).  Both run the
same kernels with the same slicing, so the results are bit-identical.  Two
calls in a row, the second on a smaller operator, so that the shared static
array still holds the first call's partial sums when the second starts; an
operator whose partial sums exceed the static array is rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import powergrid_amd as P
from powergrid_amd import synth as S


def sense_gdft(N, nshot, ns, nc, Nz=1, r2=False, seed=0):
    kx, ky, kz = S.spiral(N, nshot, ns, Nz=Nz)
    M = len(kx)
    t = np.tile(S.readout_times(nshot, ns, dwell=1e-5, te=1e-3), Nz).astype(np.float32)
    ix, iy, iz = S.image_coords(N, N, Nz)
    fm = S.field_map(N, Nz=Nz, peak_hz=150.0, seed=seed).astype(np.float32)
    sens = S.coil_maps(N, nc, Nz=Nz).astype(np.complex64)
    Ni = N * N * Nz
    if r2:
        D = P.GdftR2(M, Ni, kx, ky, kz, ix, iy, iz, fm, t, N, N, Nz)
    else:
        D = P.Gdft(M, Ni, kx, ky, kz, ix, iy, iz, fm, t)
    A = P.SENSE(D, sens, M, Ni, nc)
    x = S.ellipse_phantom(N, Nz=Nz, seed=seed + 1).astype(np.complex64)
    rng = np.random.default_rng(seed)
    y = (rng.standard_normal(M * nc) + 1j * rng.standard_normal(M * nc)).astype(np.complex64)
    return D, A, x, y


def both_ways(A, x, y, stream):
    """(existing, static) results of the forward and the adjoint"""
    xd = P.DeviceArray.from_host(x)
    yd = P.DeviceArray.from_host(y)
    out = {}
    for adj, src, n in ((False, xd, A.n_out), (True, yd, A.n_in)):
        ref = A.adjoint(src, stream=stream) if adj else A.forward(src, stream=stream)
        got = P.DeviceArray(n)
        A.apply_static(src, got, adjoint=adj, stream=stream)
        out[adj] = (ref.to_host(stream), got.to_host(stream))
    return out


def test_static_matches_workspace_path_two_calls():
    print("[no-alloc: SENSE(Gdft) 32x32 nc=4, then 16x16 nc=2]", flush=True)
    stream = P.Stream()
    big = sense_gdft(32, 4, 512, 4, seed=0)
    small = sense_gdft(16, 2, 256, 2, seed=5)
    for name, (D, A, x, y) in (("big", big), ("small", small)):
        for adj, (ref, got) in both_ways(A, x, y, stream).items():
            print(f"  {name} {'adjoint' if adj else 'forward'}: "
                  f"max |diff| {np.abs(ref - got).max():.1e}", flush=True)
            assert np.all(np.isfinite(got))
            assert np.array_equal(ref, got)


def test_static_gdft_alone_and_r2():
    print("[no-alloc: Gdft alone 16x16x4; GdftR2 through SENSE]", flush=True)
    stream = P.Stream()
    D, A, x, y = sense_gdft(16, 4, 256, 3, Nz=4, r2=True, seed=2)
    for adj, (ref, got) in both_ways(A, x, y, stream).items():
        assert np.array_equal(ref, got)
    D2, _, x2, y2 = sense_gdft(16, 2, 256, 1, Nz=4, seed=3)
    for adj, (ref, got) in both_ways(D2, x2, y2[:D2.n_out], stream).items():
        assert np.array_equal(ref, got)


def test_static_rejects_oversize_and_plain_mode():
    print("[no-alloc: size and mode checks]", flush=True)
    stream = P.Stream()
    # forward partial sums: slices x coils x M = 1 x 40 x 262144 > 8M
    N, nc = 16, 40
    kx, ky, kz = S.spiral(N, 64, 4096)
    M = len(kx)
    t = S.readout_times(64, 4096, dwell=1e-6).astype(np.float32)
    ix, iy, iz = S.image_coords(N, N, 1)
    fm = np.zeros(N * N, np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    D = P.Gdft(M, N * N, kx, ky, kz, ix, iy, iz, fm, t)
    A = P.SENSE(D, sens, M, N * N, nc)
    xd = P.DeviceArray.from_host(np.ones(N * N, np.complex64))
    yd = P.DeviceArray(M * nc)
    with pytest.raises(P.PgError, match="static scratch holds"):
        A.apply_static(xd, yd, stream=stream)
    _, A2, x2, _ = sense_gdft(16, 2, 256, 2, seed=4)
    A2.set_mode(P.MODE_PLAIN)
    with pytest.raises(P.PgError, match="PG_MODE_BATCHED"):
        A2.apply_static(P.DeviceArray.from_host(x2), P.DeviceArray(A2.n_out), stream=stream)
