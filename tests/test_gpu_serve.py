"""The serving loop (powergrid_amd.serve): a request's image is the same
alone and in a mix of concurrent requests, admission keeps the reserved
device memory within the budget, and the PCG's tolerance exit stops at the
first iteration whose gradient is below the tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import powergrid_amd as P
from powergrid_amd import serve as SV
from powergrid_amd import synth as S


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def make_request(N, nc, L, nshot, ns, Nz=1, seed=0, niter=30, tol=1e-4, beta=1e2, tag=None):
    pr = S.make_problem(N, nshot, ns, nc, L, Nz=Nz, peak_hz=100.0, seed=seed)
    M = len(pr.kx)
    G = P.Gnufft(M, 2.0, N, N, Nz, pr.kx, pr.ky, pr.kz)
    A = P.SENSE(P.TimeSegmentation(G, pr.fm, pr.t, M, pr.N, L), pr.sens, M, pr.N, nc)
    y = S.add_noise(A * pr.x_true, 40.0, seed=seed + 2).astype(np.complex64)
    return SV.ReconRequest(pr.kx, pr.ky, pr.kz, pr.t, N, N, Nz, nc, L, pr.fm, pr.sens, y,
                           beta=beta, niter=niter, tol=tol, tag=tag)


def mix():
    # matrix sizes 32..96 (a radix-3 grid among them), 3..16 coils, L 1..8,
    # one 3-D stack (4 coils x 4 segments: the deterministic tiled spread)
    return [make_request(32, 4, 4, 4, 512, seed=1, tag="a"),
            make_request(64, 8, 6, 8, 1024, seed=2, tag="b"),
            make_request(48, 3, 1, 6, 768, seed=3, tag="c"),
            make_request(32, 4, 4, 4, 512, Nz=8, seed=4, niter=12, tag="d"),
            make_request(64, 16, 8, 8, 1024, seed=5, tol=1e-3, tag="e"),
            make_request(96, 8, 8, 12, 1536, seed=6, tag="f")]


def test_serve_results_alone_and_in_the_mix():
    print("[serve alone vs mix]", flush=True)
    reqs = mix()
    st = P.Stream()
    alone = {r.tag: SV.run_request(r, st) for r in reqs}
    with SV.ReconServer(slots=4) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = {f.result().tag: f.result() for f in futs}
    assert srv.max_running >= 2
    for r in reqs:
        img, info = alone[r.tag]
        out = res[r.tag]
        print(f"  {r.tag}: {info['iterations']} iterations ({info['exit_reason']}), "
              f"latency {1e3 * out.latency:.1f} ms")
        assert out.iterations == info["iterations"] and out.exit_reason == info["exit_reason"]
        np.testing.assert_array_equal(out.image, img)


def test_serve_admission_by_memory():
    print("[serve admission]", flush=True)
    reqs = mix()[:4]
    est = [SV.estimate_bytes(r) for r in reqs]
    budget = int(1.5 * max(est))
    with SV.ReconServer(slots=4, budget_bytes=budget) as srv:
        futs = [srv.submit(r) for r in reqs]
        outs = [f.result() for f in futs]
    assert srv.peak_reserved <= budget
    assert all(o.image.size == r.N for o, r in zip(outs, reqs))
    # a request larger than the whole budget fails at once
    with SV.ReconServer(slots=1, budget_bytes=min(est) // 2) as srv:
        f = srv.submit(reqs[0])
        assert isinstance(f.exception(), MemoryError)


def test_pcg_tolerance_exit(orc):
    # the loop ends before the update of iteration k, the first with
    # ||ngrad_k|| <= tol ||ngrad_0||, ngrad = A'(y - Ax) - R'(x)
    # (solve_pwls_pcg.hpp:80-86); the iterates are those of the plain solve
    print("[pcg tolerance]", flush=True)
    N, nc, L, tol = 32, 4, 4, 1e-2
    pr = S.make_problem(N, 4, 512, nc, L, peak_hz=100.0, seed=7)
    M = len(pr.kx)
    G = P.Gnufft(M, 2.0, N, N, 1, pr.kx, pr.ky, pr.kz)
    A = P.SENSE(P.TimeSegmentation(G, pr.fm, pr.t, M, pr.N, L), pr.sens, M, pr.N, nc)
    R = P.QuadPenalty(N, N, 1, 1e2)
    y = S.add_noise(A * pr.x_true, 40.0, seed=8).astype(np.complex64)
    xk, info = P.reconSolve(y, A, R, niter=40, tol=tol)
    k = info["iterations"]
    print(f"  stopped after {k} updates ({info['exit_reason']})")
    assert info["exit_reason"] == "tolerance" and 2 <= k < 40
    xp, ip = P.reconSolve(y, A, R, niter=k)
    assert ip["exit_reason"] == "niter" and ip["iterations"] == k
    np.testing.assert_array_equal(xk, xp)
    for mode in (P.CG_GRAPH, P.CG_HOST):
        xh, ih = P.reconSolve(y, A, R, niter=40, tol=tol, mode=mode)
        assert ih["iterations"] == k
    # the FP64 restatement's gradient at the iterates k-1 and k
    Go = orc.Gnufft(M, 2.0, N, N, 1, pr.kx, pr.ky, pr.kz, prec="d")
    So = orc.SENSE(orc.TimeSegmentation(Go, pr.fm, pr.t, M, pr.N, L), pr.sens.ravel(), M,
                   pr.N, nc)
    Ro = orc.QuadPenalty(N, N, 1, 1e2, prec="d")
    yd = y.astype(np.complex128)

    def gnorm(x):
        x = np.asarray(x, np.complex128)
        return np.linalg.norm(So / (yd - So * x) - Ro.Gradient(x))

    g0 = gnorm(np.zeros(pr.N))
    xprev, _ = P.reconSolve(y, A, R, niter=k - 1)
    assert gnorm(xk) <= tol * g0 * 1.05
    assert gnorm(xprev) > tol * g0 * 0.95


def test_recon_batch_one_launch_matches_single_solves():
    # uneven reconstructions (matrix size, coils, segments, a 3-D stack) as
    # one graph launch: each image equals its own solve bit for bit
    print("[recon batch]", flush=True)
    reqs = mix()[:4]
    items, singles = [], []
    for r in reqs:
        M, N = r.M, r.N
        G = P.Gnufft(M, 2.0, r.Nx, r.Ny, r.Nz, r.kx, r.ky, r.kz)
        T = P.TimeSegmentation(G, r.fieldmap, r.tvec, M, N, r.L)
        A = P.SENSE(T, r.sens, M, N, r.nc)
        R = P.QuadPenalty(r.Nx, r.Ny, r.Nz, r.beta)
        items.append((A, r.data, R))
        singles.append((G, T))
    st = P.Stream()
    out = P.recon_solve_batch(items, niter=12, tol=1e-4, stream=st)
    for (A, y, R), (x, info) in zip(items, out):
        xs, i1 = P.reconSolve(y, A, R, niter=12, stream=st, tol=1e-4)
        assert info["iterations"] == i1["iterations"] and info["exit_reason"] == i1["exit_reason"]
        np.testing.assert_array_equal(x, xs)
    # operators that share a plan cannot run at once
    A0, y0, R0 = items[0]
    with pytest.raises(P.PgError):
        P.recon_solve_batch([(A0, y0, R0), (A0, y0, R0)], niter=2, stream=st)


# The serving workload's shapes: 2-D 128x128 (spiral 16x2048, L=8) with 8,
# 16 and 32 coils, 3-D 64x64x16 (16 partitions x 4 x 1024, 8 coils, L=8) and
# 3-D 128x128x32 (32 partitions x 8 x 2048, 8 coils, L=16), each to a
# relative gradient tolerance of 1e-4 or 30 iterations; every image is the
# same alone and in the mix, bit for bit (the 2-D kernels and the 3-D z-class
# planes sum in a fixed order).
def workload_request(N, Nz, nc, L, nsh, ns, seed, tag):
    pr = S.make_problem(N, nsh, ns, nc, L, Nz=Nz, peak_hz=120.0, seed=seed)
    M = len(pr.kx)
    G = P.Gnufft(M, 2.0, N, N, Nz, pr.kx, pr.ky, pr.kz)
    A = P.SENSE(P.TimeSegmentation(G, pr.fm, pr.t, M, pr.N, L), pr.sens, M, pr.N, nc)
    y = S.add_noise(A * pr.x_true, 40.0, seed=seed + 2)
    y = (y / np.abs(y).max()).astype(np.complex64)
    return SV.ReconRequest(pr.kx, pr.ky, pr.kz, pr.t, N, N, Nz, nc, L, pr.fm, pr.sens, y,
                           beta=1e4, niter=30, tol=1e-4, tag=tag)


def test_serve_workload_shapes_alone_and_in_the_mix():
    print("[serve: 2-D 128^2 x {8, 32} coils, 3-D 64x64x16, 3-D 128x128x32 L=16]", flush=True)
    reqs = [workload_request(128, 1, 8, 8, 16, 2048, 1, "2d8"),
            workload_request(128, 1, 32, 8, 16, 2048, 2, "2d32"),
            workload_request(64, 16, 8, 8, 4, 1024, 3, "3d64"),
            workload_request(128, 32, 8, 16, 8, 2048, 4, "3d128"),
            workload_request(128, 1, 16, 8, 16, 2048, 5, "2d16")]
    st = P.Stream()
    alone = {r.tag: SV.run_request(r, st) for r in reqs}
    with SV.ReconServer(slots=4) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = {f.result().tag: f.result() for f in futs}
    assert srv.max_running >= 2
    for r in reqs:
        img, info = alone[r.tag]
        out = res[r.tag]
        print(f"  {r.tag}: {info['iterations']} iterations ({info['exit_reason']}), "
              f"latency in the mix {1e3 * out.latency:.1f} ms", flush=True)
        assert out.iterations == info["iterations"] and out.exit_reason == info["exit_reason"]
        np.testing.assert_array_equal(out.image, img)



def test_serve_reuses_plans_with_the_same_image():
    # repeats of one protocol (same trajectory and grid count, their own data
    # and field maps): from the second on they run on the first one's idle
    # plan, and each image equals the one from a plan of its own
    print("[serve: plan reuse]", flush=True)
    base = workload_request(64, 1, 8, 8, 8, 1024, 7, "p0")
    reqs = [base]
    for j in (1, 2):
        r = workload_request(64, 1, 8, 8, 8, 1024, 7, f"p{j}")
        rng = np.random.default_rng(j)
        r.data = (base.data * (1 + 0.1 * j) + 0.01 * (rng.standard_normal(base.data.shape)
                                                       + 1j * rng.standard_normal(base.data.shape))
                  ).astype(np.complex64)
        r.fieldmap = (base.fieldmap * (1 + 0.05 * j)).astype(np.float32)
        reqs.append(r)
    st = P.Stream()
    alone = {r.tag: SV.run_request(r, st)[0] for r in reqs}
    with SV.ReconServer(slots=1) as srv:
        outs = [srv.submit(r).result() for r in reqs]
        assert srv.plan_hits == 2
    for r, o in zip(reqs, outs):
        np.testing.assert_array_equal(o.image, alone[r.tag])
