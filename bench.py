"""Benchmark of the field-corrected SENSE PWLS-PCG reconstruction on one MI355X.

Headline workload (the largest configuration that fits one GPU): 3-D
stack-of-spirals SENSE with time segmentation, 128x128x32 image, 32 kz
partitions x 8 interleaves x 2048 samples (524288 per coil, 12.5 us dwell),
32 Gaussian coils on a cylinder, smooth 3-D field map (+/-200 Hz), L=12
Hanning segments, TVPenalty (beta 1e-3, delta 1e-2), 15 PCG iterations,
kernel width 4.0 (Gnufft.cpp:84), gridos 2: 384 grids of 256x256x64 per
batched launch.  k-space is the exact field-corrected NUDFT of a
10-ellipsoid phantom (FP64 phase and sums, on the GPU) plus 40 dB noise, in
scanner units (x 1e-3): at the operator's own scale the reference's line
search stops at its 1e25 denominator bound (solve_pwls_pcg.hpp:137).

A step is one reconstruction (15 PCG iterations) with every input already
resident in HBM; `value` is PCG iterations per second of the batched path
(fused coil x segment kernels, the whole solve as one HIP graph).  The plain
path (one NUFFT at a time, host-driven loop, the reference's structure) is
timed on the same inputs.  Per-kernel durations come from HIP events
recorded by the library on its launch stream; the CPU baseline is the
oracle's restatement of the reference CPU path on this host's cores.

The 2-D field-corrected SENSE workload (128x128, spiral 16x2048, 8 coils,
L=8, 20 iterations) is reported beside it under "sense2d", with its NRMSE
against the FP64 same-operator PCG and the FP64 exact-NUDFT PCG.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import collections
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def synth_data(P, prob):
    """k-space = exact field-corrected NUDFT (FP64 trig and accumulation, on
    the GPU) of the phantom through each coil map, plus complex noise."""
    from powergrid_amd import synth as S
    M, N, nc = prob.M, prob.N, prob.nc
    D = P.Gdft(M, N, prob.kx, prob.ky, prob.kz, prob.ix, prob.iy, prob.iz, prob.fm, prob.t,
               precision=1)
    Sd = P.SENSE(D, prob.sens, M, N, nc)
    y = Sd * prob.x_true.astype(np.complex64)
    return S.add_noise(y.astype(np.complex128), 40.0, seed=11).astype(np.complex64)


def build_ops(P, prob, mode):
    G = P.Gnufft(prob.M, 2.0, prob.Nx, prob.Ny, prob.Nz, prob.kx, prob.ky, prob.kz,
                 kernel_width=prob.kernel_width)
    T = P.TimeSegmentation(G, prob.fm, prob.t, prob.M, prob.N, prob.L)
    Sg = P.SENSE(T, prob.sens, prob.M, prob.N, prob.nc)
    for op in (G, T, Sg):
        op.set_mode(mode)
    R = P.QuadPenalty(prob.Nx, prob.Ny, prob.Nz, prob.beta)
    return G, T, Sg, R


def gridding_bytes(M, G, B, d):
    """algorithmic bytes of one gridding launch: coordinates once, every grid
    once, the samples of every grid once: 4 d M + B (8 M + 8 G)"""
    return 4 * d * M + B * (8 * M + 8 * G)


def fft_bytes(G, B):
    return 16 * B * G


def host_cpu():
    """the host threads this process may run on, and the CPU model"""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_baseline(prob, y, budget_s=20.0):
    """The oracle's restatement of the reference CPU path (Gnufft/TS/SENSE +
    solve_pwls_pcg, FP32, OpenMP over samples/pixels), timed on every host
    thread this process may use."""
    from oracle import oracle as O
    threads, model = host_cpu()
    O.set_threads(threads)
    G = O.Gnufft(prob.M, 2.0, prob.Nx, prob.Ny, prob.Nz, prob.kx, prob.ky, prob.kz,
                 prob.kernel_width)
    T = O.TimeSegmentation(G, prob.fm, prob.t, prob.M, prob.N, prob.L)
    Sg = O.SENSE(T, prob.sens.ravel(), prob.M, prob.N, prob.nc)
    R = O.QuadPenalty(prob.Nx, prob.Ny, prob.Nz, prob.beta)
    # one warm-up iteration, then as many single-iteration solves as fit the budget
    O.reconSolve(y, Sg, R, prob.Nx, prob.Ny, prob.Nz, 1)
    n, t0 = 0, time.perf_counter()
    while True:
        O.reconSolve(y, Sg, R, prob.Nx, prob.Ny, prob.Nz, 2)
        n += 2
        if time.perf_counter() - t0 > budget_s / 2 or n >= 6:
            break
    dt = time.perf_counter() - t0
    # the same loops on one core (the reference's serial GCC build): one solve
    # of 2 iterations
    O.set_threads(1)
    t1 = time.perf_counter()
    O.reconSolve(y, Sg, R, prob.Nx, prob.Ny, prob.Nz, 2)
    serial = 2 / (time.perf_counter() - t1)
    O.set_threads(threads)
    # the timed solves include one A*x0 forward each; report iterations/s of
    # the whole solve call (reference loop structure)
    return dict(value=n / dt, unit="CG iterations/s", cores=threads, kind="port",
                cpu_model=model, nproc=os.cpu_count(),
                sample=f"{n} PCG iterations in {n // 2} solves of 2 iterations "
                       f"(same inputs, FP32, {threads} OpenMP threads)",
                serial_value=serial,
                serial_sample="one solve of 2 PCG iterations on 1 thread")


def cpu_baseline_sos3d(prob):
    """The oracle's FP32 restatement of the reference CPU path on the 3-D
    workload, on every host thread this process may use.  A PCG iteration
    applies the 32-coil operator forward and adjoint once each; the coils are
    independent, so the sample is one coil's SENSE forward and adjoint (12
    segment NUFFTs each, with the reference's per-call deapodization, FFT
    planning and sample-array set-up), timed once and scaled by the coil
    count.  The iteration's vector work (a few passes over 16.8M data and
    0.5M image values) is left out, in the CPU path's favour."""
    from oracle import oracle as O
    threads, model = host_cpu()
    O.set_threads(threads)
    G = O.Gnufft(prob.M, 2.0, prob.Nx, prob.Ny, prob.Nz, prob.kx, prob.ky, prob.kz,
                 prob.kernel_width)
    T = O.TimeSegmentation(G, prob.fm, prob.t, prob.M, prob.N, prob.L)
    S1 = O.SENSE(T, prob.sens[0].ravel(), prob.M, prob.N, 1)
    x = prob.x_true.astype(np.complex64)
    t0 = time.perf_counter()
    y1 = S1 * x
    t1 = time.perf_counter()
    S1 / y1
    t2 = time.perf_counter()
    per_iter = prob.nc * ((t1 - t0) + (t2 - t1))
    O.set_threads(1)
    return dict(value=1.0 / per_iter, unit="CG iterations/s", cores=threads, kind="port",
                cpu_model=model, nproc=os.cpu_count(),
                sample=f"one coil's SENSE-TS forward ({t1 - t0:.2f} s) and adjoint "
                       f"({t2 - t1:.2f} s), L=12, FP32, {threads} OpenMP threads, x {prob.nc} "
                       f"coils = {per_iter:.1f} s per PCG iteration")


def fp64_oracle_recon(prob, y, niter):
    """FP64 evaluation of the same discretised operator (same L, Hanning, LUT,
    W) -- the NRMSE gate of the reconstruction."""
    from oracle import oracle as O
    O.set_threads(host_cpu()[0])
    G = O.Gnufft(prob.M, 2.0, prob.Nx, prob.Ny, prob.Nz, prob.kx, prob.ky, prob.kz,
                 prob.kernel_width, prec="d")
    T = O.TimeSegmentation(G, prob.fm, prob.t, prob.M, prob.N, prob.L)
    Sg = O.SENSE(T, prob.sens.ravel(), prob.M, prob.N, prob.nc)
    R = O.QuadPenalty(prob.Nx, prob.Ny, prob.Nz, prob.beta, prec="d")
    x, n, _ = O.reconSolve(y.astype(np.complex128), Sg, R, prob.Nx, prob.Ny, prob.Nz, niter)
    return x


def copy_bandwidth(P, stream, nbytes=1 << 30, reps=5):
    """device-to-device copy (read + write) in GB/s: the measured HBM ceiling"""
    src = P.DeviceArray(nbytes // 8)
    dst = P.DeviceArray(nbytes // 8)
    dst.copy_from(src, stream=stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_from(src, stream=stream)
    stream.synchronize()
    return 2 * nbytes * reps / (time.perf_counter() - t0) / 1e9


def gdft_roofline(P, prob, stream, reps=3):
    """The exact field-corrected DFT (Gdft, FP32 phase and trig) through SENSE
    on the workload: flops = 16 per (sample, voxel) pair per coil and
    direction, against the FP32 vector peak."""
    M, N, nc = prob.M, prob.N, prob.nc
    D = P.Gdft(M, N, prob.kx, prob.ky, prob.kz, prob.ix, prob.iy, prob.iz, prob.fm, prob.t,
               precision=0)
    Sd = P.SENSE(D, prob.sens, M, N, nc)
    dt = _pair_ms(P, Sd, prob.x_true, stream, reps) * 1e-3
    pairs = 2.0 * M * N * nc
    tf = 16.0 * pairs / dt / 1e12
    return dict(ms_per_pair_of_ops=round(1e3 * dt, 3), pairs=pairs, achieved=round(tf, 2),
                peak=FP32_PEAK_TFLOPS, unit="TFLOP/s", frac=round(tf / FP32_PEAK_TFLOPS, 4),
                bound="fp32 valu")


def build_ops_sos3d(P, prob, mode):
    G = P.Gnufft(prob.M, 2.0, prob.Nx, prob.Ny, prob.Nz, prob.kx, prob.ky, prob.kz,
                 kernel_width=prob.kernel_width)
    T = P.TimeSegmentation(G, prob.fm, prob.t, prob.M, prob.N, prob.L)
    Sg = P.SENSE(T, prob.sens, prob.M, prob.N, prob.nc)
    for op in (G, T, Sg):
        op.set_mode(mode)
    R = P.TVPenalty(prob.Nx, prob.Ny, prob.Nz, prob.beta, prob.delta)
    return G, T, Sg, R


def sos3d_data(P, prob, stream, scale=1e-3):
    """k-space of the 3-D workload: the exact field-corrected NUDFT (Gdft
    with FP64 phase, trig and sums, on the GPU) of the phantom through each
    coil map, plus 40 dB complex noise, times `scale` (scanner units)."""
    from powergrid_amd import synth as S
    D = P.Gdft(prob.M, prob.N, prob.kx, prob.ky, prob.kz, prob.ix, prob.iy, prob.iz, prob.fm,
               prob.t, precision=1)
    Sd = P.SENSE(D, prob.sens, prob.M, prob.N, prob.nc)
    xd = P.DeviceArray.from_host(prob.x_true.astype(np.complex64))
    yd = Sd.forward(xd, stream=stream)
    y = yd.to_host(stream).astype(np.complex128)
    del yd, xd, Sd, D
    return (scale * S.add_noise(y, 40.0, seed=11)).astype(np.complex64)


def exact_cg(P, prob, yd, R, niter, stream):
    """The same PCG (same data, penalty and iteration count) with the exact
    field-corrected DFT as its operator: Gdft with FP64 phase, trig and sums
    on the GPU, through SENSE; the FP64 exact-NUDFT reconstruction the
    gridded one is compared with."""
    D = P.Gdft(prob.M, prob.N, prob.kx, prob.ky, prob.kz, prob.ix, prob.iy, prob.iz, prob.fm,
               prob.t, precision=1)
    Sd = P.SENSE(D, prob.sens, prob.M, prob.N, prob.nc)
    x, info = P.reconSolve(yd, Sd, R, niter=niter, mode=P.CG_GRAPH, stream=stream)
    out = x.to_host()
    del Sd, D
    return out, info


def direct_recon_rates(P, prob, stream, reps=5):
    """The direct-reconstruction support on the workload's trajectory:
    density compensation (one call, including its 3x-grid plan set-up), the
    dcf-weighted adjoint of every coil, and the coil-combined magnitude
    image over a 128x128x32 x 32-coil volume against the HBM peak (bytes:
    8 per complex read, 4 per voxel written)."""
    t0 = time.perf_counter()
    dcf = P.calc3DDensityCompensation(prob.kx, prob.ky, prob.kz, prob.Nx, prob.Nz)
    t_dcf = time.perf_counter() - t0
    G = P.Gnufft(prob.M, 2.0, prob.Nx, prob.Ny, prob.Nz, prob.kx, prob.ky, prob.kz)
    rng = np.random.default_rng(3)
    n = prob.M * prob.nc
    dd = P.DeviceArray.from_host(dcf)
    yd = P.DeviceArray.from_host((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64))
    P.grid_coil_images(G, dd, yd, prob.nc, stream=stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        im = P.grid_coil_images(G, dd, yd, prob.nc, stream=stream)
    stream.synchronize()
    t_grid = (time.perf_counter() - t0) / reps
    nvox, nc = 128 * 128 * 32, 32
    xd = P.DeviceArray(nvox * nc)
    xd.upload(np.ones(nvox * nc, np.complex64))
    P.calcSumOfSquaresImage(xd, 128, 32, 1, nc, stream=stream)
    stream.synchronize()
    P.profile.enable(True)
    P.profile.reset()
    for _ in range(reps):
        P.calcSumOfSquaresImage(xd, 128, 32, 1, nc, stream=stream)
    stream.synchronize()
    cnt, ms = P.profile.query("sum_of_squares")
    P.profile.enable(False)
    t_sos = ms / 1e3 / max(cnt, 1)
    gbs = (8.0 * nvox * nc + 4.0 * nvox) / t_sos / 1e9
    del im
    return dict(dcf_ms=round(1e3 * t_dcf, 2), grid_coil_images_ms=round(1e3 * t_grid, 3),
                sum_of_squares={"us": round(1e6 * t_sos, 2), "achieved": round(gbs, 1),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(gbs / HBM_PEAK_GBS, 3),
                                "volume": "128x128x32 x 32 coils"})


def _pair_ms(P, A, x, stream, reps):
    """device milliseconds of one forward + adjoint pair of operator A (inputs
    and outputs resident, allocated before the timed pairs)"""
    xd = P.DeviceArray.from_host(np.asarray(x, dtype=np.complex64).ravel())
    yd = P.DeviceArray(A.n_out)
    xo = P.DeviceArray(A.n_in)
    A.forward(xd, out=yd, stream=stream)
    A.adjoint(yd, out=xo, stream=stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        A.forward(xd, out=yd, stream=stream)
        A.adjoint(yd, out=xo, stream=stream)
    stream.synchronize()
    return 1e3 * (time.perf_counter() - t0) / reps


def _cpu_pair_ms(O, A, x, threads):
    """oracle (CPU restatement, FP32) milliseconds of one forward + adjoint pair"""
    O.set_threads(threads)
    t0 = time.perf_counter()
    y = A * x.astype(np.complex64)
    A / y
    return 1e3 * (time.perf_counter() - t0)


def rows_rates(P, S, stream, with_cpu=True):
    """The operators built beyond the headline path, each timed as a forward +
    adjoint pair on the device, with its roofline figure and the oracle's FP32
    restatement of the reference's CPU loops timed beside it (one pair, every
    host thread of this process):
      pcSenseTimeSeg: 128^2, 4 shots x 8192 samples, 8 coils, L=8 (gridding
        bytes of the 4 per-shot batched launches, each direction);
      pcSENSE (exact DFT per shot): 64^2, 4 shots x 1024, 8 coils (16 flops
        per sample-voxel pair, coil and direction);
      GdftR2: 32x32x8, 4096 samples, with SENSE over 8 coils (the same count
        plus the three sinc factors, not counted);
      min-max TimeSegmentation, exact (interptype 2) and histogram (3): the
        weight set-up at the 128^2, 16x2048, L=8 shape (the 16 interleaves
        share one readout's times, shots=16), and the SENSE pair over 8 coils,
        batched (complex weights in the fused kernels) and plain."""
    from oracle import oracle as O
    th = host_cpu()[0]
    out = {}
    # pcSenseTimeSeg
    N, Ns, nd, nc, L = 128, 4, 8192, 8, 8
    kx, ky, kz = S.spiral(N, 16, 2048)
    t = S.readout_times(Ns, nd, dwell=12.5e-6 / 4)
    fm = S.field_map(N, peak_hz=120.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    rng = np.random.default_rng(1)
    ph = rng.uniform(-0.5, 0.5, (Ns, N * N)).astype(np.float32)
    x = S.ellipse_phantom(N, seed=1)
    A = P.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens, fm, ph)
    ms = _pair_ms(P, A, x, stream, 10)
    A.set_mode(P.MODE_PLAIN)
    ms_plain = _pair_ms(P, A, x, stream, 2)
    gb = 2 * Ns * gridding_bytes(nd, 4 * N * N, nc * L, 2)
    e = dict(workload="128^2, 4 shots x 8192, 8 coils, L=8", ms_per_pair=round(ms, 3),
             ms_per_pair_plain=round(ms_plain, 3),
             gridding_GBs=round(gb / (ms * 1e-3) / 1e9, 1))
    if with_cpu:
        Oa = O.PcSenseTimeSeg(kx, ky, kz, N, N, 1, nc, t, L, 1, sens.ravel(), fm, ph.ravel())
        e["cpu_ms_per_pair"] = round(_cpu_pair_ms(O, Oa, x, th), 1)
        e["cpu_threads"] = th
    out["pcSenseTimeSeg"] = e
    # pcSENSE over the exact DFT
    N, Ns, nd, nc = 64, 4, 1024, 8
    kx, ky, kz = S.spiral(N, Ns, nd)
    t = S.readout_times(Ns, nd, dwell=12.5e-6)
    fm = S.field_map(N, peak_hz=120.0).astype(np.float32)
    sens = S.coil_maps(N, nc).astype(np.complex64)
    ph = rng.uniform(-0.5, 0.5, (Ns, N * N)).astype(np.float32)
    x = S.ellipse_phantom(N, seed=2)
    A = P.PcSENSE(kx, ky, kz, N, N, 1, nc, t, sens, fm, ph)
    ms = _pair_ms(P, A, x, stream, 5)
    tf = 16.0 * 2 * Ns * nd * N * N * nc / (ms * 1e-3) / 1e12
    e = dict(workload="64^2, 4 shots x 1024, 8 coils", ms_per_pair=round(ms, 3),
             achieved=round(tf, 2), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
             frac=round(tf / FP32_PEAK_TFLOPS, 4))
    if with_cpu:
        Oa = O.PcSENSE(kx, ky, kz, N, N, 1, nc, t, sens.ravel(), fm, ph.ravel())
        e["cpu_ms_per_pair"] = round(_cpu_pair_ms(O, Oa, x, th), 1)
    out["pcSENSE_gdft"] = e
    # GdftR2 through SENSE
    N, Nz, nc = 32, 8, 8
    kx, ky, kz = S.spiral(N, 2, 256, Nz=Nz)
    M = len(kx)
    t = np.tile(S.readout_times(2, 256, dwell=1e-5), Nz).astype(np.float32)
    ix, iy, iz = S.image_coords(N, N, Nz)
    fm = S.field_map(N, Nz=Nz, peak_hz=150.0).astype(np.float32)
    sens = S.coil_maps(N, nc, Nz=Nz).astype(np.complex64)
    x = S.ellipse_phantom(N, Nz=Nz, seed=3)
    D = P.GdftR2(M, N * N * Nz, kx, ky, kz, ix, iy, iz, fm, t, N, N, Nz)
    A = P.SENSE(D, sens, M, N * N * Nz, nc)
    ms = _pair_ms(P, A, x, stream, 3)
    tf = 16.0 * 2 * M * N * N * Nz * nc / (ms * 1e-3) / 1e12
    e = dict(workload="32x32x8, 4096 samples, 8 coils", ms_per_pair=round(ms, 3),
             achieved=round(tf, 2), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
             frac=round(tf / FP32_PEAK_TFLOPS, 4))
    if with_cpu:
        Od = O.GdftR2(M, N * N * Nz, kx, ky, kz, ix, iy, iz, fm, t, N, N, Nz)
        Oa = O.SENSE(Od, sens.ravel(), M, N * N * Nz, nc)
        e["cpu_ms_per_pair"] = round(_cpu_pair_ms(O, Oa, x, th), 1)
    out["GdftR2_sense"] = e
    # min-max time segmentation, exact (2) and histogram (3)
    prob = S.config_sense2d(seed=0)
    G = P.Gnufft(prob.M, 2.0, 128, 128, 1, prob.kx, prob.ky, prob.kz)
    if with_cpu:
        O.set_threads(th)
        Go = O.Gnufft(prob.M, 2.0, 128, 128, 1, prob.kx, prob.ky, prob.kz)
    for itype, key in ((2, "ts_minmax"), (3, "ts_histogram")):
        t0 = time.perf_counter()
        T = P.TimeSegmentation(G, prob.fm, prob.t, prob.M, prob.N, prob.L, interptype=itype,
                               shots=16)
        setup = time.perf_counter() - t0
        Sg = P.SENSE(T, prob.sens, prob.M, prob.N, prob.nc)
        ms = _pair_ms(P, Sg, prob.x_true, stream, 10)
        T.set_mode(P.MODE_PLAIN)
        Sg.set_mode(P.MODE_PLAIN)
        ms_plain = _pair_ms(P, Sg, prob.x_true, stream, 2)
        e = dict(workload="SENSE 128^2, 16x2048, 8 coils, L=8", setup_ms=round(1e3 * setup, 1),
                 ms_per_pair=round(ms, 3), ms_per_pair_plain=round(ms_plain, 3))
        if with_cpu:
            t0 = time.perf_counter()
            To = O.TimeSegmentation(Go, prob.fm, prob.t, prob.M, prob.N, prob.L,
                                    interptype=itype, shots=16)
            e["cpu_setup_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            So = O.SENSE(To, prob.sens.ravel(), prob.M, prob.N, prob.nc)
            e["cpu_ms_per_pair"] = round(_cpu_pair_ms(O, So, prob.x_true, th), 1)
        out[key] = e
    O.set_threads(1)
    return out


def serving_requests(P, S, n, seed):
    """n requests of the serving mix: 70% 2-D (the 2-D SENSE shape: 128^2,
    spiral 16x2048, L=8, with 8, 16 or 32 coils), 25% 3-D with L=8 and 5% 3-D
    with L=16, the 3-D ones 64x64x16 (16 partitions x 4 interleaves x 1024)
    or 128x128x32 (32 partitions x 8 x 2048) with 8 coils; each to a relative
    gradient tolerance of 1e-4 or 30 iterations.  The k-space of each shape
    comes from the device forward operator plus 40 dB noise, scaled to unit
    peak (the solve is invariant to the data scale, and the reference's 1e25
    bound on the line-search denominator then never ends a solve early)."""
    from powergrid_amd import serve as SV
    rng = np.random.default_rng(seed)
    kinds = rng.choice(3, size=n, p=[0.70, 0.25, 0.05])
    coils2d = rng.choice([8, 16, 32], size=n)
    big3d = rng.random(n) < 0.5
    reqs, cache = [], {}
    for j in range(n):
        if kinds[j] == 0:
            key = ("2d", int(coils2d[j]), 8)
            N, Nz, nc, L, nsh, ns = 128, 1, int(coils2d[j]), 8, 16, 2048
        else:
            L = 8 if kinds[j] == 1 else 16
            if big3d[j]:
                key = ("3d128x128x32", 8, L)
                N, Nz, nc, nsh, ns = 128, 32, 8, 8, 2048
            else:
                key = ("3d64x64x16", 8, L)
                N, Nz, nc, nsh, ns = 64, 16, 8, 4, 1024
        if key not in cache:
            pr = S.make_problem(N, nsh, ns, nc, L, Nz=Nz, peak_hz=120.0, seed=seed + len(cache))
            M = len(pr.kx)
            G = P.Gnufft(M, 2.0, N, N, Nz, pr.kx, pr.ky, pr.kz)
            A = P.SENSE(P.TimeSegmentation(G, pr.fm, pr.t, M, pr.N, L), pr.sens, M, pr.N, nc)
            y = S.add_noise(A * pr.x_true, 40.0, seed=seed + 100 + len(cache))
            y = (y / np.abs(y).max()).astype(np.complex64)
            cache[key] = (pr, y)
            del A, G
        pr, y = cache[key]
        reqs.append(SV.ReconRequest(pr.kx, pr.ky, pr.kz, pr.t, pr.Nx, pr.Ny, pr.Nz, pr.nc, pr.L,
                                    pr.fm, pr.sens, y, beta=1e4, niter=30, tol=1e-4,
                                    tag=(j, key)))
    return reqs


def serving_rates(P, S, rates=(20.0, 100.0, 200.0), n=96, slots=4, seed=11):
    """The serving run at each arrival rate: n requests of serving_requests()
    with Poisson arrivals into powergrid_amd.serve.ReconServer.  Latency runs
    from submission to the image on the host; the request's operator set-up
    is inside it.  Above the server's capacity the queue grows and latency
    with it; recons/s is then the capacity."""
    from powergrid_amd import serve as SV
    reqs = serving_requests(P, S, n, seed)
    rng = np.random.default_rng(seed + 1)
    out = {"workload": f"{n} requests per rate: 70% 2-D 128^2 spiral 16x2048 L=8 with 8/16/32 "
                       f"coils, 25% 3-D L=8 and 5% 3-D L=16 (64x64x16 or 128x128x32, 8 coils); "
                       f"to tol 1e-4 or 30 PCG iterations; {slots} slots",
           "by_rate": {}}
    for rate in rates:
        gaps = rng.exponential(1.0 / rate, size=n)
        t_start = time.perf_counter()
        with SV.ReconServer(slots=slots) as srv:
            futs = []
            t_next = t_start
            for r, g in zip(reqs, gaps):
                t_next += g
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                futs.append(srv.submit(r))
            outs = [f.result() for f in futs]
        t_end = max(o.finished for o in outs)
        lat = np.array([o.latency for o in outs]) * 1e3
        its = np.array([o.iterations for o in outs])
        per = {}
        for o in outs:
            k = o.tag[1]
            per.setdefault(f"{k[0]} {k[1]} coils L={k[2]}", []).append(o.latency * 1e3)
        out["by_rate"][f"{rate:g}/s"] = dict(
            recons_per_s=round(n / (t_end - t_start), 2),
            latency_ms_p50=round(float(np.percentile(lat, 50)), 2),
            latency_ms_p99=round(float(np.percentile(lat, 99)), 2),
            mean_iterations=round(float(its.mean()), 1),
            exit_reasons=dict(collections.Counter(o.exit_reason for o in outs)),
            max_concurrent=srv.max_running,
            latency_ms_p50_by_kind={k: round(float(np.median(v)), 2) for k, v in per.items()})
        log(f"[bench] serving {rate:g}/s: {out['by_rate'][f'{rate:g}/s']}")
    return out


def autotune_rate(P, prob, yd, stream, niter, steps=10):
    """The headline solve with pg_set_tuning("autotune", 1): the plan times its
    interpolation and spread variants at first use and keeps the faster."""
    P.set_tuning("autotune", 1)
    try:
        G, T, Sg, R = build_ops(P, prob, P.MODE_BATCHED)
        for _ in range(2):
            P.reconSolve(yd, Sg, R, niter=niter, mode=P.CG_GRAPH, stream=stream)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            x, info = P.reconSolve(yd, Sg, R, niter=niter, mode=P.CG_GRAPH, stream=stream)
        stream.synchronize()
        dt = time.perf_counter() - t0
        lds, mfma = P.op_tuning(Sg, prob.nc * prob.L)
    finally:
        P.set_tuning("autotune", 0)
    return dict(iters_per_s=round(steps * info["iterations"] / dt, 1),
                interp={0: "gather", 1: "lds tiles, bins of 128", 2: "lds tiles, 4 segments per stage, bins of 64"}.get(lds, "?"),
                spread="matrix cores" if mfma == 1 else "vector units")


def batch_rates(P, S, stream, K=96, niter=20, reps=2, seed=21, nrmse_slices=True):
    """Uneven reconstructions in one call (pg_recon_solve_batch: each slice's
    captured solve graph, launched concurrently on the library's batch
    streams): K slices with matrix sizes from {96, 128, 192}^2, 8/16/32
    coils, L in {4, 8, 12}, field maps of 50-250 Hz peak and spirals of 16
    interleaves x 512..4096 samples (8k-64k), each solved for niter PCG
    iterations (QuadPenalty 1e4, data scaled to unit peak), against the same
    solves run one after another, and against the plain form (one NUFFT at a
    time, host-driven PCG, one slice after another).  k-space is each slice's
    FP64 exact NUDFT (on the GPU) plus 40 dB noise; every slice's image is
    compared with the FP64 exact-NUDFT PCG of its data (Gdft in FP64 through
    SENSE as the operator)."""
    rng = np.random.default_rng(seed)
    items, keep, specs, grids = [], [], [], 0
    for k in range(K):
        N = int(rng.choice([96, 128, 192]))
        nc = int(rng.choice([8, 16, 32]))
        L = int(rng.choice([4, 8, 12]))
        ns = int(rng.choice([512, 1024, 2048, 4096]))
        pk = float(rng.uniform(50, 250))
        pr = S.make_problem(N, 16, ns, nc, L, peak_hz=pk, seed=seed + k)
        M = len(pr.kx)
        D = P.Gdft(M, pr.N, pr.kx, pr.ky, pr.kz, pr.ix, pr.iy, pr.iz, pr.fm, pr.t, precision=1)
        Sd = P.SENSE(D, pr.sens, M, pr.N, nc)
        xd = P.DeviceArray.from_host(pr.x_true.astype(np.complex64))
        y = Sd.forward(xd, stream=stream).to_host(stream).astype(np.complex128)
        y = S.add_noise(y, 40.0, seed=seed + 100 + k)
        y = P.DeviceArray.from_host((y / np.abs(y).max()).astype(np.complex64))
        G = P.Gnufft(M, 2.0, N, N, 1, pr.kx, pr.ky, pr.kz)
        T = P.TimeSegmentation(G, pr.fm, pr.t, M, pr.N, L)
        A = P.SENSE(T, pr.sens, M, pr.N, nc)
        R = P.QuadPenalty(N, N, 1, 1e4)
        items.append((A, y, R))
        keep.append((G, T, D, Sd, pr))
        # the segment length tau = readout / L and the largest phase a
        # segment spans, 2 pi max|fm| tau: the Hanning time segmentation is
        # an accurate model of the field only while it stays well below pi
        tau = float(pr.t.max() - pr.t.min()) / L
        specs.append(dict(N=N, nc=nc, L=L, M=M, peak_hz=round(pk, 1), tau_ms=round(1e3 * tau, 2),
                          segment_phase_rad=round(float(np.abs(pr.fm).max()) * tau, 2)))
        grids += nc * L
    P.recon_solve_batch(items, niter, stream=stream)  # capture
    for A, y, R in items:
        P.reconSolve(y, A, R, niter=niter, stream=stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = P.recon_solve_batch(items, niter, stream=stream)
    stream.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    # the same call with the solves spread over 1, 2 and 8 of the library's
    # streams (4 is the default)
    by_streams = {4: round(1e3 * t_batch, 2)}
    for nst in (1, 2, 8):
        P.set_tuning("batch_streams", nst)
        try:
            P.recon_solve_batch(items, niter, stream=stream)
            stream.synchronize()
            t0 = time.perf_counter()
            P.recon_solve_batch(items, niter, stream=stream)
            stream.synchronize()
            by_streams[nst] = round(1e3 * (time.perf_counter() - t0), 2)
        finally:
            P.set_tuning("batch_streams", 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        for A, y, R in items:
            P.reconSolve(y, A, R, niter=niter, stream=stream)
    stream.synchronize()
    t_seq = (time.perf_counter() - t0) / reps
    # the plain form, one slice after another
    t_plain = 0.0
    for (A, y, R), (G, T, D, Sd, pr) in zip(items, keep):
        for op in (G, T, A):
            op.set_mode(P.MODE_PLAIN)
        P.reconSolve(y, A, R, niter=1, mode=P.CG_HOST, stream=stream)
        stream.synchronize()
        t0 = time.perf_counter()
        P.reconSolve(y, A, R, niter=niter, mode=P.CG_HOST, stream=stream)
        stream.synchronize()
        t_plain += time.perf_counter() - t0
        for op in (G, T, A):
            op.set_mode(P.MODE_BATCHED)
    its = sum(info["iterations"] for _, info in out)
    res = dict(workload=f"{K} slices, 96-192^2, 8-32 coils, L 4-12, 8k-64k samples, 50-250 Hz, "
                        f"{niter} PCG iterations each ({grids} grids in all)",
               ms_one_call=round(1e3 * t_batch, 2), ms_one_by_one=round(1e3 * t_seq, 2),
               ms_plain_one_by_one=round(1e3 * t_plain, 1),
               ms_one_call_by_streams=by_streams,
               cg_iters_per_s=round(its / t_batch, 1),
               speedup_vs_one_by_one=round(t_seq / t_batch, 3),
               speedup_vs_plain=round(t_plain / t_batch, 1))
    if nrmse_slices:
        per = []
        for (x, info), (A, y, R), (G, T, D, Sd, pr), sp in zip(out, items, keep, specs):
            xe, ie = P.reconSolve(y, Sd, R, niter=niter, mode=P.CG_GRAPH, stream=stream)
            per.append(dict(sp, nrmse_vs_fp64_exact_nudft_cg=round(nrmse(x.to_host(),
                                                                         xe.to_host()), 5),
                            exact_cg_iterations=ie["iterations"]))
        e = np.array([p["nrmse_vs_fp64_exact_nudft_cg"] for p in per])
        ph = np.array([p["segment_phase_rad"] for p in per])
        res["nrmse_vs_fp64_exact_nudft_cg"] = dict(
            median=round(float(np.median(e)), 5), max=round(float(e.max()), 5),
            median_segment_phase_below_1_rad=round(float(np.median(e[ph < 1.0])), 5)
            if np.any(ph < 1.0) else None,
            median_segment_phase_above_pi=round(float(np.median(e[ph > np.pi])), 5)
            if np.any(ph > np.pi) else None)
        res["per_slice"] = per
    return res


def pmc_traffic(summary, kernel):
    """HBM bytes per launch of `kernel` from a committed FETCH_SIZE/WRITE_SIZE
    summary under profiles/ (counters scaled by a device copy of known size),
    or None when it holds no row for it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 7 and cells[0] == kernel:
            try:
                return float(cells[6]) * 1e6
            except ValueError:
                return None
    return None


def timed_solves(P, yd, Sg, R, niter, steps, warmup, stream):
    """`warmup` untimed and `steps` timed reconstructions (whole-solve graph);
    returns (it/s, seconds, per-solve latencies, x, info)"""
    for _ in range(warmup):
        x, info = P.reconSolve(yd, Sg, R, niter=niter, mode=P.CG_GRAPH, stream=stream)
    stream.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        x, info = P.reconSolve(yd, Sg, R, niter=niter, mode=P.CG_GRAPH, stream=stream)
        lat.append(time.perf_counter() - ts)
    stream.synchronize()
    dt = time.perf_counter() - t0
    return steps * info["iterations"] / dt, dt, lat, x, info


def kernel_times(P, Sg, x_host, stream, nrep, M, Gsz, B, nc, L, d, plane_g=None):
    """Per-kernel durations of the batched operator (HIP events the library
    records on its launch stream, nrep forward + adjoint pairs), each with its
    algorithmic bytes and fraction of the HBM peak.  plane_g: on z-class
    planes, the points per grid that the FFT passes transform (nzc*Gx*Gy
    instead of the full grid Gsz), whose bytes and fraction are added."""
    P.profile.enable(True)
    P.profile.reset()
    xd = P.DeviceArray.from_host(np.asarray(x_host, np.complex64))
    for _ in range(nrep):
        ad = Sg.forward(xd, stream=stream)
        Sg.adjoint(ad, stream=stream)
    stream.synchronize()
    kern = {}
    for name in ("fwd_prep_fft", "interp", "spread_values", "spread", "adj_fft_post"):
        n, ms = P.profile.query(name)
        kern[name] = ms / max(n, 1)
    P.profile.enable(False)
    gb = gridding_bytes(M, Gsz, B, d)
    fb = fft_bytes(Gsz, B)
    # per-sample spread values: B complex written, plus the coil data and AA read
    vb = 8 * B * M + 8 * nc * M + 4 * L * M
    kinfo = {
        "interp": dict(ms=kern["interp"], bytes=gb),
        "spread": dict(ms=kern["spread"], bytes=gb),
        "spread_values": dict(ms=kern["spread_values"], bytes=vb),
        "fwd_prep_fft": dict(ms=kern["fwd_prep_fft"], bytes=fb),
        "adj_fft_post": dict(ms=kern["adj_fft_post"], bytes=fb),
    }
    for k, v in kinfo.items():
        v["GBs"] = v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0
        v["frac"] = v["GBs"] / HBM_PEAK_GBS
        extra = ""
        if plane_g and k in ("fwd_prep_fft", "adj_fft_post") and v["ms"] > 0:
            v["plane_bytes"] = fft_bytes(plane_g, B)
            v["plane_frac"] = v["plane_bytes"] / (v["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
            extra = f"; {100 * v['plane_frac']:.1f}% by the class planes' bytes"
        log(f"[bench] {k:13s} {v['ms'] * 1e3:9.1f} us  {v['GBs']:8.1f} GB/s  "
            f"({100 * v['frac']:.1f}% of HBM{extra})")
    return kinfo


def roofline(kinfo, L, summary, plane_bytes=None):
    """the roofline object of the dominant gridding kernel.  The interpolation
    stages all L segments of a tile (k_interp_lds) below 4 segments and four
    at a time (k_interp_lds4) from 4 up; the spread of a time-segmented batch
    stages compact sample rows in LDS (k_spread_patch_st<1, 32>, or <1, 48>
    for row slices wider than 32 floats).  plane_bytes: on z-class planes the
    launch moves 4*2*M + B*(8*M + 8*nzc*Gx*Gy) bytes (its x-y grids are the
    nzc planes, the z taps having been applied in the z stage), reported
    beside the full-grid figure of 4*d*M + B*(8*M + 8*G)."""
    dom = max(("interp", "spread"), key=lambda k: kinfo[k]["ms"])
    kname = {"interp": "k_interp_lds4<false>" if L >= 4 else "k_interp_lds<8, false>",
             "spread": "k_spread_patch_st<1, 32>"}[dom]
    tr = pmc_traffic(summary, kname)
    if tr is None and dom == "spread":  # row slices wider than 32 floats
        kname = "k_spread_patch_st<1, 48>"
        tr = pmc_traffic(summary, kname)
    out = dict(bound="hbm", achieved=round(kinfo[dom]["GBs"], 1), peak=HBM_PEAK_GBS,
               unit="GB/s", frac=round(kinfo[dom]["frac"], 4),
               traffic=None if tr is None else round(tr), kernel=kname,
               algorithmic_bytes=kinfo[dom]["bytes"], avg_launch_ms=round(kinfo[dom]["ms"], 5))
    if plane_bytes:
        gbs = plane_bytes / (kinfo[dom]["ms"] * 1e-3) / 1e9
        out["plane_bytes"] = plane_bytes
        out["achieved_plane_bytes"] = round(gbs, 1)
        out["frac_plane_bytes"] = round(gbs / HBM_PEAK_GBS, 4)
    return out


def nrmse(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def sense2d_section(P, S, args, stream):
    """The 2-D field-corrected SENSE workload: batched and plain rates,
    kernels, autotuner, NRMSE against the FP64 same-operator PCG (at beta 1e4
    and 1e-3) and the FP64 exact-NUDFT PCG, CPU baseline."""
    prob = S.config_sense2d(seed=0)
    if args.beta is not None:
        prob.beta = args.beta
    niter = prob.niter
    log(f"[bench] 2-D workload 128x128 spiral 16x2048 nc={prob.nc} L={prob.L} niter={niter}")
    y = synth_data(P, prob)
    yd = P.DeviceArray.from_host(y)
    G, T, Sg, R = build_ops(P, prob, P.MODE_BATCHED)
    value, dt, lat, x, info = timed_solves(P, yd, Sg, R, niter, args.sense2d_steps, 3, stream)
    x_gpu = x.to_host()
    log(f"[bench] 2-D batched: {value:.1f} it/s, {1e3 * dt / args.sense2d_steps:.2f} ms/recon")
    tuned = autotune_rate(P, prob, yd, stream, niter)
    log(f"[bench] 2-D autotuned: {tuned}")
    M, Gsz, B = prob.M, 4 * prob.N, prob.nc * prob.L
    kinfo = kernel_times(P, Sg, x_gpu, stream, 20, M, Gsz, B, prob.nc, prob.L, 2)
    G2, T2, Sg2, R2 = build_ops(P, prob, P.MODE_PLAIN)
    P.reconSolve(yd, Sg2, R2, niter=2, mode=P.CG_HOST, stream=stream)
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        xp, infop = P.reconSolve(yd, Sg2, R2, niter=niter, mode=P.CG_HOST, stream=stream)
    stream.synchronize()
    plain = args.plain_steps * infop["iterations"] / (time.perf_counter() - t0)
    log(f"[bench] 2-D plain: {plain:.1f} it/s")
    out = {
        "workload": "2D field-corrected SENSE 128x128, spiral 16x2048, 8 coils, L=8 Hanning, "
                    f"20 PCG iterations, QuadPenalty {prob.beta:g}, W=4, os=2",
        "iters_per_s": round(value, 2), "ms_per_recon": round(1e3 * dt / args.sense2d_steps, 3),
        "plain_iters_per_s": round(plain, 2), "speedup_vs_plain": round(value / plain, 2),
        "recon_latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                             "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "nrmse_plain_vs_batched": nrmse(xp.to_host(), x_gpu),
        "kernels_us": {k: round(v["ms"] * 1e3, 2) for k, v in kinfo.items()},
        "autotuned": tuned,
        "roofline": roofline(kinfo, prob.L, "sense2d_pmc_traffic.md"),
        "grids_per_launch": B,
    }
    if not args.no_oracle:
        t0 = time.perf_counter()
        x64 = fp64_oracle_recon(prob, y, niter)
        out["nrmse_vs_fp64_oracle"] = nrmse(x_gpu, x64)
        out["nrmse_vs_phantom"] = nrmse(x_gpu, prob.x_true)
        log(f"[bench] 2-D fp64 oracle recon {time.perf_counter() - t0:.1f}s, "
            f"nrmse {out['nrmse_vs_fp64_oracle']:.2e}")
        xe, _ = exact_cg(P, prob, yd, R, niter, stream)
        out["nrmse_vs_fp64_exact_nudft_cg"] = nrmse(x_gpu, xe)
        out["nrmse_fp64_oracle_vs_exact_nudft_cg"] = nrmse(x64, xe)
        if args.beta is None:
            # the same workload at the configuration's nearly unregularised
            # beta, where the 20-iteration PCG amplifies FP32 rounding
            lowb = 1e-3
            R_low = P.QuadPenalty(prob.Nx, prob.Ny, prob.Nz, lowb)
            xl, _ = P.reconSolve(yd, Sg, R_low, niter=niter, mode=P.CG_GRAPH, stream=stream)
            xl = xl.to_host()
            beta_keep, prob.beta = prob.beta, lowb
            xl64 = fp64_oracle_recon(prob, y, niter)
            prob.beta = beta_keep
            out["nrmse_vs_fp64_oracle_beta_1e-3"] = nrmse(xl, xl64)
            xle, _ = exact_cg(P, prob, yd, R_low, niter, stream)
            out["nrmse_vs_fp64_exact_nudft_cg_beta_1e-3"] = nrmse(xl, xle)
        out["nrmse_vs_fp64_exact_nudft_cg_by_segments"] = ts_model_error(P, prob, yd, xe, stream)
        log(f"[bench] 2-D time-segmentation model error {out['nrmse_vs_fp64_exact_nudft_cg_by_segments']}")
    if not args.no_cpu:
        cb = cpu_baseline(prob, y)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = round(value / cb["value"], 1)
        log(f"[bench] 2-D cpu baseline {cb['value']:.3f} it/s on {cb['cores']} threads")
    extras = dict(prob=prob, yd=yd, Sg=Sg)
    return out, extras


def plain2d_section(P, S, args, stream):
    """The plain 2-D workload: 64x64 ellipse phantom, spiral 8x512, 1 coil
    (unit sensitivity), no field map, KB width 6 oversampled points
    (kernelWidth 3.0, beta by the reference's formula), 10 PCG iterations,
    QuadPenalty beta 1e-3; k-space = FP64 exact NUDFT (on the GPU) + 40 dB
    noise.  GPU rates (whole-solve graph; plain: host-driven loop), and the
    reconstruction against the reference's own FP32 CG image (the oracle's
    FP32 restatement), the FP64 same-operator PCG and the FP64 exact-NUDFT
    PCG.  At beta 1e-3 the 10 iterations amplify FP32 rounding (the FP32
    restatement itself lands ~1e-2 from FP64); the same solve at beta 1e4 is
    reported beside it."""
    from oracle import oracle as O
    prob = S.make_problem(64, 8, 512, 1, 1, field=False, niter=10, beta=1e-3, kernel_width=3.0,
                          seed=0)
    N, M = prob.N, prob.M
    ones = np.ones(N, np.complex64)
    D = P.Gdft(M, N, prob.kx, prob.ky, prob.kz, prob.ix, prob.iy, prob.iz, np.zeros(N), prob.t,
               precision=1)
    y0 = (D * prob.x_true.astype(np.complex64)).astype(np.complex128)
    y = S.add_noise(y0, 40.0, seed=11).astype(np.complex64)
    yd = P.DeviceArray.from_host(y)
    G = P.Gnufft(M, 2.0, 64, 64, 1, prob.kx, prob.ky, prob.kz, kernel_width=3.0)
    Sg = P.SENSE(G, ones, M, N, 1)
    out = {"workload": "plain 2D Gnufft 64x64, spiral 8x512, 1 coil, no field map, KB width 6 "
                       "(kernelWidth 3.0), 10 PCG iterations, QuadPenalty 1e-3"}
    res = {}
    for beta in (1e-3, 1e4):
        R = P.QuadPenalty(64, 64, 1, beta)
        value, dt, lat, x, info = timed_solves(P, yd, Sg, R, prob.niter, 50, 3, stream)
        xg = x.to_host()
        e = {"iters_per_s": round(value, 1),
             "recon_latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                                  "p99": round(1e3 * float(np.percentile(lat, 99)), 3)}}
        if not args.no_oracle:
            O.set_threads(host_cpu()[0])
            ref = {}
            for prec in "fd":
                Go = O.Gnufft(M, 2.0, 64, 64, 1, prob.kx, prob.ky, prob.kz, 3.0, prec=prec)
                So = O.SENSE(Go, ones, M, N, 1)
                Ro = O.QuadPenalty(64, 64, 1, beta, prec=prec)
                yy = y if prec == "f" else y.astype(np.complex128)
                ref[prec] = O.reconSolve(yy, So, Ro, 64, 64, 1, prob.niter)[0]
            xe, _ = exact_cg_op(P, D, ones, prob, yd, R, prob.niter, stream)
            e.update(nrmse_vs_fp32_reference_cg=nrmse(xg, ref["f"]),
                     nrmse_vs_fp64_oracle=nrmse(xg, ref["d"]),
                     nrmse_fp32_reference_vs_fp64_oracle=nrmse(ref["f"], ref["d"]),
                     nrmse_vs_fp64_exact_nudft_cg=nrmse(xg, xe),
                     nrmse_vs_phantom=nrmse(xg, prob.x_true))
        res[beta] = (e, R)
    e, R = res[1e-3]
    out.update(e)
    out["beta_1e4"] = res[1e4][0]
    G2 = P.Gnufft(M, 2.0, 64, 64, 1, prob.kx, prob.ky, prob.kz, kernel_width=3.0)
    S2 = P.SENSE(G2, ones, M, N, 1)
    for op in (G2, S2):
        op.set_mode(P.MODE_PLAIN)
    P.reconSolve(yd, S2, R, niter=2, mode=P.CG_HOST, stream=stream)
    t0 = time.perf_counter()
    for _ in range(3):
        _, infop = P.reconSolve(yd, S2, R, niter=prob.niter, mode=P.CG_HOST, stream=stream)
    stream.synchronize()
    out["plain_iters_per_s"] = round(3 * infop["iterations"] / (time.perf_counter() - t0), 1)
    out["speedup_vs_plain"] = round(out["iters_per_s"] / out["plain_iters_per_s"], 2)
    if not args.no_cpu:
        threads, model = host_cpu()
        O.set_threads(threads)
        Go = O.Gnufft(M, 2.0, 64, 64, 1, prob.kx, prob.ky, prob.kz, 3.0)
        So = O.SENSE(Go, ones, M, N, 1)
        Ro = O.QuadPenalty(64, 64, 1, 1e-3)
        O.reconSolve(y, So, Ro, 64, 64, 1, 1)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < 3.0:
            O.reconSolve(y, So, Ro, 64, 64, 1, prob.niter)
            n += prob.niter
        cpu = n / (time.perf_counter() - t0)
        O.set_threads(1)
        out["cpu_baseline"] = dict(value=round(cpu, 2), unit="CG iterations/s", cores=threads,
                                   kind="port", cpu_model=model, nproc=os.cpu_count(),
                                   sample=f"{n} PCG iterations in solves of {prob.niter} "
                                          f"(FP32, {threads} OpenMP threads)")
        out["speedup_vs_cpu"] = round(out["iters_per_s"] / cpu, 1)
    return out


def exact_cg_op(P, D, sens, prob, yd, R, niter, stream):
    """the PCG of `exact_cg` over an existing FP64 Gdft D through SENSE"""
    Sd = P.SENSE(D, sens, prob.M, prob.N, len(sens) // prob.N)
    x, info = P.reconSolve(yd, Sd, R, niter=niter, mode=P.CG_GRAPH, stream=stream)
    return x.to_host(), info


def ts_model_error(P, prob, yd, xe, stream, Ls=(8, 16, 32)):
    """The 2-D SENSE reconstruction (beta as the workload) with L Hanning
    segments against xe, the FP64 exact-NUDFT PCG of the same data: the
    time-segmentation model's own error, which falls as L grows."""
    R = P.QuadPenalty(prob.Nx, prob.Ny, prob.Nz, prob.beta)
    out = {}
    for L in Ls:
        G = P.Gnufft(prob.M, 2.0, prob.Nx, prob.Ny, prob.Nz, prob.kx, prob.ky, prob.kz)
        T = P.TimeSegmentation(G, prob.fm, prob.t, prob.M, prob.N, L)
        Sg = P.SENSE(T, prob.sens, prob.M, prob.N, prob.nc)
        x, _ = P.reconSolve(yd, Sg, R, niter=prob.niter, mode=P.CG_GRAPH, stream=stream)
        out[f"L={L}"] = nrmse(x.to_host(), xe)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed reconstructions of the 3-D workload (15 PCG iterations each)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-oracle", action="store_true",
                    help="skip the FP64 oracle and exact-NUDFT reconstructions")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--plain-iters-3d", type=int, default=2,
                    help="PCG iterations of the timed plain 3-D solve")
    ap.add_argument("--sense2d-steps", type=int, default=20)
    ap.add_argument("--beta", type=float, default=None, help="2-D QuadPenalty weight override")
    ap.add_argument("--no-sense2d", action="store_true", help="skip the 2-D SENSE workload")
    ap.add_argument("--no-plain2d", action="store_true", help="skip the plain 2-D workload")
    ap.add_argument("--no-gdft", action="store_true", help="skip the Gdft roofline")
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the density-compensation / gridded-coil-image timings")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the uneven one-launch batch measurement")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving-loop measurement")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the pcSENSE / GdftR2 / min-max interpolator timings")
    ap.add_argument("--counters", type=int, default=0, metavar="REPS",
                    help="hardware-counter pass (for rocprofv3 --pmc): one 256 MiB device-to-"
                         "device copy (byte-counter calibration), then REPS forward+adjoint "
                         "pairs of the batched operator; no timing, no JSON line")
    ap.add_argument("--counters-2d", action="store_true",
                    help="the counter pass on the 2-D workload instead of the 3-D one")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="a library tuning switch set before any plan is made "
                         "(powergrid_amd.set_tuning); repeatable")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import powergrid_amd as P
    from powergrid_amd import synth as S
    for kv in args.tune:
        k, v = kv.split("=", 1)
        P.set_tuning(k, int(v))
        log(f"[bench] tuning {k}={int(v)}")
    stream = P.Stream()

    if args.counters:
        prob = S.config_sense2d(seed=0) if args.counters_2d else S.config_sos3d(seed=0)
        if args.counters_2d:
            _, _, Sg, _ = build_ops(P, prob, P.MODE_BATCHED)
        else:
            _, _, Sg, _ = build_ops_sos3d(P, prob, P.MODE_BATCHED)
        nbytes = 1 << 28
        src = P.DeviceArray(nbytes // 8)
        dst = P.DeviceArray(nbytes // 8)
        dst.copy_from(src, stream=stream)
        stream.synchronize()
        xd = P.DeviceArray.from_host(prob.x_true.astype(np.complex64))
        for _ in range(args.counters):
            ad = Sg.forward(xd, stream=stream)
            Sg.adjoint(ad, stream=stream)
        stream.synchronize()
        log(f"[bench] counter pass: copy {nbytes} B, {args.counters} forward+adjoint pairs")
        return

    # ---------------- headline: the 3-D stack-of-spirals workload ---------
    prob = S.config_sos3d(seed=0)
    niter = prob.niter
    log(f"[bench] 3-D workload {prob.Nx}x{prob.Ny}x{prob.Nz}, M={prob.M}, nc={prob.nc}, "
        f"L={prob.L}, TV, niter={niter}")
    t0 = time.perf_counter()
    G, T, Sg, R = build_ops_sos3d(P, prob, P.MODE_BATCHED)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    y = sos3d_data(P, prob, stream)
    t_data = time.perf_counter() - t0
    log(f"[bench] plan set-up {t_setup:.2f} s, exact NUDFT data {t_data:.2f} s, "
        f"z-class planes {P.zclasses(G)}")
    yd = P.DeviceArray.from_host(y)
    value, dt, lat, x, info = timed_solves(P, yd, Sg, R, niter, args.steps, args.warmup, stream)
    iters_done = info["iterations"]
    x_gpu = x.to_host()
    log(f"[bench] batched: {value:.2f} it/s, {1e3 * dt / args.steps:.1f} ms/recon, "
        f"exit={info['exit_reason']} iters={iters_done}")
    M, B = prob.M, prob.nc * prob.L
    Gsz = 8 * prob.N  # 256 x 256 x 64
    nzc = P.zclasses(G)
    kinfo = kernel_times(P, Sg, x_gpu, stream, 3, M, Gsz, B, prob.nc, prob.L, 3,
                         plane_g=nzc * 4 * prob.Nx * prob.Ny if nzc else None)
    plane_b = (4 * 2 * M + B * (8 * M + 8 * nzc * 4 * prob.Nx * prob.Ny)) if nzc else None
    roof = roofline(kinfo, prob.L, "sos3d_pmc_traffic.md", plane_b)

    # plain path: one NUFFT at a time, host-driven PCG (the reference's loop)
    G2, T2, Sg2, R2 = build_ops_sos3d(P, prob, P.MODE_PLAIN)
    P.reconSolve(yd, Sg2, R2, niter=1, mode=P.CG_HOST, stream=stream)
    stream.synchronize()
    t0 = time.perf_counter()
    xp, infop = P.reconSolve(yd, Sg2, R2, niter=args.plain_iters_3d, mode=P.CG_HOST,
                             stream=stream)
    stream.synchronize()
    plain = infop["iterations"] / (time.perf_counter() - t0)
    log(f"[bench] plain: {plain:.3f} it/s")
    del G2, T2, Sg2, R2
    # the batched solve stopped after the same iteration count, for the
    # plain-vs-batched distance
    xb_short, _ = P.reconSolve(yd, Sg, R, niter=args.plain_iters_3d, mode=P.CG_GRAPH,
                               stream=stream)

    out = {
        "metric": "SENSE-TS CG iters/s batched vs plain at NRMSE<=1e-3; p50/p99 recon latency",
        "value": round(value, 3),
        "unit": "CG iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3D stack-of-spirals SENSE 128x128x32, 32 partitions x 8 interleaves x "
                        "2048 samples, 32 coils, L=12 Hanning, TVPenalty beta 1e-3 delta 1e-2, "
                        "15 PCG iterations, W=4, os=2; data = exact field-corrected NUDFT "
                        "(FP64) + 40 dB noise, x 1e-3 (scanner units: at the operator's own "
                        "scale the reference's 1e25 line-search bound, "
                        "solve_pwls_pcg.hpp:137, ends the solve at once)",
            "step": "one reconstruction (15 PCG iterations), inputs resident in HBM",
            "cg_scalars": "f64", "grids_per_launch": B, "zclass_planes": P.zclasses(G),
        },
        "plain_iters_per_s": round(plain, 4),
        "speedup_vs_plain": round(value / plain, 1),
        "recon_latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                             "p99": round(1e3 * float(np.percentile(lat, 99)), 2)},
        "exit_reason": info["exit_reason"],
        "nrmse_plain_vs_batched": nrmse(xp.to_host(), xb_short.to_host()),
        "plain_vs_batched_iterations": args.plain_iters_3d,
        "nrmse_vs_phantom": nrmse(x_gpu / 1e-3, prob.x_true),
        "kernels_us": {k: round(v["ms"] * 1e3, 1) for k, v in kinfo.items()},
        "roofline": roof,
        "plan_setup_s": round(t_setup, 2),
    }
    if not args.no_oracle:
        t0 = time.perf_counter()
        xe, infe = exact_cg(P, prob, yd, R, niter, stream)
        out["nrmse_vs_fp64_exact_nudft_cg"] = nrmse(x_gpu, xe)
        log(f"[bench] exact-NUDFT PCG {time.perf_counter() - t0:.1f} s, nrmse "
            f"{out['nrmse_vs_fp64_exact_nudft_cg']:.2e}")
    if not args.no_cpu:
        t0 = time.perf_counter()
        cb = cpu_baseline_sos3d(prob)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = round(value / cb["value"], 1)
        log(f"[bench] cpu baseline {cb['value']:.5f} it/s on {cb['cores']} threads "
            f"({time.perf_counter() - t0:.1f} s)")
    del G, T, Sg, R, yd
    out["hbm_copy_GBs"] = round(copy_bandwidth(P, stream), 1)
    log(f"[bench] device copy {out['hbm_copy_GBs']} GB/s")

    # ---------------- the 2-D workload and the rows beside the path -------
    if not args.no_sense2d:
        out["sense2d"], ex = sense2d_section(P, S, args, stream)
        prob2 = ex["prob"]
        if not args.no_gdft:
            out["gdft_roofline"] = gdft_roofline(P, prob2, stream)
            log(f"[bench] gdft {out['gdft_roofline']}")
        if not args.no_direct:
            out["direct_recon"] = direct_recon_rates(P, prob2, stream)
            log(f"[bench] direct recon {out['direct_recon']}")
    if not args.no_plain2d:
        out["plain2d"] = plain2d_section(P, S, args, stream)
        log(f"[bench] plain 2-D {out['plain2d']}")
    if not args.no_batch:
        out["uneven_batch"] = batch_rates(P, S, stream)
        log(f"[bench] uneven batch {out['uneven_batch']}")
    if not args.no_serve:
        out["serving"] = serving_rates(P, S)
        log(f"[bench] serving {out['serving']}")
    if not args.no_rows:
        out["rows"] = rows_rates(P, S, stream, with_cpu=not args.no_cpu)
        log(f"[bench] rows {out['rows']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
