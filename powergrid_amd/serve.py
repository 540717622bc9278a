"""Continuous-batching reconstruction service on one GPU.

A ReconServer takes reconstruction requests -- field-corrected SENSE
(Gnufft + TimeSegmentation + SENSE) of any matrix size, coil count, segment
count and trajectory length, each with its own k-space, sensitivities and
field map -- and runs them as they arrive.  `slots` worker threads each own
a library stream.  A request is admitted, in arrival order, when a slot is
free and its estimated device footprint fits what is left of the HBM budget
(a fraction of the free device memory when the server starts).  An admitted
request builds its operators, uploads its inputs (they stay resident until
it finishes), runs its whole PCG as one graph launch on the slot's stream
(to a relative gradient tolerance or niter iterations), copies the image
back and releases its memory.  The graph launches of admitted requests
overlap on the device, and the operator set-up of one request (host work in
the library, run with the GIL released) overlaps the others' device work.

Every request has its own operators and workspaces, and every kernel of the
2-D and multi-grid 3-D paths sums in a fixed order, so a request's image does
not depend on what else is running.

There is no counterpart in the reference, whose drivers reconstruct one
dataset per process (PowerGridSenseMPI.cpp:198-320 hands slices to MPI
ranks).
"""
from __future__ import annotations

import collections
import ctypes as C
import hashlib
import math
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from . import (CG_GRAPH, SENSE, DeviceArray, Gnufft, QuadPenalty, Stream, TimeSegmentation,
               TVPenalty, solve_pwls_pcg)
from ._lib import check, lib


@dataclass
class ReconRequest:
    """One reconstruction: the arguments of reconSolve over
    SENSE(TimeSegmentation(Gnufft)) (reconSolve.h:63-82), plus the stopping
    rule.  Arrays are host arrays: kx, ky, kz, tvec of M samples; fieldmap
    of N = Nx*Ny*Nz voxels (rad/s); sens nc maps of N (coil-major); data
    M*nc (coil-major)."""
    kx: np.ndarray
    ky: np.ndarray
    kz: np.ndarray
    tvec: np.ndarray
    Nx: int
    Ny: int
    Nz: int
    nc: int
    L: int
    fieldmap: np.ndarray
    sens: np.ndarray
    data: np.ndarray
    beta: float = 1e4
    penalty: str = "quad"  # or "tv"
    delta: float = 1e-2
    niter: int = 30
    tol: float = 1e-4
    interptype: int = 1
    gridos: float = 2.0
    kernel_width: float = 4.0
    tag: Any = None

    @property
    def M(self):
        return int(np.asarray(self.kx).size)

    @property
    def N(self):
        return self.Nx * self.Ny * self.Nz


@dataclass
class ReconResult:
    image: np.ndarray
    iterations: int
    exit_reason: str
    submitted: float  # time.perf_counter() stamps
    started: float
    finished: float
    device_bytes: int
    tag: Any = None

    @property
    def latency(self):
        return self.finished - self.submitted


def mem_info():
    """(free, total) device memory in bytes"""
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().pg_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def plan_bytes(req: ReconRequest) -> int:
    """Device bytes of the request's Gnufft plan: its batched grids and FFT
    intermediates (B = nc*L grids of the oversampled size), per-sample spread
    values and tap tables, with a quarter in reserve."""
    d = 3 if req.Nz > 1 else 2
    G = 1
    for n in (req.Nx, req.Ny, req.Nz if d == 3 else 1):
        G *= int(math.ceil(req.gridos * n))
    B = req.nc * max(1, req.L)
    M = req.M
    taps = int(math.ceil(req.kernel_width * req.gridos))
    grids = 8 * B * G * 2  # grids and the FFT intermediates
    vals = 8 * B * M
    tables = M * (4 * d + 4 * d * taps + 64 * 4)
    return int(1.25 * (grids + vals + tables))


def estimate_bytes(req: ReconRequest) -> int:
    """Device bytes a request holds at its peak: its plan (plan_bytes), the
    data-space CG vectors, the segment weights, the image-side maps and the
    CG image vectors, with a quarter in reserve."""
    M, N = req.M, req.N
    data = M * req.nc * (8 + 16 + 8 + 8 + 8)  # y, Ax (FP64), A dir, residual, scratch
    image = N * (8 * (req.nc + 2 * max(1, req.L)) + 16 * 8)
    seg = 4 * M * max(1, req.L)
    return plan_bytes(req) + int(1.25 * (data + image + seg)) + (32 << 20)


def plan_key(req: ReconRequest):
    """Requests with equal keys can run on the same Gnufft plan: the same
    trajectory (its bytes hashed), geometry, kernel width and grid count."""
    h = hashlib.blake2b(digest_size=16)
    for a in (req.kx, req.ky, req.kz):
        h.update(np.ascontiguousarray(np.asarray(a, np.float32)).tobytes())
    return (req.M, float(req.gridos), req.Nx, req.Ny, req.Nz, float(req.kernel_width),
            req.nc * max(1, req.L), h.hexdigest())


def run_request(req: ReconRequest, stream: Optional[Stream] = None, keep: Optional[list] = None,
                plan: Optional[Gnufft] = None):
    """Build the request's operators, solve and return (image, info); every
    copy and launch goes on `stream`.  With `keep` (a list) the device
    objects are appended to it instead of being released here, so that the
    caller can release them after handing the image on.  `plan`: a Gnufft of
    the request's trajectory and geometry (plan_key) to use instead of a new
    one; its tables are a function of the trajectory alone and every launch
    overwrites the workspaces it reads, so the image is the same."""
    M, N = req.M, req.N
    G = plan if plan is not None else Gnufft(M, req.gridos, req.Nx, req.Ny, req.Nz, req.kx,
                                             req.ky, req.kz, kernel_width=req.kernel_width)
    T = TimeSegmentation(G, req.fieldmap, req.tvec, M, N, req.L, interptype=req.interptype)
    A = SENSE(T, req.sens, M, N, req.nc)
    if req.penalty == "tv":
        R = TVPenalty(req.Nx, req.Ny, req.Nz, req.beta, req.delta)
    else:
        R = QuadPenalty(req.Nx, req.Ny, req.Nz, req.beta)
    y = DeviceArray.from_host(np.asarray(req.data, np.complex64), stream=stream)
    x, info = solve_pwls_pcg(None, A, None, y, R, req.niter, CG_GRAPH, stream, req.tol)
    img = x.to_host(stream)
    if keep is not None:
        keep.extend((x, y, A, T, G, R))
    del x, y, A, T, G, R
    return img, info


class ReconServer:
    """See the module docstring.  submit() returns a Future of a
    ReconResult; requests whose estimate exceeds the whole budget fail with
    MemoryError at once.

    keep_plans: a finished request's Gnufft plan (its tap tables, work lists
    and workspaces) stays on the device, idle, and the next request with the
    same trajectory, geometry and grid count (plan_key) runs on it instead of
    building its own.  Idle plans keep their bytes reserved in the budget;
    the least recently used are released when an admission needs the room,
    and they never hold more than half of the budget."""

    def __init__(self, slots: int = 4, mem_fraction: float = 0.8,
                 budget_bytes: Optional[int] = None, keep_plans: bool = True):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        free, _ = mem_info()
        self.budget = int(budget_bytes) if budget_bytes else int(free * mem_fraction)
        self.keep_plans = keep_plans
        self.reserved = 0
        self.peak_reserved = 0
        self.max_running = 0
        self.plan_hits = 0
        self.pooled = 0  # bytes reserved by idle plans
        self._running = 0
        self._pool: collections.OrderedDict = collections.OrderedDict()  # key -> [(plan, bytes)]
        self._to_free: list = []
        self._cv = threading.Condition()
        self._queue: collections.deque = collections.deque()
        self._closed = False
        self._threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(slots)]
        for t in self._threads:
            t.start()

    def submit(self, req: ReconRequest) -> Future:
        fut: Future = Future()
        need = estimate_bytes(req)
        if need > self.budget:
            fut.set_exception(MemoryError(
                f"request needs ~{need} device bytes, the budget is {self.budget}"))
            return fut
        key = plan_key(req) if self.keep_plans else None
        with self._cv:
            if self._closed:
                raise RuntimeError("server closed")
            self._queue.append((req, need, fut, time.perf_counter(), key))
            self._cv.notify_all()
        return fut

    def close(self):
        """Finish the queued requests, stop the workers and release the idle
        plans."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        with self._cv:
            for plans in self._pool.values():
                for plan, nb in plans:
                    self._to_free.append(plan)
                    self.reserved -= nb
            self._pool.clear()
            self.pooled = 0
        self._drain()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _evict_lru(self, keep_key=None) -> bool:
        """Release the least recently used idle plan (not of keep_key); under
        the lock.  The plan itself is freed later, outside the lock."""
        for key in self._pool:
            if key == keep_key:
                continue
            plans = self._pool[key]
            plan, nb = plans.pop(0)
            if not plans:
                del self._pool[key]
            self.pooled -= nb
            self.reserved -= nb
            self._to_free.append(plan)
            return True
        return False

    def _drain(self):
        """free the released plans (outside the lock: a free waits for the
        device)"""
        with self._cv:
            plans, self._to_free = self._to_free, []
        del plans

    def _take(self):
        # arrival order: the head waits for memory rather than being passed
        with self._cv:
            while True:
                if self._queue:
                    req, need, fut, t_sub, key = self._queue[0]
                    idle = self._pool.get(key) if key is not None else None
                    held = idle[-1][1] if idle else 0  # an idle plan's bytes are reserved
                    extra = need - held
                    while self.reserved + extra > self.budget and self._evict_lru(key):
                        pass
                    if self.reserved + extra <= self.budget:
                        self._queue.popleft()
                        plan = None
                        if idle:
                            plan, _ = idle.pop()
                            if not idle:
                                del self._pool[key]
                            self.pooled -= held
                            self.plan_hits += 1
                        self.reserved += extra
                        self.peak_reserved = max(self.peak_reserved, self.reserved)
                        self._running += 1
                        self.max_running = max(self.max_running, self._running)
                        return req, need, fut, t_sub, key, plan
                elif self._closed:
                    return None
                self._cv.wait()

    def _worker(self):
        stream = Stream()
        while True:
            self._drain()
            item = self._take()
            if item is None:
                return
            req, need, fut, t_sub, key, plan = item
            t0 = time.perf_counter()
            keep: list = []
            ok = False
            try:
                img, info = run_request(req, stream, keep, plan)
                ok = True
                fut.set_result(ReconResult(img, info["iterations"], info["exit_reason"], t_sub,
                                           t0, time.perf_counter(), need, req.tag))
            except BaseException as e:  # reported through the future
                fut.set_exception(e)
            finally:
                # the plan goes back to the pool (its bytes stay reserved); the
                # other operators and buffers go after the image is handed on
                # (their release waits for the device)
                used = keep[4] if len(keep) == 6 else plan
                keep.clear()
                plan = None
                with self._cv:
                    self._running -= 1
                    nb = plan_bytes(req)
                    if ok and key is not None and used is not None and not self._closed_pool():
                        self._pool.setdefault(key, []).append((used, nb))
                        self._pool.move_to_end(key)
                        self.pooled += nb
                        self.reserved -= need - nb
                        while self.pooled > self.budget // 2 and self._evict_lru(None):
                            pass
                    else:
                        if used is not None:
                            self._to_free.append(used)
                        self.reserved -= need
                    self._cv.notify_all()
                used = None

    def _closed_pool(self):
        return not self.keep_plans


__all__ = ["ReconRequest", "ReconResult", "ReconServer", "estimate_bytes", "mem_info",
           "plan_bytes", "plan_key", "run_request"]
