"""Host logic of the serving loop (powergrid_amd.serve) with the device work
replaced by a stand-in: arrival-order admission, the memory budget, the
immediate failure of a request larger than the budget, and draining on
close.  CPU only."""
import threading
import time

import numpy as np
import pytest

from powergrid_amd import serve as SV


def req(N, nc, L, M, tag):
    z = np.zeros(M, np.float32)
    return SV.ReconRequest(z, z, z, z, N, N, 1, nc, L, np.zeros(N * N, np.float32),
                           np.zeros(N * N * nc, np.complex64), np.zeros(M * nc, np.complex64),
                           tag=tag)


class FakePlan:
    """stands in for a Gnufft plan; records its release"""
    released = []

    def __init__(self, origin):
        self.origin = origin

    def __del__(self):
        FakePlan.released.append(self.origin)


@pytest.fixture
def fake_device(monkeypatch):
    log = []
    lock = threading.Lock()

    def run(r, stream=None, keep=None, plan=None):
        with lock:
            log.append(("start", r.tag))
            log.append(("plan", r.tag, None if plan is None else plan.origin))
        time.sleep(0.02)
        with lock:
            log.append(("end", r.tag))
        if keep is not None:
            # x, y, A, T, G, R as run_request leaves them
            keep.extend((None, None, None, None, plan if plan is not None else FakePlan(r.tag),
                         None))
        return np.full(r.N, r.nc, np.complex64), {"iterations": 3, "exit_reason": "niter"}

    monkeypatch.setattr(SV, "run_request", run)
    monkeypatch.setattr(SV, "Stream", lambda: None)
    monkeypatch.setattr(SV, "mem_info", lambda: (1 << 40, 1 << 40))
    return log


def test_estimate_grows_with_grids_and_samples():
    a = SV.estimate_bytes(req(64, 8, 8, 8192, "a"))
    assert SV.estimate_bytes(req(64, 16, 8, 8192, "b")) > a
    assert SV.estimate_bytes(req(128, 8, 8, 8192, "c")) > a
    assert SV.estimate_bytes(req(64, 8, 8, 32768, "d")) > a
    # the batched grids dominate: 64 grids of 128^2 complex64 twice over
    assert a > 2 * 8 * 64 * 128 * 128


def test_admission_order_and_budget(fake_device):
    reqs = [req(64, 8, 8, 8192, i) for i in range(8)]
    need = SV.estimate_bytes(reqs[0])
    with SV.ReconServer(slots=4, budget_bytes=2 * need + need // 2) as srv:
        futs = [srv.submit(r) for r in reqs]
        outs = [f.result() for f in futs]
    assert srv.peak_reserved <= srv.budget
    assert srv.max_running == 2
    starts = [e[1] for e in fake_device if e[0] == "start"]
    assert starts == list(range(8))  # arrival order
    assert all(o.iterations == 3 and o.latency >= o.finished - o.started for o in outs)
    assert [o.tag for o in outs] == list(range(8))


def test_head_of_line_waits_for_memory(fake_device):
    small = SV.estimate_bytes(req(32, 2, 2, 1024, "s"))
    big = req(128, 16, 8, 32768, "big")
    with SV.ReconServer(slots=4, budget_bytes=SV.estimate_bytes(big) + small) as srv:
        f1 = srv.submit(req(32, 2, 2, 1024, "s1"))
        f2 = srv.submit(big)
        f3 = srv.submit(req(32, 2, 2, 1024, "s2"))
        for f in (f1, f2, f3):
            f.result()
    order = [e[1] for e in fake_device if e[0] == "start"]
    assert order.index("big") < order.index("s2")


def test_too_large_request_fails_at_once(fake_device):
    r = req(64, 8, 8, 8192, "x")
    with SV.ReconServer(slots=2, budget_bytes=SV.estimate_bytes(r) - 1) as srv:
        f = srv.submit(r)
        assert isinstance(f.exception(timeout=1), MemoryError)
    assert fake_device == []


def test_close_drains_the_queue(fake_device):
    srv = SV.ReconServer(slots=1, budget_bytes=1 << 40)
    futs = [srv.submit(req(32, 2, 2, 1024, i)) for i in range(5)]
    srv.close()
    assert all(f.done() and not f.exception() for f in futs)
    with pytest.raises(RuntimeError):
        srv.submit(req(32, 2, 2, 1024, "late"))


def traj_req(seed, tag, nc=8, L=8, M=8192, N=64):
    r = req(N, nc, L, M, tag)
    rng = np.random.default_rng(seed)
    r.kx = rng.uniform(-N / 2, N / 2, M).astype(np.float32)
    return r


def test_idle_plans_serve_requests_with_the_same_trajectory(fake_device):
    # one slot: each request finishes before the next starts, so from the
    # second on each runs on the first one's plan; another trajectory (or
    # grid count) gets its own
    reqs = [traj_req(1, f"a{i}") for i in range(3)] + [traj_req(2, "b"), traj_req(1, "c", L=4)]
    with SV.ReconServer(slots=1, budget_bytes=1 << 40) as srv:
        for r in reqs:
            srv.submit(r).result()
        assert srv.plan_hits == 2
        assert srv.pooled == SV.plan_bytes(reqs[0]) * 2 + SV.plan_bytes(reqs[4])
    plans = {e[1]: e[2] for e in fake_device if e[0] == "plan"}
    assert plans == {"a0": None, "a1": "a0", "a2": "a0", "b": None, "c": None}
    assert SV.plan_key(reqs[0]) == SV.plan_key(reqs[1]) != SV.plan_key(reqs[3])
    assert srv.reserved == 0 and srv.pooled == 0


def test_idle_plans_give_way_to_admissions(fake_device):
    # room for one request: b's admission releases a's idle plan, which is
    # freed outside the lock
    a, b = traj_req(1, "a"), traj_req(2, "b")
    FakePlan.released.clear()
    with SV.ReconServer(slots=1, budget_bytes=SV.estimate_bytes(a) + SV.estimate_bytes(b) // 2) as srv:
        srv.submit(a).result()
        assert srv.pooled == SV.plan_bytes(a)
        srv.submit(b).result()
        assert srv.plan_hits == 0
        assert srv.peak_reserved <= srv.budget
        assert srv.pooled == SV.plan_bytes(b)
    assert "a" in FakePlan.released and "b" in FakePlan.released
    assert srv.reserved == 0


def test_keep_plans_off(fake_device):
    reqs = [traj_req(1, f"a{i}") for i in range(3)]
    with SV.ReconServer(slots=1, budget_bytes=1 << 40, keep_plans=False) as srv:
        for r in reqs:
            srv.submit(r).result()
        assert srv.plan_hits == 0 and srv.pooled == 0
    assert all(e[2] is None for e in fake_device if e[0] == "plan")

