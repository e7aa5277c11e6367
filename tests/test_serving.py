"""Serving-loop scheduler on the CPU (simulated clock, fake executor): admission by
device-memory budget, FIFO order, rejection of requests that can never fit,
failure of one admission not affecting others, continuous batching, and the
per-request choice of form."""
import pytest

from pynof_amd.serving import Executor, Request, Server, poisson_requests


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-3)


class FakeExecutor(Executor):
    def __init__(self, clock, cost_per_build=0.01, fail_rids=()):
        self.clock = clock
        self.cost = cost_per_build
        self.fail = set(fail_rids)
        self.admitted, self.released, self.forms = [], [], []
        self.live = set()

    def footprint(self, req):
        return req.nbf * 1_000_000

    def admit(self, req):
        if req.rid in self.fail:
            raise MemoryError("injected allocation failure")
        self.admitted.append(req.rid)
        self.live.add(req.rid)

    def step(self, reqs):
        assert all(r.rid in self.live for r in reqs)
        self.forms.append({r.rid: r.form_for_build(r.builds_done) for r in reqs})
        for r in reqs:
            r.energies.append(float(r.rid))
        self.clock.t += self.cost * len(reqs)

    def release(self, req):
        self.live.discard(req.rid)
        self.released.append(req.rid)


def make(rid, nbf, arrival=0.0, tol=1e-9, n=5):
    return Request(rid=rid, nbf=nbf, naux=8, nbf5=nbf, ndoc=1, tol=tol, n_builds=n, arrival=arrival)


def test_budget_and_fifo():
    clk = Clock()
    ex = FakeExecutor(clk)
    srv = Server(ex, hbm_budget=250e6, clock=clk, sleep=clk.sleep)
    for i in range(4):
        srv.submit(make(i, 100))
    peak = []
    orig = ex.step

    def step(reqs):
        peak.append(len(reqs))
        orig(reqs)

    ex.step = step
    srv.run(t0=0.0)
    assert [r.rid for r in srv.finished] == [0, 1, 2, 3]
    assert ex.admitted == [0, 1, 2, 3]
    assert max(peak) == 2  # 2 x 100 MB fit the 250 MB budget, a third does not
    assert srv.in_use == 0 and not ex.live


def test_oversized_rejected_others_served():
    clk = Clock()
    ex = FakeExecutor(clk)
    srv = Server(ex, hbm_budget=150e6, clock=clk, sleep=clk.sleep)
    srv.submit(make(0, 100))
    srv.submit(make(1, 500))
    srv.submit(make(2, 100))
    srv.run(t0=0.0)
    assert [r.rid for r in srv.rejected] == [1] and "exceeds" in srv.rejected[0].error
    assert sorted(r.rid for r in srv.finished) == [0, 2]


def test_admission_failure_isolated():
    clk = Clock()
    ex = FakeExecutor(clk, fail_rids=(1,))
    srv = Server(ex, hbm_budget=1e9, clock=clk, sleep=clk.sleep)
    for i in range(3):
        srv.submit(make(i, 50))
    srv.run(t0=0.0)
    assert [r.rid for r in srv.rejected] == [1] and "injected" in srv.rejected[0].error
    assert sorted(r.rid for r in srv.finished) == [0, 2]
    assert srv.in_use == 0


def test_continuous_batching_joins_mid_flight():
    clk = Clock()
    ex = FakeExecutor(clk, cost_per_build=0.01)
    srv = Server(ex, hbm_budget=1e9, clock=clk, sleep=clk.sleep)
    srv.submit(make(0, 50, arrival=0.0, n=10))
    srv.submit(make(1, 50, arrival=0.035, n=4))
    srv.run(t0=0.0)
    assert srv.step_sizes[0] == 1 and 2 in srv.step_sizes
    first_joint = srv.step_sizes.index(2)
    assert 3 <= first_joint <= 5  # joined at the first step after its arrival
    r1 = next(r for r in srv.finished if r.rid == 1)
    assert r1.done < next(r for r in srv.finished if r.rid == 0).done  # finished while 0 was still running


def test_form_per_tolerance():
    from pynof_amd.serving import SPLIT_ENERGY_ERROR_EH, split_error_bound

    # the final build is refined when the split error bound at the size exceeds the tolerance
    assert [make(0, 10, tol=1e-9, n=3).form_for_build(i) for i in range(3)] == ["split", "split", "split_refine"]
    assert [make(0, 10, tol=1e-6, n=3).form_for_build(i) for i in range(3)] == ["split", "split", "split_refine"]
    assert [make(0, 10, tol=1.0, n=3).form_for_build(i) for i in range(3)] == ["split"] * 3
    big = split_error_bound(800)
    assert make(0, 800, tol=0.5 * big, n=2).form_for_build(1) == "split_refine"
    assert make(0, 800, tol=2.0 * big, n=2).form_for_build(1) == "split"
    # the bound is the measured error times the margin at the tabulated sizes, monotone in between
    for nbf, e in SPLIT_ENERGY_ERROR_EH:
        assert abs(split_error_bound(nbf) - 10.0 * e) <= 1e-12 * e
    assert split_error_bound(24) <= split_error_bound(100) <= split_error_bound(800)
    assert split_error_bound(10) == split_error_bound(24) and split_error_bound(2000) == split_error_bound(800)
    # every size of the serving mix refines both tolerance classes
    assert min(split_error_bound(n) for n in range(24, 801)) > 1e-6


def test_timeout():
    clk = Clock()
    ex = FakeExecutor(clk, cost_per_build=1.0)
    srv = Server(ex, hbm_budget=1e9, clock=clk, sleep=clk.sleep)
    srv.submit(make(0, 10, n=100))
    with pytest.raises(TimeoutError):
        srv.run(t0=0.0, timeout=5.0)
    # the unfinished request is released (its device memory freed) and marked
    assert not ex.live and ex.released == [0] and srv.in_use == 0 and not srv.active
    assert [r.rid for r in srv.rejected] == [0] and "not finished" in srv.rejected[0].error


def test_poisson_requests_shape():
    rs = poisson_requests(200, rate=10.0, seed=5000)
    assert all(24 <= r.nbf <= 800 for r in rs)
    assert all(r.naux % 8 == 0 and r.naux >= 3.6 * r.nbf for r in rs)
    arr = [r.arrival for r in rs]
    assert arr == sorted(arr)
    assert rs[0].nbf == poisson_requests(1, rate=10.0, seed=5000)[0].nbf
    med = sorted(r.nbf for r in rs)[100]
    assert 100 <= med <= 220
    assert {r.tol for r in rs} == {1e-6, 1e-9}


def test_poisson_requests_same_requests_at_every_rate():
    a, b = poisson_requests(50, rate=0.0, seed=5000), poisson_requests(50, rate=7.5, seed=5000)
    assert [(r.nbf, r.tol, r.seed) for r in a] == [(r.nbf, r.tol, r.seed) for r in b]
    assert all(r.arrival == 0.0 for r in a) and b[-1].arrival > 0.0


def test_served_request_is_not_resubmitted():
    """A served Request object carries its finished build count; submitting it again
    raises instead of looping (the count would never meet n_builds again)."""
    clock = Clock()
    ex = FakeExecutor(clock)
    srv = Server(ex, hbm_budget=1e12, clock=clock, sleep=clock.sleep)
    r = make(0, 40)
    srv.submit(r)
    srv.run(t0=0.0)
    assert r.builds_done == r.n_builds
    with pytest.raises(ValueError):
        Server(FakeExecutor(clock), hbm_budget=1e12, clock=clock, sleep=clock.sleep).submit(r)


class ThreadedExecutor(Executor):
    """Admissions on the server's worker thread (real clock): request 1's admission is slow."""

    concurrent_admission = True

    def __init__(self, fail_rids=()):
        import threading

        self.fail = set(fail_rids)
        self.live = set()
        self.admit_threads = set()
        self.in_admission = threading.Event()
        self.steps_during_admission = 0
        self.order = []

    def footprint(self, req):
        return 1000

    def admit(self, req):
        import threading
        import time

        self.admit_threads.add(threading.current_thread().name)
        if req.rid in self.fail:
            raise MemoryError("injected allocation failure")
        if req.rid == 1:
            self.in_admission.set()
            time.sleep(0.05)
            self.in_admission.clear()
        self.live.add(req.rid)

    def step(self, reqs):
        import time

        assert all(r.rid in self.live for r in reqs)
        if self.in_admission.is_set():
            self.steps_during_admission += 1
        for r in reqs:
            if r.builds_done == 0:
                self.order.append(r.rid)
            r.energies.append(float(r.rid))
        time.sleep(0.002)

    def release(self, req):
        self.live.discard(req.rid)


def test_concurrent_admission_overlaps_steps():
    ex = ThreadedExecutor(fail_rids=(2,))
    srv = Server(ex, hbm_budget=1e6)
    reqs = [make(0, 10, arrival=0.0, n=40), make(1, 10, arrival=0.005, n=5), make(2, 10, arrival=0.006, n=5),
            make(3, 10, arrival=0.007, n=5)]
    for r in reqs:
        srv.submit(r)
    srv.run(timeout=30)
    assert sorted(r.rid for r in srv.finished) == [0, 1, 3]
    assert [r.rid for r in srv.rejected] == [2] and "injected" in srv.rejected[0].error
    assert srv.in_use == 0 and not srv.admitting
    assert ex.steps_during_admission > 0  # request 0 kept stepping while request 1 was admitted
    assert all(name.startswith("admit") for name in ex.admit_threads)
    assert ex.order == [0, 1, 3]  # requests join in arrival order
    assert all(len(r.energies) == r.n_builds for r in srv.finished)


def test_small_rotation_exponential_matches_expm():
    import numpy as np
    import scipy.linalg
    import torch

    from pynof_amd.serving import _expm_small

    g = torch.Generator().manual_seed(3)
    for n in (1, 24, 150):
        a = torch.randn((n, n), dtype=torch.float64, generator=g)
        k = a - a.T
        if n > 1:
            k *= 1e-2 / torch.linalg.norm(k)
        e = _expm_small(k).numpy()
        assert np.max(np.abs(e - scipy.linalg.expm(k.numpy()))) < 1e-16
        assert np.max(np.abs(e.T @ e - np.eye(n))) < 1e-14


class SlowFailExecutor(ThreadedExecutor):
    """Request 1's admission is still running when the loop stops, and then fails."""

    def admit(self, req):
        import time

        if req.rid == 1:
            time.sleep(0.1)
            raise MemoryError("injected late allocation failure")
        self.live.add(req.rid)


def test_loop_stop_settles_failed_admission_in_flight():
    ex = SlowFailExecutor()
    srv = Server(ex, hbm_budget=1e6)
    srv.submit(make(0, 10, arrival=0.0, n=100000))
    srv.submit(make(1, 10, arrival=0.001, n=5))
    with pytest.raises(TimeoutError):
        srv.run(timeout=0.03)
    assert srv.in_use == 0 and not srv.admitting and not srv.active
    errs = {r.rid: r.error for r in srv.rejected}
    assert set(errs) == {0, 1}
    assert "injected late" in errs[1] and "not finished" in errs[0]
    assert ex.live == set()  # request 0 was released


class ReleaseFailExecutor(FakeExecutor):
    def release(self, req):
        super().release(req)
        raise RuntimeError("injected release failure")


def test_failed_release_keeps_the_books():
    clock = Clock()
    srv = Server(ReleaseFailExecutor(clock), hbm_budget=1e12, clock=clock, sleep=clock.sleep)
    srv.submit(make(0, 40, n=2))
    with pytest.raises(RuntimeError):
        srv.run(t0=0.0)
    assert [r.rid for r in srv.finished] == [0] and srv.in_use == 0 and not srv.active


class LaneExecutor(FakeExecutor):
    """Asynchronous steps on a simulated clock: a lane's step takes cost_per_build times the
    largest nbf of its batch / 100, and lanes run at the same time."""

    async_steps = True

    def __init__(self, clock):
        super().__init__(clock)
        self.launches = []  # (time, lane, rids)

    def step_async(self, reqs, lane):
        assert all(r.rid in self.live and r.lane == lane for r in reqs)
        for r in reqs:
            r.energies.append(float(r.rid))
        self.launches.append((self.clock.t, lane, tuple(r.rid for r in reqs)))
        return self.clock.t + self.cost * max(r.nbf for r in reqs) / 100.0

    def step_done(self, handle):
        return self.clock.t >= handle

    def wait_any(self, handles, timeout):
        self.clock.t = max(self.clock.t, min(min(handles), self.clock.t + timeout))


def test_lanes_step_independently():
    clk = Clock()
    ex = LaneExecutor(clk)
    srv = Server(ex, hbm_budget=1e12, clock=clk, sleep=clk.sleep, lanes=(128, 256))
    big, small = make(0, 800, n=4), make(1, 40, arrival=0.001, n=4)
    mid = make(2, 200, arrival=0.002, n=4)
    for r in (big, small, mid):
        srv.submit(r)
    assert (big.lane, small.lane, mid.lane) == (2, 0, 1)
    srv.run(t0=0.0)
    assert sorted(r.rid for r in srv.finished) == [0, 1, 2] and srv.in_use == 0
    assert all(len(r.energies) == r.n_builds for r in srv.finished)
    # the small request steps while the big one's step (0.08 per build) is in flight: it is
    # done long before the big request, not after four of its steps
    assert small.done < 0.04 < big.done
    assert {lane for _, lane, _ in ex.launches} == {0, 1, 2}
    for _, lane, rids in ex.launches:  # a lane's batch holds only its own size class
        assert all({0: 2, 1: 0, 2: 1}[rid] == lane for rid in rids)


def test_lanes_need_async_executor():
    clk = Clock()
    with pytest.raises(ValueError):
        Server(FakeExecutor(clk), lanes=(128,))
    with pytest.raises(ValueError):
        Server(LaneExecutor(clk), lanes=(256, 128))


def test_lanes_timeout_waits_for_inflight_steps():
    clk = Clock()
    ex = LaneExecutor(clk)
    srv = Server(ex, hbm_budget=1e12, clock=clk, sleep=clk.sleep, lanes=(128,))
    srv.submit(make(0, 800, n=1000))
    with pytest.raises(TimeoutError):
        srv.run(t0=0.0, timeout=0.5)
    assert srv.in_use == 0 and not srv.active and [r.rid for r in srv.rejected] == [0]
    assert ex.live == set()


class _FakeEngine:
    """A stand-in for JKEngine in the pool policy test: its device bytes and whether it
    was closed."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.closed = False

    @property
    def footprint(self):
        return self.nbytes

    def close(self):
        self.closed = True


def test_engine_pool_policy(monkeypatch):
    """DeviceExecutor's engine pool (no device needed): an idle engine serves a request whose
    engine bytes it holds with at most 50% to spare (the smallest such), otherwise a new
    engine is made (an idle engine is never grown); released engines stay idle within the
    count and byte caps and are closed beyond them; close() empties the pool."""
    import threading

    from pynof_amd import serving as S

    ex = object.__new__(S.DeviceExecutor)
    ex._idle, ex._idle_lock, ex.device = [], threading.Lock(), 0
    ex.pool_engines, ex.pool_bytes, ex.pool_spare = 2, 300.0, 1.5
    made = []

    class NewEngine(_FakeEngine):
        def __init__(self, device):
            super().__init__(0)
            made.append(self)

    monkeypatch.setattr("pynof_amd.engine.JKEngine", NewEngine)
    reset = []  # a reused engine gets the built-in tile shapes back
    monkeypatch.setattr("pynof_amd.autotune.apply", lambda eng, arith, entry: reset.append((eng, arith, entry)))
    need = {}
    monkeypatch.setattr(S.DeviceExecutor, "_engine_bytes", lambda self, r: need[r])
    a, b = _FakeEngine(100), _FakeEngine(140)
    ex._put_engine(a)
    ex._put_engine(b)
    c = _FakeEngine(60)
    ex._put_engine(c)  # a third idle engine is over the count cap: closed
    assert c.closed and [e for _, e in ex._idle] == [a, b]
    need["r1"] = 90  # both hold 90 within 50%: the smaller one
    assert ex._take_engine("r1") is a and reset == [(a, "fp64", None), (a, "split", None)]
    need["r2"] = 150  # nothing holds 150: a new engine (140 is not grown)
    assert ex._take_engine("r2") is made[-1] and [e for _, e in ex._idle] == [b]
    need["r3"] = 90  # 140 is more than 50% over 90: a new engine
    assert ex._take_engine("r3") is made[-1] and len(made) == 2
    ex._put_engine(_FakeEngine(150))  # 140 + 150 bytes: within both caps
    third = _FakeEngine(5)
    ex._put_engine(third)  # a third idle engine: over the count cap
    ex.pool_engines = 3
    heavy = _FakeEngine(20)
    ex._put_engine(heavy)  # 140 + 150 + 20 bytes exceed the byte cap
    assert third.closed and heavy.closed and len(ex._idle) == 2
    idle = [e for _, e in ex._idle]
    ex.close()
    assert all(e.closed for e in idle) and ex._idle == []


def test_admitted_footprint_charges_the_budget():
    """An executor whose admission holds more than the estimate (a reused, larger engine)
    has the difference charged to the budget when the request activates, and released."""

    class Holder(FakeExecutor):
        def admitted_footprint(self, req):
            return req.footprint + 30

    clk = Clock()
    ex = Holder(clk)
    srv = Server(ex, hbm_budget=100e6, clock=clk, sleep=clk.sleep)
    reqs = [make(i, 10, n=2) for i in range(3)]
    for r in reqs:
        srv.submit(r)
    est = reqs[0].footprint
    srv._admit(clk())
    assert len(srv.active) == 3 and all(r.footprint == est + 30 for r in srv.active)
    assert srv.in_use == 3 * (est + 30)
    srv.run(timeout=60)
    assert len(srv.finished) == 3 and srv.in_use == 0
