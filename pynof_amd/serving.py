"""Continuous-batching serving loop for orbital-optimisation requests.

A request is one orbital optimisation as PyNOF's ADAM loop runs it
(minimization.py:200-253): a fixed number of JK builds (20 here) with the
orbitals rotated between builds, C <- C exp(kappa) (utils.py:550-565). The
server:

* admits requests first-come-first-served while the device-memory budget
  holds their b tensor, its split planes and their workspace (footprint from
  the C ABI's pnjk_workspace_bytes); a request that could never fit is
  rejected, one that does not fit now waits until memory frees up; with an
  executor that allows it (``concurrent_admission``), admissions (upload of b,
  set-up of C, H, n) run on a worker thread and their own stream while the
  active requests keep stepping, and a request joins the batch at the first
  step after its admission completes;
* keeps each admitted request's b, C, H and occupations resident on the device
  until it completes;
* runs continuous batching: every scheduler step advances every active request
  by one build, all of them together in grouped launches (one per form), and
  requests that arrive while a step runs join at the next step;
* picks each request's form from its energy tolerance: every build but the last
  runs split; the final build, whose energy the request returns, runs split too
  only when the split form's energy error at the request's size is below the
  tolerance (``split_error_bound``, from errors measured on the MI355X), and
  otherwise as split_refine, whose FP64 MO path gives the FP64 energy. For the
  synthetic systems served here the split error (1e-4 to 1e-2 Eh over nbf 24 to
  800) is above both the 1e-6 and the 1e-9 Eh tolerance, so every request
  refines its final build.

The scheduler is independent of the device: `Executor` is the interface, with
`DeviceExecutor` driving pynof_amd on the GPU.
"""
from __future__ import annotations

import dataclasses
import math
import threading
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait

import numpy as np

# Split-form (BF16x3, FP32 accumulation) energy error against the FP64 form in Eh, measured on
# one MI355X with the serving generators (b decay 0.05, naux = ceil8(3.6 nbf), nbf5 = nbf,
# seeded C, H, n): the largest of three seeds per size. |E| grows from 37-54 Eh at nbf = 24 to
# 1.5e5 Eh at nbf = 800, and the error with it.
SPLIT_ENERGY_ERROR_EH = ((24, 1.44e-4), (40, 3.58e-4), (60, 3.86e-4), (96, 1.31e-3), (150, 1.63e-3),
                         (240, 1.88e-3), (400, 3.85e-3), (600, 4.30e-3), (800, 1.02e-2))
SPLIT_ERROR_MARGIN = 10.0


def split_error_bound(nbf: int) -> float:
    """Bound on the split form's energy error (Eh) for an instance of nbf basis functions:
    the measured errors interpolated log-log in nbf (held at the ends), times a margin of 10
    (three seeds per size are few)."""
    xs = np.log([n for n, _ in SPLIT_ENERGY_ERROR_EH])
    ys = np.log([e for _, e in SPLIT_ENERGY_ERROR_EH])
    return SPLIT_ERROR_MARGIN * float(np.exp(np.interp(math.log(max(nbf, 1)), xs, ys)))


@dataclasses.dataclass
class Request:
    rid: int
    nbf: int
    naux: int
    nbf5: int
    ndoc: int
    tol: float = 1e-9
    n_builds: int = 20
    seed: int = 0
    arrival: float = 0.0
    # filled by the server
    started: float | None = None   # admission (b upload) began
    admitted: float | None = None
    done: float | None = None
    builds_done: int = 0
    energies: list = dataclasses.field(default_factory=list)
    error: str | None = None
    footprint: int = 0
    lane: int = 0                  # size class (Server lanes)

    def form_for_build(self, i: int) -> str:
        if i == self.n_builds - 1 and self.tol < split_error_bound(self.nbf):
            return "split_refine"
        return "split"

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class Executor:
    """Device side of the server. An executor whose ``admit`` may run on another thread
    than ``step`` and ``release`` (at most one admission at a time) sets
    ``concurrent_admission = True``."""

    concurrent_admission = False

    def footprint(self, req: Request) -> int:
        raise NotImplementedError

    def admit(self, req: Request) -> None:
        raise NotImplementedError

    def admitted_footprint(self, req: Request) -> int:
        """Device bytes the admitted request holds (at least its estimated footprint): an
        executor that reuses buffers may hold more than the estimate."""
        return req.footprint

    def step(self, reqs: list[Request]) -> None:
        """Runs build number req.builds_done of every request in `reqs` and appends
        each request's energy to req.energies."""
        raise NotImplementedError

    def release(self, req: Request) -> None:
        raise NotImplementedError

    # Optional asynchronous steps (``async_steps = True``), used by a Server with lanes: a
    # lane's step is enqueued and returns a handle at once, so steps of different lanes run
    # on the device together. Such a Server sets ``lanes_active`` before it admits anything.
    async_steps = False
    lanes_active = False

    def step_async(self, reqs: list[Request], lane: int):
        """Enqueues build number req.builds_done of every request in `reqs` (all of one lane);
        returns a handle. Energies are appended to req.energies (they may be read once the
        step is done)."""
        raise NotImplementedError

    def step_done(self, handle) -> bool:
        raise NotImplementedError

    def wait_any(self, handles, timeout: float) -> None:
        """Returns when one of the handles' steps is done or after `timeout` seconds."""
        raise NotImplementedError


class Server:
    def __init__(self, executor: Executor, hbm_budget: float = 200e9, max_active: int = 64, clock=time.perf_counter,
                 sleep=time.sleep, lanes=None):
        """lanes: ascending nbf bounds of the size classes, e.g. (128, 256) for nbf < 128,
        128 <= nbf < 256 and nbf >= 256. With lanes (and an executor with asynchronous steps)
        each size class steps on its own, as soon as its previous step is done, so the builds
        of small requests run beside those of large ones instead of waiting for them; without,
        every step advances every active request together."""
        self.ex = executor
        self.lanes = tuple(int(b) for b in lanes) if lanes else ()
        if self.lanes and not getattr(executor, "async_steps", False):
            raise ValueError("lanes need an executor with asynchronous steps")
        if list(self.lanes) != sorted(set(self.lanes)):
            raise ValueError("lane bounds must be strictly ascending")
        if self.lanes:
            executor.lanes_active = True
        self.budget = float(hbm_budget)
        self.max_active = int(max_active)
        self.clock = clock
        self.sleep = sleep
        self.pending: deque[Request] = deque()
        self.admitting: deque = deque()  # (request, future) of admissions in flight, FIFO
        self._pool: ThreadPoolExecutor | None = None
        self.active: list[Request] = []
        self.finished: list[Request] = []
        self.rejected: list[Request] = []
        self.in_use = 0
        self.steps = 0
        self.step_sizes: list[int] = []

    def submit(self, req: Request) -> None:
        if req.builds_done or req.energies or req.done is not None:
            raise ValueError(f"request {req.rid} was already served; submit a fresh Request")
        req.footprint = int(self.ex.footprint(req))
        req.lane = sum(req.nbf >= b for b in self.lanes)
        if req.footprint > self.budget:
            req.error = f"footprint {req.footprint} B exceeds the device budget {int(self.budget)} B"
            self.rejected.append(req)
            return
        self.pending.append(req)

    def _admit(self, now: float) -> None:
        while (self.pending and self.pending[0].arrival <= now
               and len(self.active) + len(self.admitting) < self.max_active):
            r = self.pending[0]
            if self.in_use + r.footprint > self.budget:
                break  # first-come-first-served: later requests wait behind it
            self.pending.popleft()
            r.started = self.clock()
            self.in_use += r.footprint
            if self._pool is not None:
                self.admitting.append((r, self._pool.submit(self.ex.admit, r)))
                continue
            try:
                self.ex.admit(r)
            except Exception as exc:  # allocation or upload failure: fail this request only
                self._fail(r, exc)
                continue
            self._activate(r)
        while self.admitting and self.admitting[0][1].done():
            r, fut = self.admitting.popleft()
            exc = fut.exception()
            if exc is not None:
                self._fail(r, exc)
            else:
                self._activate(r)

    def _activate(self, r: Request) -> None:
        held = int(self.ex.admitted_footprint(r))  # the books take what admission really holds
        if held > r.footprint:
            self.in_use += held - r.footprint
            r.footprint = held
        r.admitted = self.clock()
        self.active.append(r)

    def _fail(self, r: Request, exc: BaseException) -> None:
        r.error = f"admission failed: {exc}"
        self.in_use -= r.footprint
        self.rejected.append(r)

    def run(self, t0: float | None = None, timeout: float = 1e9) -> None:
        """Serves until every submitted request is done or rejected. Arrival times
        are relative to t0 (default: now)."""
        base = self.clock() if t0 is None else t0
        for r in self.pending:
            r.arrival += base
        end = base + timeout
        if getattr(self.ex, "concurrent_admission", False):
            self._pool = ThreadPoolExecutor(1, thread_name_prefix="admit")
        try:
            self._loop(end)
        finally:
            if self._pool is not None:
                self._pool.shutdown(wait=True)
                self._pool = None
            self._abandon()

    def _abandon(self) -> None:
        """After the loop stopped on an error (a time-out, a failed step): releases the
        requests it admitted but did not finish, so their device memory is freed, and lists
        them with the rejected ones (done stays None); a normal end leaves nothing to release.
        Lane steps still in flight are waited for first (their requests' buffers are in use)."""
        for _, h in getattr(self, "_inflight", {}).values():
            try:
                while not self.ex.step_done(h):
                    self.ex.wait_any([h], 0.01)
            except Exception:  # best effort: the step's failure is not this request's to report
                pass
        self._inflight = {}
        left = []
        for r, fut in self.admitting:  # the pool has shut down: every admission has ended
            exc = fut.exception()
            if exc is not None:
                self._fail(r, exc)
            else:
                left.append(r)
        left += self.active
        self.admitting.clear()
        self.active = []
        for r in left:
            r.error = r.error or "not finished: the serving loop stopped"
            try:
                self.ex.release(r)
            except Exception:  # best effort: keep releasing the others
                pass
            self.in_use -= r.footprint
            self.rejected.append(r)

    def _finish_step(self, batch: list[Request]) -> None:
        self.steps += 1
        self.step_sizes.append(len(batch))
        t = self.clock()
        for r in batch:
            r.builds_done += 1
            if r.builds_done >= r.n_builds:
                r.done = t
                self.active.remove(r)
                try:
                    self.ex.release(r)
                finally:  # the request has left the device's books even if release failed
                    self.in_use -= r.footprint
                    self.finished.append(r)

    def _loop(self, end: float) -> None:
        if self.lanes:
            return self._loop_lanes(end)
        while self.pending or self.active or self.admitting:
            now = self.clock()
            if now > end:
                raise TimeoutError("serving loop exceeded its time limit")
            self._admit(now)
            if not self.active:
                if self.admitting:
                    wait([self.admitting[0][1]], timeout=0.05, return_when=FIRST_COMPLETED)
                elif self.pending:
                    self.sleep(max(0.0, min(self.pending[0].arrival - now, 0.05)))
                continue
            batch = list(self.active)
            self.ex.step(batch)
            self._finish_step(batch)

    def _loop_lanes(self, end: float) -> None:
        inflight: dict[int, tuple] = {}  # lane -> (batch, handle)
        self._inflight = inflight
        while self.pending or self.active or self.admitting or inflight:
            now = self.clock()
            if now > end:
                raise TimeoutError("serving loop exceeded its time limit")
            self._admit(now)
            launched = False
            for lane in range(len(self.lanes) + 1):
                if lane in inflight:
                    continue
                batch = [r for r in self.active if r.lane == lane]
                if batch:
                    inflight[lane] = (batch, self.ex.step_async(batch, lane))
                    launched = True
            finished = False
            for lane in list(inflight):
                batch, h = inflight[lane]
                if self.ex.step_done(h):
                    del inflight[lane]
                    self._finish_step(batch)
                    finished = True
            if launched or finished:
                continue
            if inflight:
                self.ex.wait_any([h for _, h in inflight.values()], 0.002)
            elif self.admitting:
                wait([self.admitting[0][1]], timeout=0.05, return_when=FIRST_COMPLETED)
            elif self.pending:
                self.sleep(max(0.0, min(self.pending[0].arrival - now, 0.05)))

    def stats(self) -> dict:
        lat = np.array([r.latency for r in self.finished]) if self.finished else np.zeros(0)
        out = {"finished": len(self.finished), "rejected": len(self.rejected), "steps": self.steps,
               "mean_batch": float(np.mean(self.step_sizes)) if self.step_sizes else 0.0}
        if lat.size:
            out.update(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                       mean_s=float(lat.mean()))
        return out


def poisson_requests(count: int, rate: float, seed: int = 5000, median_nbf: int = 150, lo: int = 24, hi: int = 800,
                     sigma: float = 0.6, tol_tight_fraction: float = 0.5, n_builds: int = 20) -> list[Request]:
    """Requests with Poisson arrivals (rate per second) and lognormal nbf (median
    `median_nbf`, clipped to [lo, hi]); naux = 3.6 nbf rounded up to a multiple of
    8, nbf5 = nbf; half the requests ask for 1e-9 Eh, half for 1e-6 Eh. Sizes and
    tolerances come from their own random stream, so every rate (including 0, all
    at t = 0) serves the same requests."""
    from .synth import ceil8, default_ndoc

    rng = np.random.default_rng(seed)
    arrivals = np.random.default_rng([seed, 1])
    t = 0.0
    out = []
    for i in range(count):
        t += arrivals.exponential(1.0 / rate) if rate > 0 else 0.0
        nbf = int(np.clip(round(median_nbf * math.exp(sigma * rng.standard_normal())), lo, hi))
        tol = 1e-9 if rng.random() < tol_tight_fraction else 1e-6
        out.append(Request(rid=i, nbf=nbf, naux=ceil8(3.6 * nbf), nbf5=nbf, ndoc=default_ndoc(nbf), tol=tol,
                           n_builds=n_builds, seed=6000 + i, arrival=t))
    return out


def _expm_small(k, eye=None):
    """exp(k) for a matrix of norm ||k||_F <= 1e-2, as its Taylor series to degree 8 in
    Horner form: the first omitted term is below 1e-2^9 / 9! < 1e-23, far under FP64
    rounding. Unlike torch.linalg.matrix_exp, which picks its degree from the norm on the
    host, it never waits for the device, so the loop queues every request's rotation
    without a synchronisation. Each Horner step is one addmm (eye + (k p) / j)."""
    import torch

    if eye is None:
        eye = torch.eye(k.shape[0], dtype=k.dtype, device=k.device)
    p = torch.add(eye, k, alpha=1.0 / 8.0)
    for j in (7.0, 6.0, 5.0, 4.0, 3.0, 2.0, 1.0):
        p = torch.addmm(eye, k, p, alpha=1.0 / j)
    return p


class DeviceExecutor(Executor):
    """Runs requests on one GPU: one JKEngine per request (its b resident), grouped
    builds per step. C is rotated on the device by exp(kappa_i) with a seeded
    antisymmetric kappa_i, ||kappa_i||_F = 1e-2.

    With ``occupation=True`` each build is followed by an occupation step on the
    device (GNOF, softmax parameters; pnjk_occupation_step): the request's energy
    is calcocce at its current parameters gamma (pnof.py:994-1001), and gamma takes
    one gradient step, gamma -= occ_lr * calcoccg, before the next build. The
    occupation block is the first ndoc (1 + ncwo) orbitals, ncwo = (nbf5 - ndoc) // ndoc.
    J_MO, K_MO, H_core and gamma never leave the GPU.

    Admission may run on the server's worker thread: it works on a stream of its own and
    synchronises it before returning, so the request's tensors are ready for the build
    stream; those tensors are then marked as used on the build stream, so that when a step
    replaces one (the rotated C, the updated gamma) the caching allocator does not hand its
    memory back to the admission stream before the build stream's queued reads of it have
    run. The per-request state table is guarded by a lock.

    ``keep_final_C`` keeps each request's final C and the outputs named in ``keep_outs``
    (default: all) on the host at release.

    Engines are pooled: a released request's engine keeps its device buffers (they are
    grow-only) and serves a later request, so admission creates no streams and allocates
    nothing when an idle engine fits, and release frees nothing (device frees and stream
    creation synchronise with the whole GPU, so they stalled every lane's launches). An
    idle engine is reused for a request whose engine bytes it holds with at most 50% to
    spare (the smallest such); otherwise the request gets a new engine (an idle engine is
    never grown: growing frees its buffers). A reused engine's spare bytes are charged to the
    request (``admitted_footprint``), so the server's budget covers every active engine's real
    size. At most ``pool_engines`` engines (default: the server's default active count) and
    ``pool_bytes`` of device memory stay idle (beyond that a released engine is closed), so
    the device holds at most the budget plus ``pool_bytes`` (plus one admission's spare,
    charged as it completes). ``close()`` frees the pool."""

    concurrent_admission = True

    def __init__(self, device: int = 0, with_jk: bool = True, keep_final_C: bool = False,
                 occupation: bool = False, occ_lr: float = 0.05, keep_outs=None, pool_engines: int = 64,
                 pool_bytes: float = 32e9):
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.keep_outs = keep_outs
        self._build_stream = torch.cuda.current_stream(self.dev)  # the stream step() runs on
        self._lane_streams: dict[int, object] = {}  # lane -> stream of its asynchronous steps
        self.device = device
        self.with_jk = with_jk
        self.keep_final_C = keep_final_C
        self.occupation = occupation
        self.occ_lr = occ_lr
        self.state: dict[int, dict] = {}
        self._state_lock = threading.Lock()
        self._admit_stream = None
        self._groups: dict = {}
        self._idle: list = []  # (device bytes, engine) of pooled engines
        self._idle_lock = threading.Lock()
        self.pool_engines = int(pool_engines)
        self.pool_bytes = float(pool_bytes)
        self.pool_spare = 1.5

    def _engine_bytes(self, req: Request) -> int:
        """Device bytes a request's engine holds: packed b and the split_refine workspace
        with the split planes."""
        from . import _lib

        L = _lib.lib()
        b_bytes = req.nbf * (req.nbf + 1) // 2 * L.pnjk_packed_ld(req.nbf, req.naux) * 8
        return int(b_bytes + L.pnjk_workspace_bytes(req.nbf, req.naux, req.nbf5, 2))

    def _take_engine(self, req: Request):
        from .engine import JKEngine

        need = self._engine_bytes(req)
        pick = None
        with self._idle_lock:
            fits = [x for x in self._idle if need <= x[0] <= self.pool_spare * need]
            if fits:
                pick = min(fits, key=lambda x: x[0])
                self._idle.remove(pick)
        if pick is None:
            return JKEngine(self.device)
        from . import autotune

        for arith in ("fp64", "split"):  # the last request's tile shapes do not carry over
            autotune.apply(pick[1], arith, None)
        return pick[1]

    def _put_engine(self, eng) -> None:
        fp = eng.footprint
        with self._idle_lock:
            if len(self._idle) < self.pool_engines and sum(x[0] for x in self._idle) + fp <= self.pool_bytes:
                self._idle.append((fp, eng))
                return
        eng.close()

    def close(self) -> None:
        """Closes the pooled engines (their device memory is freed)."""
        with self._idle_lock:
            idle, self._idle = self._idle, []
        for _, e in idle:
            e.close()

    def footprint(self, req: Request) -> int:
        from . import _lib

        L = _lib.lib()
        ld = L.pnjk_packed_ld(req.nbf, req.naux)
        ntri = req.nbf * (req.nbf + 1) // 2
        b_bytes = ntri * ld * 8
        ws = L.pnjk_workspace_bytes(req.nbf, req.naux, req.nbf5, 2)  # split_refine needs the most
        outs = (2 * req.nbf5 * req.nbf * req.nbf + 2 * req.nbf5 * req.nbf5 + req.nbf5) * 8
        return int(b_bytes + ws + outs + 3 * req.nbf * req.nbf * 8)

    def admit(self, req: Request) -> None:
        torch = self.torch
        if self._admit_stream is None:
            self._admit_stream = torch.cuda.Stream(self.dev)
        with torch.cuda.device(self.dev), torch.cuda.stream(self._admit_stream):
            st = self._admit_on_stream(req)
            self._admit_stream.synchronize()
        used_on = self._stream_of(req.lane) if self.lanes_active else self._build_stream
        for t in [st["C"], st["H"], st["n"], st.get("gamma"), *st["outs"].values()]:
            if t is not None:
                t.record_stream(used_on)
        with self._state_lock:
            self.state[req.rid] = st

    def admitted_footprint(self, req: Request) -> int:
        """The estimate plus what the request's engine holds beyond its own engine bytes (a
        pooled engine may be up to half as large again)."""
        with self._state_lock:
            eng = self.state[req.rid]["eng"]
        return req.footprint + max(0, eng.footprint - self._engine_bytes(req))

    def _admit_on_stream(self, req: Request) -> dict:
        from .engine import JKEngine
        from .synth import device_b_packed, make_H, make_n

        torch = self.torch
        eng = self._take_engine(req)
        btri = device_b_packed(req.nbf, req.naux, req.seed, 0.05, JKEngine.packed_ld(req.nbf, req.naux), self.dev)
        eng.upload_b_packed(btri, req.nbf, req.naux)
        del btri
        g = torch.Generator(device=self.dev)
        g.manual_seed(req.seed)
        C = torch.linalg.qr(torch.randn((req.nbf, req.nbf), dtype=torch.float64, device=self.dev, generator=g))[0]
        H = torch.from_numpy(make_H(req.nbf, req.seed + 1)).to(self.dev)
        n = make_n(req.nbf5, req.ndoc, req.seed + 2)
        n5, nbf = req.nbf5, req.nbf
        outs = {k: torch.empty(s, dtype=torch.float64, device=self.dev) for k, s in
                (("J_MO", (n5, n5)), ("K_MO", (n5, n5)), ("H_core", (n5,)))}
        if self.with_jk:
            outs["J"] = torch.empty((n5, nbf, nbf), dtype=torch.float64, device=self.dev)
            outs["K"] = torch.empty((n5, nbf, nbf), dtype=torch.float64, device=self.dev)
        st = {"eng": eng, "C": C, "H": H, "n": torch.from_numpy(n).to(self.dev), "outs": outs, "gen": g}
        if self.occupation:
            ncwo = max(0, (n5 - req.ndoc) // req.ndoc)
            pn = JKEngine.pnof(0, req.ndoc, 0, ncwo, ipnof=8, occ_method="Softmax")
            gamma = torch.full((req.ndoc * ncwo,), -2.0, dtype=torch.float64, device=self.dev)
            st.update(pnof=pn, gamma=gamma, occ_out=None)
        return st

    def _group(self, reqs):
        from .batch import JKGroup

        key = tuple(r.rid for r in reqs)
        grp = self._groups.get(key)
        if grp is None:
            with self._state_lock:
                live = set(self.state)
            for k in list(self._groups):
                if not set(k) <= live:
                    self._groups.pop(k).close()
            grp = JKGroup([self.state[r.rid]["eng"] for r in reqs])
            self._groups[key] = grp
        return grp

    async_steps = True

    def _stream_of(self, lane: int):
        with self._state_lock:
            st = self._lane_streams.get(lane)
            if st is None:
                st = self._lane_streams[lane] = self.torch.cuda.Stream(self.dev)
            return st

    def step(self, reqs):
        self._enqueue(reqs)
        self.torch.cuda.current_stream(self.dev).synchronize()

    def step_async(self, reqs, lane: int):
        torch = self.torch
        s = self._stream_of(lane)
        with torch.cuda.stream(s):
            self._enqueue(reqs)
            ev = torch.cuda.Event()
            ev.record(s)
        return ev

    def step_done(self, handle) -> bool:
        return handle.query()

    def wait_any(self, handles, timeout: float) -> None:
        t_end = time.perf_counter() + timeout
        while time.perf_counter() < t_end:
            if any(h.query() for h in handles):
                return
            time.sleep(2e-4)

    def _enqueue(self, reqs):
        """Enqueues one build of every request in `reqs` on the current stream, then each
        request's energy and the orbital rotation for its next build."""
        torch = self.torch
        by_form: dict[str, list] = {}
        for r in reqs:
            by_form.setdefault(r.form_for_build(r.builds_done), []).append(r)
        outputs = ("J", "K", "J_MO", "K_MO", "H_core") if self.with_jk else ("J_MO", "K_MO", "H_core")
        for form, rs in by_form.items():
            st = [self.state[r.rid] for r in rs]
            if len(rs) == 1:  # one request (every final refinement, most steps at low load): no group plan
                st[0]["eng"].build(st[0]["C"], rs[0].nbf5, H=st[0]["H"], precision=form, outputs=outputs,
                                   out=st[0]["outs"])
                continue
            self._group(rs).build([s["C"] for s in st], [r.nbf5 for r in rs], [s["H"] for s in st],
                                  precision=form, outputs=outputs, outs=[s["outs"] for s in st])
        for r in reqs:  # energies (FP64, on the device), then the orbital rotation for the next build
            s = self.state[r.rid]
            o, n = s["outs"], s["n"]
            if self.occupation:
                m = s["pnof"].nbf5
                s["occ_out"] = s["eng"].occupation_step(
                    s["gamma"], s["pnof"], o["J_MO"][:m, :m].contiguous(), o["K_MO"][:m, :m].contiguous(),
                    o["H_core"][:m].contiguous(), out=s["occ_out"])
                e = s["occ_out"]["E"][0].clone()
                if r.builds_done + 1 < r.n_builds:
                    s["gamma"] = s["gamma"] - self.occ_lr * s["occ_out"]["grad"]
            else:
                e = 2.0 * torch.dot(n, o["H_core"]) + n @ (2.0 * o["J_MO"] - o["K_MO"]) @ n
            r.energies.append(e)
            if r.builds_done + 1 < r.n_builds:
                a = torch.randn((r.nbf, r.nbf), dtype=torch.float64, device=self.dev, generator=s["gen"])
                k = a - a.T
                k *= 1e-2 / torch.linalg.norm(k)
                if "eye" not in s:
                    s["eye"] = torch.eye(r.nbf, dtype=torch.float64, device=self.dev)
                s["C"] = s["C"] @ _expm_small(k, s["eye"])

    def release(self, req: Request) -> None:
        with self._state_lock:
            s = self.state.pop(req.rid)
        if req.energies:  # one device-to-host copy for the request's energies
            req.energies = self.torch.stack(req.energies).cpu().tolist()
        if self.keep_final_C:
            req.final_C = s["C"].cpu().numpy()
            req.final_outs = {k: v.cpu().numpy() for k, v in s["outs"].items()
                              if self.keep_outs is None or k in self.keep_outs}
            if self.occupation:
                req.final_gamma = s["gamma"].cpu().numpy()
        for k in list(self._groups):
            if req.rid in k:
                self._groups.pop(k).close()
        self._put_engine(s["eng"])
