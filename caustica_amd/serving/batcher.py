"""Continuous batching of CW solves of mixed grid sizes on one GPU.

Requests arrive over time. Each is planned on the host (coefficient maps, k tables,
source; a small thread pool so planning overlaps device work), then admitted into the
session of its padded shape when a member slot is free and its HBM footprint fits the
budget (footprint = MemberPlan.device_bytes(), the device-side inventory of a member;
upstream's planner estimate for the same solve, planner/model.py:86-115, is larger
because it counts dense operator volumes the device never builds). Every session
advances all of its members together -- one batched launch sequence per step, each
member at its own step counter, so members join and leave at any step boundary --
and different shapes run on their own streams, so their kernels overlap on the
device. A member whose schedule is done is read back and retired on a read-back
thread (its copies run on the session's read-back stream while the serving loop keeps
admitting and advancing the others); its slot and memory are reused by the next
admission of that shape once it is retired.

A member's trajectory does not depend on what it shares launches with: every kernel
reads only the member's own arrays and scalars, so a result is bit-identical alone
and in the mix (tests/test_gpu_serving.py).
"""

from __future__ import annotations

import dataclasses
import hashlib
import queue
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from caustica_amd.solvers.base import SolverDivergedError, SolverResult
from caustica_amd.solvers.kspace.native import KSpaceSession
from caustica_amd.solvers.kspace.operators import MemberPlan, plan_member


@dataclass
class SolveRequest:
    """One CW solve with a pinned schedule (convergence_tol == 0: its step count is known)."""

    rid: int
    arrival_s: float
    solver: str
    grid: object
    medium: object
    source: object
    spec: object
    harmonics: tuple = (1,)
    reference_point: tuple | None = None
    record_region: tuple | None = None
    label: str = ""


@dataclass
class SolveResponse:
    rid: int
    result: SolverResult | None
    error: str | None
    arrival_s: float
    admit_s: float
    done_s: float
    shape: tuple
    batch_peak: int = 1

    @property
    def latency_s(self) -> float:
        return self.done_s - self.arrival_s


@dataclass
class _Live:
    req: SolveRequest
    plan: MemberPlan
    slot: int
    total_steps: int
    rec_steps: int
    admit_s: float
    done_steps: int = 0
    peak_batch: int = 1


@dataclass
class _Group:
    session: KSpaceSession
    voxels: int
    live: dict = field(default_factory=dict)  # slot -> _Live
    busy: bool = False                        # a chunk is queued on the session stream
    reading: int = 0                          # finished members still being read back
    pooled: int = 0                           # the session's retired blocks [bytes], last seen


def schedule_of(plan: MemberPlan, spec) -> tuple[int, int]:
    """(total steps, record steps) of a pinned schedule (engine.py:484-556 with tol = 0)."""
    if spec.convergence_tol != 0.0:
        raise ValueError("served solves need a pinned schedule (convergence_tol == 0)")
    settle = plan.eff_max
    if spec.t_end_min_us is not None:
        settle = max(settle, int(np.ceil(spec.t_end_min_us * 1e-6 * plan.f0)) - spec.n_record_periods)
    rec = spec.n_record_periods * plan.spp
    return settle * plan.spp + rec, rec


_EMPTY = np.zeros(0, np.float32)


def light_plan(plan: MemberPlan) -> MemberPlan:
    """The plan without its host coefficient maps (uploaded already): what a live member
    keeps for read-back and accounting (footprint, record box, schedule, nonlinearity)."""
    return dataclasses.replace(plan, dt_over_rho=_EMPTY, rhoc2_dt=_EMPTY, absorb=_EMPTY,
                               beta2_dt=None if plan.beta2_dt is None else _EMPTY)


def result_digest(res: SolverResult) -> str:
    """128-bit digest over every phasor and p_max of a result (xxh3 when the xxhash module
    is there, BLAKE2b otherwise): equal digests = bitwise equal results."""
    try:
        import xxhash

        h = xxhash.xxh3_128()
    except ImportError:  # pragma: no cover
        h = hashlib.blake2b(digest_size=16)
    for k in sorted(res.phasors):
        h.update(memoryview(np.ascontiguousarray(res.phasors[k])).cast("B"))
    h.update(memoryview(np.ascontiguousarray(res.p_max)).cast("B"))
    return h.hexdigest()


def finish_result(sess: KSpaceSession, slot: int, plan: MemberPlan, req: SolveRequest,
                  total_steps: int, rec_steps: int, periods: int) -> SolverResult:
    """Read a finished member back into the upstream result shape (engine.py:558-586)."""
    peaks = sess.peaks(slot, 0, plan.eff_max)
    history, prev = [], None
    for p in range(1, plan.eff_max + 1):
        pk = float(peaks[p - 1])
        if not np.isfinite(pk):
            raise SolverDivergedError(f"request {req.rid} diverged in settling period {p}")
        rel = abs(pk - prev) / max(prev, 1e-9) if prev is not None else None
        if p >= plan.tof_periods:
            history.append((p, pk, float("nan") if rel is None else rel))
        prev = pk
    phasors, pmax = {}, None
    for hi, h in enumerate(plan.harmonics):
        ph, pm = sess.result(slot, hi, 2.0 / rec_steps, with_pmax=(hi == 0))
        phasors[h] = ph.reshape(plan.rec_shape)
        if hi == 0:
            pmax = pm.reshape(plan.rec_shape)
    if not np.isfinite(pmax).all():
        raise SolverDivergedError(f"request {req.rid} diverged in the record window")
    return SolverResult(
        phasor=phasors[1], p_max=pmax, region=plan.rec, dt=plan.dt, spp=plan.spp,
        steps_total=total_steps, t_end_s=total_steps * plan.dt, tof_periods=plan.tof_periods,
        converged_period=plan.eff_max, settle_capped=True, convergence_history=history,
        phasors=phasors,
        meta={"solver": req.solver, "backend": "hip", "padded_shape": plan.padded,
              "c_ref": plan.c_ref, "nonlinear_active": plan.beta2_dt is not None,
              "source": plan.label, "served": True},
    )


class ContinuousBatcher:
    """Serve a stream of CW solve requests with per-shape batched launches.

    hbm_budget: bytes of device memory members may hold at once (default: 80 % of
    the free device memory at construction). chunk_steps: steps one 256^3 member
    advances per chunk; a group's chunk is scaled to the same voxel-steps (small grids
    and small batches take more steps per chunk), so every chunk is a few milliseconds
    of device work and the host revisits a group only when its stream has drained.

    max_launch_voxels caps the voxels one launch advances in a shape group (see
    __init__); requests beyond the cap wait for a slot of their shape.

    A request that cannot be served fails alone, as an error response: planning errors
    (a schedule that is not pinned, a shape without kernels), a footprint larger than
    the whole budget, an allocation or upload failure at admission, a diverged field at
    read-back. The loop always ends with one response per request.
    """

    def __init__(self, hbm_budget: int | None = None, max_batch: int = 16, chunk_steps: int = 48,
                 planners: int = 4, device: int = 0, max_launch_voxels: int | None = 6 << 20):
        from caustica_amd import _native as N

        self.device = device
        if hbm_budget is None:
            free, _total = N.device_memory(device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        # members of one shape advanced together are capped at this many voxels per launch
        # (at least one member; None: max_batch members of any size). The default, three
        # 128^3 grids, lets the small shapes share launches and runs 192^3 and larger one
        # member at a time: with shapes overlapping on their own streams, larger batches
        # lost more to cache reuse than they gained (a 96-request mixed queue at 32 req/s:
        # 5.35 solves/s uncapped, 5.50 at 16.8 M voxels, 5.60-5.65 at 2.1-6.3 M), and the
        # first of several large requests no longer waits for the others' steps
        self.max_launch_voxels = max_launch_voxels
        self.chunk_steps = int(chunk_steps)
        self.groups: dict[tuple, _Group] = {}
        self.in_use = 0
        self.peak_in_use = 0  # high-water mark of admitted footprints (bytes)
        self.planners = ThreadPoolExecutor(max_workers=planners)
        # read-back and retirement of finished members, off the serving loop
        self.readers = ThreadPoolExecutor(max_workers=2)
        # host time by activity [s] (admission and launches on the serving loop, read-back
        # on the read-back threads), and counts (what keeps the loop from refilling
        # drained groups)
        self.stats = {"admit_s": 0.0, "readback_s": 0.0, "launch_s": 0.0, "plan_wait_s": 0.0,
                      "admitted": 0, "chunks": 0, "idle_polls": 0}

    def _group(self, plan: MemberPlan) -> _Group:
        g = self.groups.get(plan.padded3)
        if g is None:
            from caustica_amd.solvers.kspace.engine import TUNE_MIN_VOXELS
            from caustica_amd.tune import autotune

            g = _Group(KSpaceSession(plan.padded3, self.max_batch, self.device),
                       int(np.prod(plan.padded3)))
            # the same per-shape variant choice as the plain solve path (cached), so a
            # served result is bit-identical to the same solve run alone
            if g.voxels >= TUNE_MIN_VOXELS:
                autotune(g.session, plan)
            self.groups[plan.padded3] = g
        return g

    def tune(self, plan: MemberPlan) -> None:
        """Create (and autotune) the shape group of ``plan`` ahead of the first request."""
        self._group(plan)

    def _plan(self, req: SolveRequest):
        """Host planning (thread pool): the member plan, its pinned schedule and its
        footprint, refused here if it can never fit the budget."""
        plan = plan_member(req.grid, req.medium, req.source, req.spec, harmonics=req.harmonics,
                           record_region=req.record_region, reference_point=req.reference_point,
                           nonlinear=req.solver == "westervelt")
        total, rec = schedule_of(plan, req.spec)
        need = plan.device_bytes()
        if need > self.hbm_budget:
            raise MemoryError(f"the solve needs {need} bytes of device memory, more than the whole "
                              f"serving budget of {self.hbm_budget} bytes")
        return plan, total, rec, need

    def close(self) -> None:
        self.planners.shutdown(wait=True)
        self.readers.shutdown(wait=True)
        for g in self.groups.values():
            g.session.close()
        self.groups.clear()

    def _pooled(self) -> int:
        """Bytes of retired blocks the sessions hold (a session being read back counts with
        its last value)."""
        for x in self.groups.values():
            if x.session.lock.acquire(blocking=False):
                try:
                    x.pooled = x.session.pooled_bytes()
                finally:
                    x.session.lock.release()
        return sum(x.pooled for x in self.groups.values())

    def serve(self, requests: list[SolveRequest], t0: float | None = None,
              on_done=None) -> list[SolveResponse]:
        """Replay ``requests`` by their arrival times (seconds after t0) and serve them.

        Returns one response per request (in completion order)."""
        pending = sorted(requests, key=lambda r: r.arrival_s)
        t0 = time.perf_counter() if t0 is None else t0
        now = lambda: time.perf_counter() - t0  # noqa: E731
        planned: queue.Queue = queue.Queue()
        waiting: list = []  # (req, (plan, total, rec, need)) planned, not admitted
        responses: list[SolveResponse] = []
        n_total = len(pending)
        nxt = 0
        in_flight_plans = 0

        def fail(req, exc, admit_s=None):
            t = now()
            resp = SolveResponse(req.rid, None, f"{type(exc).__name__}: {exc}", req.arrival_s,
                                 t if admit_s is None else admit_s, t, tuple(req.grid.shape))
            responses.append(resp)
            if on_done is not None:
                on_done(resp)

        def submit_plan(req):
            fut = self.planners.submit(self._plan, req)
            fut.add_done_callback(lambda f, r=req: planned.put((r, f)))

        read_done: queue.Queue = queue.Queue()

        def readback(g, lv):
            """Read a finished member back and retire it (read-back thread). The session's
            lock is held throughout: the serving loop skips the session meanwhile, so no
            step sequence is being captured on it while its events are waited for."""
            tr = time.perf_counter()
            err, res, retire_err = None, None, None
            with g.session.lock:
                try:
                    res = finish_result(g.session, lv.slot, lv.plan, lv.req, lv.total_steps,
                                        lv.rec_steps, 0)
                except Exception as exc:  # noqa: BLE001 - a failed solve fails alone
                    err = f"{type(exc).__name__}: {exc}"
                try:
                    g.session.remove(lv.slot)
                except Exception as exc:  # noqa: BLE001 - reported with the response
                    retire_err = exc
            read_done.put((g, lv, res, err, retire_err, time.perf_counter() - tr, now()))

        while len(responses) < n_total:
            # 1. arrivals -> planner pool
            t = now()
            while nxt < n_total and pending[nxt].arrival_s <= t:
                submit_plan(pending[nxt])
                in_flight_plans += 1
                nxt += 1
            # 2. planned -> waiting
            while True:
                try:
                    req, fut = planned.get_nowait()
                except queue.Empty:
                    break
                in_flight_plans -= 1
                try:
                    waiting.append((req, fut.result()))
                except Exception as exc:  # noqa: BLE001 - a bad request fails alone
                    fail(req, exc)
            # 3. groups whose chunk has drained: hand the members whose schedule is done to
            # the read-back threads (they hold their session slots until retired)
            finished = []
            for g in self.groups.values():
                if g.busy:
                    if not g.session.idle():
                        continue
                    g.busy = False
                for slot in [s_ for s_, lv in g.live.items() if lv.done_steps >= lv.total_steps]:
                    finished.append((g, g.live.pop(slot)))
            for g, lv in finished:
                g.reading += 1
                self.readers.submit(readback, g, lv)
            # retired members: their memory and slots are free again
            while True:
                try:
                    g, lv, res, err, retire_err, dt_read, t_done = read_done.get_nowait()
                except queue.Empty:
                    break
                g.reading -= 1
                self.in_use -= lv.plan.device_bytes()
                self.stats["readback_s"] += dt_read
                if retire_err is not None and err is None:
                    err = f"{type(retire_err).__name__}: {retire_err}"
                resp = SolveResponse(lv.req.rid, res, err, lv.req.arrival_s, lv.admit_s, t_done,
                                     tuple(lv.req.grid.shape), lv.peak_batch)
                responses.append(resp)
                finished.append(None)  # progress this round
                if on_done is not None:
                    on_done(resp)
            # 4. admission in arrival order, by HBM budget and free slots; a running group
            # takes new members too (their upload does not wait for its chunk, and they
            # join its next one)
            still = []
            for req, (plan, total, rec, need) in waiting:
                try:
                    g = self._group(plan)
                except Exception as exc:  # noqa: BLE001 - e.g. a shape without kernels
                    fail(req, exc)
                    continue
                pooled = self._pooled()
                if self.in_use + need + pooled > self.hbm_budget and pooled:
                    for x in self.groups.values():
                        if not x.busy and x.session.lock.acquire(blocking=False):
                            try:
                                x.session.trim()  # retired blocks give way
                            finally:
                                x.session.lock.release()
                    pooled = self._pooled()
                occupied = len(g.live) + g.reading
                cap = self.max_batch
                if self.max_launch_voxels is not None:
                    cap = max(1, min(cap, int(self.max_launch_voxels) // g.voxels))
                if occupied < cap and self.in_use + need + pooled <= self.hbm_budget:
                    if not g.session.lock.acquire(blocking=False):
                        still.append((req, (plan, total, rec, need)))  # being read back
                        continue
                    ta = time.perf_counter()
                    try:
                        slot = g.session.add(plan, max_periods=plan.eff_max + 4)
                        g.session.schedule(slot, total - rec)
                    except Exception as exc:  # noqa: BLE001 - allocation / upload failure
                        fail(req, exc)
                        continue
                    finally:
                        g.session.lock.release()
                        self.stats["admit_s"] += time.perf_counter() - ta
                    self.stats["admitted"] += 1
                    g.live[slot] = _Live(req, light_plan(plan), slot, total, rec, now())
                    self.in_use += need
                    self.peak_in_use = max(self.peak_in_use, self.in_use)
                else:
                    still.append((req, (plan, total, rec, need)))
            waiting = still
            # 5. queue the next chunk of every idle group with live members; groups of
            # different shapes run on their own streams, so their chunks overlap
            launched = False
            for g in self.groups.values():
                if g.busy or not g.live:
                    continue
                if not g.session.lock.acquire(blocking=False):
                    continue  # a member is being read back: refill on a later round
                budget = self.chunk_steps * (256 ** 3)
                k = max(1, budget // (g.voxels * len(g.live)))
                k = min(k, min(lv.total_steps - lv.done_steps for lv in g.live.values()))
                tl = time.perf_counter()
                try:
                    g.session.advance(list(g.live), k)
                    self.stats["chunks"] += 1
                    self.stats["launch_s"] += time.perf_counter() - tl
                except Exception as exc:  # noqa: BLE001 - the group's members fail together
                    for slot, lv in list(g.live.items()):
                        g.live.pop(slot)
                        self._retire(g, slot, lv)
                        fail(lv.req, exc, lv.admit_s)
                    continue
                finally:
                    g.session.lock.release()
                g.busy = True
                launched = True
                for lv in g.live.values():
                    lv.done_steps += k
                    lv.peak_batch = max(lv.peak_batch, len(g.live))
            if not launched and not finished:
                self.stats["idle_polls"] += 1
                if not any(g.busy for g in self.groups.values()) and nxt < n_total \
                        and not waiting and in_flight_plans == 0:
                    time.sleep(max(0.0, min(pending[nxt].arrival_s - now(), 0.002)))
                else:
                    if in_flight_plans and not any(g.busy or g.live for g in self.groups.values()):
                        self.stats["plan_wait_s"] += 0.0002  # the device waits for planning
                    time.sleep(0.0002)
        return responses

    def _retire(self, g: _Group, slot: int, lv: _Live) -> None:
        try:
            g.session.remove(slot)
        finally:
            self.in_use -= lv.plan.device_bytes()


def serve_serial(requests: list[SolveRequest], t0: float | None = None, on_done=None,
                 wait_for_arrivals: bool = True) -> list[SolveResponse]:
    """The plain form: one solve at a time through the solver API, in arrival order.

    With ``wait_for_arrivals=False`` the solves run back to back (each response's
    admit/done times are then its service interval, the input of ``replay_fifo``)."""
    import caustica_amd.solvers as solvers

    pending = sorted(requests, key=lambda r: r.arrival_s)
    t0 = time.perf_counter() if t0 is None else t0
    out = []
    for req in pending:
        wait = req.arrival_s - (time.perf_counter() - t0)
        if wait > 0 and wait_for_arrivals:
            time.sleep(wait)
        start = time.perf_counter() - t0
        err, res = None, None
        try:
            res = solvers.get(req.solver)().run(req.grid, req.medium, req.source, req.spec,
                                                backend="hip", harmonics=req.harmonics,
                                                reference_point=req.reference_point,
                                                record_region=req.record_region)
        except Exception as exc:  # noqa: BLE001
            err = f"{type(exc).__name__}: {exc}"
        resp = SolveResponse(req.rid, res, err, req.arrival_s, start, time.perf_counter() - t0,
                             tuple(req.grid.shape))
        out.append(resp)
        if on_done is not None:
            on_done(resp)
    return out


def replay_fifo(back_to_back: list[SolveResponse], arrivals: dict) -> list[SolveResponse]:
    """The plain form's timeline at other arrival times, from service times measured once.

    One solve at a time in arrival order is a single FIFO server: solve i starts at
    max(arrival_i, done_{i-1}) and takes its measured service time (done - admit of a
    back-to-back run). ``arrivals`` maps rid -> arrival time [s]."""
    order = sorted(back_to_back, key=lambda r: (arrivals[r.rid], r.rid))
    out, free_at = [], 0.0
    for r in order:
        a = arrivals[r.rid]
        start = max(a, free_at)
        done = start + (r.done_s - r.admit_s)
        free_at = done
        out.append(dataclasses.replace(r, arrival_s=a, admit_s=start, done_s=done))
    return out


def summarize(responses: list[SolveResponse]) -> dict:
    ok = [r for r in responses if r.error is None]
    lat = np.array([r.latency_s for r in ok]) if ok else np.zeros(1)
    first = min(r.arrival_s for r in responses)
    last = max(r.done_s for r in responses)
    return {
        "requests": len(responses),
        "errors": len(responses) - len(ok),
        "solves_per_s": len(ok) / max(last - first, 1e-9),
        "makespan_s": last - first,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "max_batch_seen": max((r.batch_peak for r in responses), default=1),
    }
