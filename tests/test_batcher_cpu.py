"""Host logic of the continuous batcher on the CPU, with the device session replaced.

A stand-in for KSpaceSession (no GPU) records what the batcher asks of it. The batcher
must turn every request it cannot serve into an error response and still answer every
other request: a footprint above the whole HBM budget (it must not wait forever for
room), a schedule that is not pinned (planning error), an allocation failure at
admission, a diverged field at read-back. Admission into a group whose chunk is in
flight is allowed: the new member joins the next chunk.
"""

import threading
import time

import numpy as np
import pytest

from caustica_amd import workloads
from caustica_amd.serving import batcher as B


class FakeSession:
    instances: list = []
    fail_add_for: set = set()      # member plans (by id) whose add() raises
    nan_for: set = set()           # plans whose peaks come back non-finite
    busy_polls = 0                 # idle() reports busy this many times per chunk
    slow_read: dict = {}           # padded shape -> seconds a read-back takes
    events: list = []              # (time, what, padded shape)

    def __init__(self, padded3, max_members=1, device=0):
        self.padded3 = tuple(padded3)
        self.max_members = max_members
        self.plans = {}
        self.next_slot = 0
        self.log = []
        self._busy = 0
        self.lock = threading.RLock()
        FakeSession.instances.append(self)

    def add(self, plan, max_periods=None):
        if id(plan) in FakeSession.fail_add_for:
            raise RuntimeError("hipMalloc: out of memory")
        if len(self.plans) >= self.max_members:  # as cst_member_add: every slot is taken
            raise RuntimeError("session is full")
        slot = self.next_slot
        self.next_slot += 1
        self.plans[slot] = plan
        self.log.append(("add", slot, self._busy > 0))
        return slot

    def schedule(self, slot, record_from):
        self.log.append(("schedule", slot, record_from))

    def advance(self, slots, k):
        self.log.append(("advance", tuple(slots), k))
        FakeSession.events.append((time.perf_counter(), "advance", self.padded3))
        self._busy = FakeSession.busy_polls

    def idle(self):
        if self._busy > 0:
            self._busy -= 1
            return False
        return True

    def peaks(self, slot, first, count):
        FakeSession.events.append((time.perf_counter(), "read start", self.padded3))
        time.sleep(FakeSession.slow_read.get(self.padded3, 0.0))
        FakeSession.events.append((time.perf_counter(), "read end", self.padded3))
        v = np.nan if id(self.plans[slot]) in FakeSession.nan_for else 1.0
        return np.full(count, v, np.float32)

    def result(self, slot, hi, scale, with_pmax=True, copy=True):
        n = int(np.prod(self.plans[slot].rec_shape))
        return np.ones(n, np.complex64), (np.ones(n, np.float32) if with_pmax else None)

    def remove(self, slot):
        self.plans.pop(slot)
        self.log.append(("remove", slot))

    def pooled_bytes(self):
        return 0

    def trim(self):
        return 0

    def close(self):
        pass


@pytest.fixture(autouse=True)
def fake_session(monkeypatch):
    FakeSession.instances = []
    FakeSession.fail_add_for = set()
    FakeSession.nan_for = set()
    FakeSession.busy_polls = 0
    FakeSession.slow_read = {}
    FakeSession.events = []
    monkeypatch.setattr(B, "KSpaceSession", FakeSession)
    yield


def _requests(n, spacing=0.0):
    rng = np.random.default_rng(3)
    return [workloads.serving_request(i, rng, spacing * i, shape=(48, 48, 64), solver="linear")
            for i in range(n)]


@pytest.mark.timeout(60)
def test_every_request_answered():
    reqs = _requests(4)
    b = B.ContinuousBatcher(hbm_budget=1 << 40, max_batch=2, planners=2)
    try:
        resp = b.serve(reqs)
    finally:
        b.close()
    assert sorted(r.rid for r in resp) == [0, 1, 2, 3]
    assert all(r.error is None for r in resp)
    assert b.in_use == 0


@pytest.mark.timeout(60)
def test_request_larger_than_the_budget_fails_alone():
    reqs = _requests(3)
    b = B.ContinuousBatcher(hbm_budget=1 << 20, max_batch=2, planners=2)  # below one member
    try:
        resp = b.serve(reqs)
    finally:
        b.close()
    assert len(resp) == 3
    for r in resp:
        assert r.result is None and "MemoryError" in r.error and "budget" in r.error


@pytest.mark.timeout(60)
def test_unpinned_schedule_fails_alone():
    reqs = _requests(3)
    reqs[1].spec = reqs[1].spec.model_copy(update={"convergence_tol": 0.01})
    b = B.ContinuousBatcher(hbm_budget=1 << 40, max_batch=2, planners=2)
    try:
        resp = {r.rid: r for r in b.serve(reqs)}
    finally:
        b.close()
    assert "pinned schedule" in resp[1].error
    assert resp[0].error is None and resp[2].error is None


@pytest.mark.timeout(60)
def test_admission_failure_and_divergence_fail_alone(monkeypatch):
    reqs = _requests(4)
    orig = B.ContinuousBatcher._plan

    def plan_and_mark(self, req):
        out = orig(self, req)
        if req.rid == 1:
            FakeSession.fail_add_for.add(id(out[0]))
        if req.rid == 2:
            FakeSession.nan_for.add(id(out[0]))
        return out

    monkeypatch.setattr(B.ContinuousBatcher, "_plan", plan_and_mark)
    b = B.ContinuousBatcher(hbm_budget=1 << 40, max_batch=4, planners=1)
    try:
        resp = {r.rid: r for r in b.serve(reqs)}
    finally:
        b.close()
    assert "out of memory" in resp[1].error
    assert "diverged" in resp[2].error
    assert resp[0].error is None and resp[3].error is None
    assert b.in_use == 0


@pytest.mark.timeout(60)
def test_admission_into_a_running_group():
    FakeSession.busy_polls = 50  # each chunk stays in flight for a while
    reqs = _requests(3, spacing=0.002)
    b = B.ContinuousBatcher(hbm_budget=1 << 40, max_batch=4, chunk_steps=1, planners=1)
    try:
        resp = b.serve(reqs)
    finally:
        b.close()
    assert all(r.error is None for r in resp)
    log = FakeSession.instances[0].log
    assert any(e[0] == "add" and e[2] for e in log), "no member was admitted while a chunk ran"
    assert max(len(e[1]) for e in log if e[0] == "advance") >= 2


def test_fifo_replay_of_the_plain_form():
    """A single FIFO server: start = max(arrival, previous done), done = start + service,
    checked against an event-by-event simulation with random arrivals and services."""
    rng = np.random.default_rng(0)
    n = 200
    service = rng.uniform(0.05, 0.5, n)
    back = []
    t = 0.0
    for i in range(n):
        back.append(B.SolveResponse(i, None, None, 0.0, t, t + service[i], (8, 8, 8)))
        t += service[i]
    arrivals = np.cumsum(rng.exponential(0.25, n))
    out = {r.rid: r for r in B.replay_fifo(back, dict(enumerate(arrivals)))}
    clock = 0.0
    for i in range(n):  # arrival order is rid order here
        clock = max(clock, arrivals[i]) + service[i]
        assert out[i].done_s == pytest.approx(clock, abs=1e-9)
        assert out[i].done_s - out[i].admit_s == pytest.approx(service[i], abs=1e-12)
        assert out[i].latency_s >= service[i] - 1e-12
    # below capacity the latency is the service time; saturated, solves/s = 1 / mean service
    sat = B.replay_fifo(back, {i: 0.0 for i in range(n)})
    assert B.summarize(sat)["solves_per_s"] == pytest.approx(n / service.sum(), rel=1e-9)


@pytest.mark.timeout(60)
def test_slots_of_retiring_members_are_not_reused_early():
    """A member whose schedule is done holds its session slot until it is read back and
    removed: admissions in the same loop turn must not count that slot as free."""
    FakeSession.busy_polls = 50  # chunks take a while: the queue is planned before anyone finishes
    reqs = _requests(9)
    b = B.ContinuousBatcher(hbm_budget=1 << 40, max_batch=2, chunk_steps=1, planners=4)
    try:
        resp = b.serve(reqs)
    finally:
        b.close()
    assert [r.error for r in resp if r.error] == []
    assert len(resp) == 9


@pytest.mark.timeout(60)
def test_other_shapes_advance_while_a_member_is_read_back():
    """Read-back runs off the serving loop: while a finished member of one shape is read
    back (slowly here), members of another shape keep receiving chunks."""
    rng = np.random.default_rng(5)
    a = workloads.serving_request(0, rng, 0.0, shape=(48, 48, 64), solver="linear")
    b = [workloads.serving_request(i, rng, 0.0, shape=(64, 64, 64), solver="linear") for i in (1, 2)]
    FakeSession.slow_read = {(48, 48, 64): 0.5}
    FakeSession.busy_polls = 20
    bat = B.ContinuousBatcher(hbm_budget=1 << 40, max_batch=4, planners=2, chunk_steps=1)
    try:
        resp = bat.serve([a] + b)
    finally:
        bat.close()
    assert sorted(r.rid for r in resp) == [0, 1, 2] and all(r.error is None for r in resp)
    t0 = next(t for t, what, shp in FakeSession.events if what == "read start" and shp == (48, 48, 64))
    t1 = next(t for t, what, shp in FakeSession.events if what == "read end" and shp == (48, 48, 64))
    during = [t for t, what, shp in FakeSession.events if what == "advance" and shp == (64, 64, 64) and t0 < t < t1]
    assert during, "the other shape's group got no chunk while the member was read back"
    assert bat.in_use == 0
