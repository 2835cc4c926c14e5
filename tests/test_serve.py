"""CPU: the serving scheduler (admission by memory, coalescing, slots, arrival order) with a fake
executor and a fake clock; no device involved."""
import numpy as np
import pytest

from gquery_amd.serve import Request, Scheduler, percentile, request_bytes, run_open_loop


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeExecutor:
    """Each launch takes `per_query` seconds per query on the fake clock."""

    def __init__(self, clock, dims=(2, 3), per_query=1e-6):
        self.clock, self.dims, self.per_query = clock, list(dims), per_query
        self.launches = []

    def launch(self, reqs, cls=0):
        h = {"reqs": list(reqs), "cls": cls, "end": self.clock() + self.per_query * sum(r.n for r in reqs)}
        self.launches.append(h)
        return h

    def done(self, h):
        return self.clock() >= h["end"]

    def finish(self, h):
        for r in h["reqs"]:
            r.result = ("ok", r.rid)


def req(rid, n, scene=0, kind="cp"):
    return Request(rid=rid, scene=scene, kind=kind, x=np.zeros((n, 2 if scene == 0 else 3), np.float32))


def test_pending_requests_are_coalesced_into_one_launch():
    clk = FakeClock()
    ex = FakeExecutor(clk)
    s = Scheduler(ex, memory_budget=1 << 30, slots=1, clock=clk)
    assert s.classes == {"cp": 0, "ray": 0, "sil": 0}
    for i in range(5):
        s.submit(req(i, 100))
    s.poll()
    assert len(ex.launches) == 1 and len(ex.launches[0]["reqs"]) == 5


def test_memory_budget_limits_admission_and_frees_on_completion():
    clk = FakeClock()
    ex = FakeExecutor(clk)
    one = request_bytes(req(0, 1000), 2)
    s = Scheduler(ex, memory_budget=2 * one + 10, slots=4, clock=clk)
    for i in range(3):
        s.submit(req(i, 1000))
    s.poll()
    assert [len(h["reqs"]) for h in ex.launches] == [2]   # the third does not fit
    assert s.in_use == 2 * one
    clk.t = 1.0
    s.poll()                                               # first launch retires, third admitted
    assert [len(h["reqs"]) for h in ex.launches] == [2, 1]
    assert sorted(r.rid for r in s.completed) == [0, 1]
    clk.t = 2.0
    s.poll()
    assert s.idle() and s.in_use == 0 and len(s.completed) == 3


def test_slots_bound_concurrent_launches():
    clk = FakeClock()
    ex = FakeExecutor(clk)
    s = Scheduler(ex, memory_budget=1 << 30, slots=2, max_batch_queries=100, clock=clk)
    for i in range(6):
        s.submit(req(i, 100))
    s.poll()
    assert len(ex.launches) == 2 and len(s.inflight) == 2 and len(s.pending[0]) == 4


def test_arrival_order_no_overtaking():
    clk = FakeClock()
    ex = FakeExecutor(clk)
    small = request_bytes(req(0, 10), 2)
    big = req(1, 100000)
    s = Scheduler(ex, memory_budget=request_bytes(big, 2) + small - 1, slots=4, clock=clk)
    s.submit(req(0, 10))
    s.poll()
    s.submit(big)
    s.submit(req(2, 10))
    s.poll()
    # the big request does not fit beside 0; the small one behind it waits although it would fit
    assert [[r.rid for r in h["reqs"]] for h in ex.launches] == [[0]]
    assert [r.rid for r in s.pending[0]] == [1, 2]
    clk.t = 1.0
    s.poll()
    assert [[r.rid for r in h["reqs"]] for h in ex.launches] == [[0], [1]]
    clk.t = 2.0
    s.poll()
    assert [[r.rid for r in h["reqs"]] for h in ex.launches] == [[0], [1], [2]]


def test_oversized_request_is_rejected():
    clk = FakeClock()
    s = Scheduler(FakeExecutor(clk), memory_budget=1000, clock=clk)
    with pytest.raises(MemoryError):
        s.submit(req(0, 10000))
    with pytest.raises(ValueError):
        Scheduler(FakeExecutor(clk), memory_budget=0)


def test_open_loop_latencies_on_a_fake_clock():
    clk = FakeClock()
    ex = FakeExecutor(clk, per_query=1e-3)
    s = Scheduler(ex, memory_budget=1 << 30, slots=1, clock=clk)
    reqs = [req(i, 10) for i in range(4)]
    arrivals = [0.0, 0.001, 0.05, 0.051]

    def tick(dt):
        clk.t += max(dt, 1e-4)

    # drive the fake clock forward inside the loop
    orig_poll = s.poll

    def poll():
        n = orig_poll()
        clk.t += 1e-3
        return n

    s.poll = poll
    run_open_loop(s, reqs, arrivals, clock=clk, sleep=tick)
    lat = [r.latency for r in reqs]
    assert all(l is not None and l > 0 for l in lat)
    assert all(r.result == ("ok", r.rid) for r in reqs)
    assert percentile(lat, 50) <= percentile(lat, 99)


def test_silhouettes_have_their_own_queue_and_slots():
    clk = FakeClock()
    ex = FakeExecutor(clk, per_query=1e-3)
    s = Scheduler(ex, memory_budget=1 << 30, slots=(1, 1), max_batch_queries=1000, clock=clk)
    s.submit(req(0, 1000, kind="sil"))     # long
    s.submit(req(1, 10, kind="cp"))
    s.submit(req(2, 10, kind="sil"))
    s.submit(req(3, 10, kind="ray"))
    s.poll()
    # one launch per class: the silhouette batch does not block the closest-point / ray one
    assert sorted((h["cls"], tuple(r.rid for r in h["reqs"])) for h in ex.launches) == [(0, (1, 3)), (1, (0,))]
    clk.t = 0.5
    s.poll()
    assert sorted(r.rid for r in s.completed) == [1, 3]
    clk.t = 1.1
    s.poll()
    clk.t = 2.0
    s.poll()
    assert sorted(r.rid for r in s.completed) == [0, 1, 2, 3]


class SizedExecutor(FakeExecutor):
    """An executor that reports each request's device footprint (as GpuExecutor does from
    gq_batch_sizes): 8 bytes per query of records plus a fixed 1 MB of workspace."""

    def footprint(self, r):
        return 8 * r.n + (1 << 20)


def test_admission_charges_the_footprint_the_executor_reports():
    clk = FakeClock()
    ex = SizedExecutor(clk)
    # room for exactly three reported footprints of 1000-query requests; the size-only estimate of
    # the same requests is far smaller, so admitting by it would take all ten at once
    fp = 8 * 1000 + (1 << 20)
    assert request_bytes(req(0, 1000), 2) * 10 < 3 * fp
    s = Scheduler(ex, memory_budget=3 * fp + 100, slots=4, clock=clk)
    for i in range(10):
        s.submit(req(i, 1000))
    assert all(r.footprint == fp for q in s.pending for r in q)
    s.poll()
    assert [len(h["reqs"]) for h in ex.launches] == [3] and s.in_use == 3 * fp
    clk.t = 1.0
    s.poll()
    assert s.in_use == 3 * fp and len(s.completed) == 3   # the first three freed, the next three in
    with pytest.raises(MemoryError):
        s.submit(req(99, 10 ** 6))                          # reported footprint above the whole budget
