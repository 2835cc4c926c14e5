"""Serving loop: resident scenes, continuous batching of query requests, admission by free memory.

Requests (scene, kind, points[, directions]) arrive at any time.  The loop admits pending
requests in arrival order while their device footprint (inputs, records and their share of the
batch workspace) fits the memory budget and a launch slot is free, coalesces everything admitted
at once into one batch (gq_batch_enqueue: one sort pass and one traversal launch per query group),
and enqueues it on the next of `slots` streams, so launches that overlap in time overlap on the
device.  Requests that arrive while launches run wait only until a slot and memory free up.
Completion is detected per launch with a stream event; a request's latency is completion minus
arrival.  Each result equals the single-batch call on the same inputs bit for bit (queries are
independent and every kernel breaks ties to the lowest primitive slot), whatever it was batched
with.

The scheduling (`Scheduler`) is plain Python over an executor interface, so its admission and
ordering rules are tested on the CPU with a fake executor; `GpuExecutor` is the HIP one.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Any, Callable, List, Optional


@dataclasses.dataclass
class Request:
    rid: int
    scene: int                 # index into the server's scenes
    kind: str                  # "cp", "ray" or "sil"
    x: Any                     # (n, dim) device tensor
    v: Any = None              # rays: (n, dim) directions
    arrival: float = 0.0       # seconds (perf_counter clock)
    admitted: Optional[float] = None
    done: Optional[float] = None
    result: Any = None
    footprint: Optional[int] = None   # device bytes charged while in flight (set on submission)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def latency(self) -> Optional[float]:
        return None if self.done is None else self.done - self.arrival


# bytes per query of the records of each kind (valid, p, n, uv, d, prim_id, on_boundary), by dim
def record_bytes(kind: str, dim: int) -> int:
    if kind == "sil":
        return 1 + 4 * dim + 4 + 4
    return 1 + 4 * dim + 4 * dim + 4 * (dim - 1) + 4 + 4 + (1 if kind == "ray" else 0)


def request_bytes(req: Request, dim: int) -> int:
    """A size-only estimate of a request's device footprint while in flight (inputs, records and
    48 bytes of scratch per query), for executors that cannot ask the library; GpuExecutor charges
    the records plus the workspace gq_batch_sizes reports instead (GpuExecutor.footprint)."""
    inputs = 4 * dim * req.n * (2 if req.kind == "ray" else 1)
    return inputs + record_bytes(req.kind, dim) * req.n + 48 * req.n + 4096


# Requests are scheduled in cost classes with their own queues and launch slots, so a long
# silhouette batch never holds up the closest-point and ray requests queued behind it.
DEFAULT_CLASSES = {"cp": 0, "ray": 0, "sil": 1}


class Scheduler:
    """Admission and coalescing.  `executor` provides launch(requests, cls) -> handle,
    done(handle) -> bool, finish(handle) (stores results in the requests), and dims[scene].
    Within a class, requests are admitted in arrival order (no overtaking); a launch takes the
    queue head and every following request of the class that fits the memory budget, up to
    max_batch_queries queries."""

    def __init__(self, executor, memory_budget: int, slots=(2, 1), max_batch_queries: int = 1 << 20,
                 classes=None, clock: Callable[[], float] = time.perf_counter):
        slots = (slots,) if isinstance(slots, int) else tuple(slots)
        if memory_budget <= 0 or not slots or min(slots) <= 0:
            raise ValueError("memory_budget and slots must be positive")
        self.ex = executor
        self.budget = int(memory_budget)
        self.slots = slots
        self.classes = dict(classes) if classes is not None else (
            DEFAULT_CLASSES if len(slots) > 1 else {k: 0 for k in DEFAULT_CLASSES})
        self.max_batch_queries = int(max_batch_queries)
        self.clock = clock
        self.pending = [collections.deque() for _ in slots]
        self.inflight: List[tuple] = []     # (handle, requests, bytes, class)
        self.in_use = 0
        self.completed: List[Request] = []

    def footprint(self, req: Request) -> int:
        """Device bytes charged for a request: the executor's figure where it has one (records plus
        the workspace the library reports), else the size-only estimate.  Computed once per request."""
        if req.footprint is None:
            f = getattr(self.ex, "footprint", None)
            req.footprint = int(f(req)) if f is not None else request_bytes(req, self.ex.dims[req.scene])
        return req.footprint

    def submit(self, req: Request):
        fp = self.footprint(req)
        if fp > self.budget:
            raise MemoryError(f"request {req.rid} needs {fp} bytes, more than the whole budget {self.budget}")
        if req.kind not in self.classes:
            raise ValueError(f"unknown query kind {req.kind!r}")
        if not req.arrival:
            req.arrival = self.clock()
        self.pending[self.classes[req.kind]].append(req)

    def poll(self) -> int:
        """Retire finished launches, then admit and launch what fits.  Returns #completed now."""
        n_done = 0
        still = []
        for h, reqs, b, c in self.inflight:
            if self.ex.done(h):
                self.ex.finish(h)
                now = self.clock()
                for r in reqs:
                    r.done = now
                self.completed.extend(reqs)
                self.in_use -= b
                n_done += len(reqs)
            else:
                still.append((h, reqs, b, c))
        self.inflight = still
        for c, q in enumerate(self.pending):
            while q and sum(1 for f in self.inflight if f[3] == c) < self.slots[c]:
                batch, used, nq = [], 0, 0
                while q:
                    r = q[0]
                    fp = self.footprint(r)
                    if self.in_use + used + fp > self.budget or (batch and nq + r.n > self.max_batch_queries):
                        break
                    batch.append(q.popleft())
                    used += fp
                    nq += r.n
                if not batch:
                    break
                now = self.clock()
                for r in batch:
                    r.admitted = now
                h = self.ex.launch(batch, c)
                self.inflight.append((h, batch, used, c))
                self.in_use += used
        return n_done

    def idle(self) -> bool:
        return not any(self.pending) and not self.inflight


class GpuExecutor:
    """Launches a list of requests as one gquery_amd.batch.Batch on a free stream slot of its class.
    A request is charged its records plus the batch workspace gq_batch_sizes reports for it alone
    (the coalesced batch needs at most the sum: it shares the job and tile tables).  Workspaces and
    pinned staging buffers are pooled and reused once the launch that last used them has completed."""

    def __init__(self, shapes, slots=(2, 1), method: str = "engine", device=None):
        import torch
        self.torch = torch
        self.shapes = list(shapes)
        self.dims = [getattr(s, "scene", s).dim for s in self.shapes]
        self.method = method
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        slots = (slots,) if isinstance(slots, int) else tuple(slots)
        # one stream per launch slot of each class, default priority (no priority argument)
        self.streams = [[torch.cuda.Stream(self.device) for _ in range(n)] for n in slots]
        self._stream_ev = [[None] * n for n in slots]   # last launch's completion event per stream
        self._turn = [0] * len(slots)
        self._buffers = []   # [workspace, staging, completion event of the last launch using them]
        self.launches = 0
        self.buffer_allocations = 0

    def footprint(self, req) -> int:
        from .batch import Batch
        dev, _ = Batch.sizes([(self.shapes[req.scene], req.kind, req.x, req.v)], method=self.method)
        return record_bytes(req.kind, self.dims[req.scene]) * req.n + dev

    def _take_buffers(self, dev_bytes: int, host_bytes: int):
        """A pooled (workspace, staging) pair no launch is using, grown if too small."""
        torch = self.torch
        free = [b for b in self._buffers if b[2] is None or b[2].query()]
        for b in free:
            if b[0].numel() >= dev_bytes and b[1].numel() >= host_bytes:
                return b
        if free:
            b = free[0]
        else:
            b = [None, None, None]
            self._buffers.append(b)
        if b[0] is None or b[0].numel() < dev_bytes:
            b[0] = torch.empty(max(dev_bytes, 1), dtype=torch.uint8, device=self.device)
            self.buffer_allocations += 1
        if b[1] is None or b[1].numel() < host_bytes:
            b[1] = torch.empty(max(host_bytes, 1), dtype=torch.uint8, pin_memory=True)
            self.buffer_allocations += 1
        return b

    def launch(self, reqs, cls: int = 0):
        from .batch import Batch
        torch = self.torch
        pool = self.streams[cls]
        # a stream of the class whose previous launch has finished, else the next in turn
        k = next((i for i, e in enumerate(self._stream_ev[cls]) if e is None or e.query()),
                 self._turn[cls] % len(pool))
        self._turn[cls] += 1
        st = pool[k]
        self.launches += 1
        jobs = [(self.shapes[r.scene], r.kind, r.x, r.v) for r in reqs]
        dev, host = Batch.sizes(jobs, method=self.method)
        buf = self._take_buffers(dev, host)
        # launches overlap on the device whenever there is more than one slot: mark the jobs shared
        # so small launch groups keep the full-launch packet budgets (less total work)
        shared = int(sum(len(p) for p in self.streams) > 1)
        jobs = [dict(shape=s, kind=k, x=x, v=v, shared=shared) for s, k, x, v in jobs]
        # the request inputs were written on the default stream: order this launch after them
        st.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(st):
            b = Batch(jobs, method=self.method, device=self.device, workspace=buf[0], staging=buf[1])
            b.run()
            ev = torch.cuda.Event()
            ev.record(st)
        buf[2] = ev
        self._stream_ev[cls][k] = ev
        return (b, ev, reqs)

    def done(self, h) -> bool:
        return h[1].query()

    def finish(self, h):
        b, ev, reqs = h
        for r, res in zip(reqs, b.results):
            r.result = res


def percentile(values, q):
    vals = sorted(values)
    if not vals:
        return None
    k = (len(vals) - 1) * q / 100.0
    lo, hi = int(k), min(int(k) + 1, len(vals) - 1)
    return vals[lo] + (vals[hi] - vals[lo]) * (k - lo)


def run_open_loop(sched: Scheduler, requests, arrivals, clock=time.perf_counter, sleep=None):
    """Feeds `requests` at offsets `arrivals` (seconds from now) and runs the loop until all are
    done.  Returns the requests with their latencies filled in."""
    t0 = clock()
    i = 0
    order = sorted(range(len(requests)), key=lambda k: arrivals[k])
    while i < len(order) or not sched.idle():
        now = clock()
        while i < len(order) and t0 + arrivals[order[i]] <= now:
            r = requests[order[i]]
            r.arrival = t0 + arrivals[order[i]]
            sched.submit(r)
            i += 1
        sched.poll()
        if sleep is not None and sched.idle() and i < len(order):
            sleep(max(0.0, min(t0 + arrivals[order[i]] - clock(), 1e-4)))
    return requests
