"""Served encrypted kNN: a resident Hamming index, continuous batching, memory admission.

The evaluator is encode_hamming_server (hamming/paillier.py:214-230, over the
extensions' pointwise mulmod: paillier_GPU_client.cu:986-991,
paillier_GPU_lookup_client.cu:2312-2329): a query ciphertext y is multiplied
into every index ciphertext x_i mod n^2, giving the encrypted Hamming
distances that the client decrypts. The reference runs one call per query
from Python; here a request loop keeps the index resident in HBM and serves
many queries at once. Two batching modes:

* ``fifo`` (the default): admitted queries run in arrival order. Whenever the
  device queue has room (``depth`` launches in flight), ONE launch takes the
  admitted queries not yet launched (up to ``MQ_MAX``) over the whole index.
  Inside the launch the workgroups take the tiles query by query, so the
  queries complete one after another (each overlapping only the start of the
  next), with no gap between them as between launches, and the launch pays
  one tail instead of one per query. A launch of several queries stamps the
  device clock when each query's last tile is done
  (``Modulus.mulmod_multi_stamped_dev``); the stamps are converted to host
  time against the event recorded after the launch. A launch of one query is
  the plain launch (completion = its event). Queries arriving while the queue
  is full are admitted and prepared (operand tables) up to ``max_active``
  (which may exceed one launch's ``MQ_MAX``), so a prepared batch is ready
  when a slot frees.
* ``scan``: a cyclic scan over ``nchunks`` index ranges. Every iteration
  multiplies the next range by all active queries; a query arriving while
  others run joins at the next iteration, at whatever range the scan is on,
  and completes after ``nchunks`` iterations (processor sharing).

The evaluator is bound by the integer multiply-add rate (32 multiply-adds per
byte), so a query costs the same in a batch as alone. With equal work per
query, FIFO order gives the lowest latency; processor sharing (``scan``)
makes every active query wait for all the others. ``fifo`` keeps each
query's own completion time and removes the gaps between launches.

Memory admission: each active query holds a result buffer (one ciphertext per
index entry). A waiting query is admitted only when a released buffer can be
reused, or free device memory minus a reserve covers a new one. At most
``max_active`` are active (``MQ_MAX`` operands per launch; fifo mode admits
up to ``FIFO_MAX_ACTIVE``, scan mode ``MQ_MAX``, since every active query of a
scan joins each launch).

Latency is measured to device completion from the request's arrival time.
The schedulers and the admission policy are plain Python, independent of the
device, so their invariants are tested on the CPU.
"""

from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

MQ_MAX = 32  # operands per launch (MultiQuery in the kernel)
FIFO_MAX_ACTIVE = 4 * MQ_MAX  # admitted (launched or waiting for a launch) in fifo mode


@dataclass
class Request:
    rid: int
    y: int                      # query ciphertext (below n^2)
    arrival: float              # host clock, seconds
    keep: bool = False          # keep the result buffer for the caller
    start_chunk: int = -1
    left: int = 0
    admitted: Optional[float] = None
    launched: Optional[float] = None  # host clock when its (last) launch was queued
    done: Optional[float] = None  # device completion, host clock
    result: object = None         # (limbs, stride) device tensor when kept
    chunks_seen: list = field(default_factory=list)

    @property
    def latency(self) -> Optional[float]:
        return None if self.done is None else self.done - self.arrival


class MemoryAdmission:
    """Admit a request when a pooled result buffer is free or free device
    memory minus `reserve` covers a new one. `free_bytes` is injectable."""

    def __init__(self, free_bytes: Callable[[], int], request_bytes: int, reserve: int = 0):
        self.free_bytes = free_bytes
        self.request_bytes = int(request_bytes)
        self.reserve = int(reserve)
        self.pooled = 0     # released buffers ready for reuse
        self.held = 0       # buffers owned by active requests
        self.allocated = 0  # buffers ever allocated and not trimmed

    def try_reserve(self) -> Optional[str]:
        """'pool' (reuse), 'new' (allocate one) or None (wait)."""
        if self.pooled > 0:
            self.pooled -= 1
            self.held += 1
            return "pool"
        if self.free_bytes() - self.reserve >= self.request_bytes:
            self.held += 1
            self.allocated += 1
            return "new"
        return None

    def release(self) -> None:
        assert self.held > 0
        self.held -= 1
        self.pooled += 1


class ScanScheduler:
    """Continuous batching over a cyclic scan of `nchunks` index ranges."""

    def __init__(self, nchunks: int, max_active: int = MQ_MAX):
        if nchunks < 1:
            raise ValueError("nchunks must be >= 1")
        if not 1 <= max_active <= MQ_MAX:
            raise ValueError(f"max_active must be in [1, {MQ_MAX}]")
        self.nchunks = nchunks
        self.max_active = max_active
        self.waiting: collections.deque[Request] = collections.deque()
        self.active: list[Request] = []
        self.cursor = 0

    def submit(self, req: Request) -> None:
        self.waiting.append(req)

    def admit(self, admission: MemoryAdmission, now: float, on_admit: Callable[[Request, str], None]) -> int:
        """FIFO admission while slots and memory allow; returns how many joined."""
        n = 0
        while self.waiting and len(self.active) < self.max_active:
            how = admission.try_reserve()
            if how is None:
                break
            req = self.waiting.popleft()
            req.start_chunk, req.left, req.admitted = self.cursor, self.nchunks, now
            on_admit(req, how)
            self.active.append(req)
            n += 1
        return n

    def next_iteration(self) -> tuple[int, list[Request], list[Request]]:
        """(range, queries in this launch, queries finishing with it)."""
        if not self.active:
            raise RuntimeError("no active request")
        chunk = self.cursor
        batch = list(self.active)
        finished = []
        for r in batch:
            r.chunks_seen.append(chunk)
            r.left -= 1
            if r.left == 0:
                finished.append(r)
        self.active = [r for r in self.active if r.left > 0]
        self.cursor = (self.cursor + 1) % self.nchunks
        return chunk, batch, finished

    def idle(self) -> bool:
        return not self.waiting and not self.active


class FifoBatcher:
    """FIFO continuous batching: admitted requests are launched in arrival
    order; a launch takes the admitted requests not yet launched, at most
    MQ_MAX; at most `max_active` are admitted (launched or not)."""

    def __init__(self, max_active: int = 2 * MQ_MAX):
        if not 1 <= max_active <= FIFO_MAX_ACTIVE:
            raise ValueError(f"max_active must be in [1, {FIFO_MAX_ACTIVE}]")
        self.max_active = max_active
        self.waiting: collections.deque[Request] = collections.deque()
        self.pending: list[Request] = []   # admitted, not launched
        self.running = 0                   # launched, not complete

    def submit(self, req: Request) -> None:
        self.waiting.append(req)

    def admit(self, admission: MemoryAdmission, now: float, on_admit: Callable[[Request, str], None],
              limit: int = FIFO_MAX_ACTIVE) -> int:
        n = 0
        while self.waiting and len(self.pending) + self.running < self.max_active and n < limit:
            how = admission.try_reserve()
            if how is None:
                break
            req = self.waiting.popleft()
            req.admitted = now
            on_admit(req, how)
            self.pending.append(req)
            n += 1
        return n

    def next_batch(self) -> list[Request]:
        batch, self.pending = self.pending[:MQ_MAX], self.pending[MQ_MAX:]
        self.running += len(batch)
        return batch

    def complete(self, batch: list[Request]) -> None:
        self.running -= len(batch)
        assert self.running >= 0

    @property
    def active(self) -> list[Request]:
        return self.pending

    def idle(self) -> bool:
        return not self.waiting and not self.pending and self.running == 0


class KnnServer:
    """Resident encrypted index + a scheduler driving the multi-query kernel.

    index: device tensor (limbs, stride) int32, interleaved ciphertexts under
    modulus n^2 (a ``Modulus``); `count` entries. mode "fifo": one launch
    covers the whole index for a batch of queries in arrival order; mode
    "scan": one launch covers one range of `chunk` entries for all active
    queries.
    """

    def __init__(self, ctx, modulus, index, count: int, chunk: Optional[int] = None, max_active: int = MQ_MAX,
                 reserve_bytes: int = 4 << 30, free_bytes: Optional[Callable[[], int]] = None, depth: int = 2,
                 clock: Callable[[], float] = time.perf_counter, mode: str = "fifo", prep_per_pump: int = 4):
        import torch

        if mode not in ("fifo", "scan"):
            raise ValueError("mode must be 'fifo' or 'scan'")
        self.torch = torch
        self.mode = mode
        self.ctx, self.mod, self.index = ctx, modulus, index
        self.limbs = modulus.limbs
        self.count = int(count)
        self.stride = int(index.shape[1])
        if index.shape[0] != self.limbs or self.stride < self.count:
            raise ValueError("index must be (limbs, stride >= count)")
        self.chunk = self.count if mode == "fifo" else int(chunk or self.count)
        self.nchunks = math.ceil(self.count / self.chunk)
        self.sched = FifoBatcher(max_active) if mode == "fifo" else ScanScheduler(self.nchunks, max_active)
        self.device = torch.device(f"cuda:{ctx.device}")
        self.result_bytes = self.limbs * self.stride * 4
        if free_bytes is None:
            free_bytes = lambda: torch.cuda.mem_get_info(self.device)[0]  # noqa: E731
        self.admission = MemoryAdmission(free_bytes, self.result_bytes, reserve_bytes)
        self.pool: list = []
        self.depth = depth
        self.prep_per_pump = max(1, int(prep_per_pump))
        self.prepared = 0
        self.clock = clock
        self.stream = torch.cuda.ExternalStream(ctx.stream_ptr, device=self.device) if ctx.stream_ptr else \
            torch.cuda.default_stream(self.device)
        self.inflight: collections.deque = collections.deque()  # (event, finished requests)
        self.completed: list[Request] = []
        self.iterations = 0
        self.launched_pairs = 0
        self._next_id = 0
        ctx.sync()
        torch.cuda.synchronize(self.device)
        self.base_event = torch.cuda.Event(enable_timing=True)
        self._anchor()
        # query ciphertexts are staged in pinned host rows, so the copy that
        # feeds a query's operand preparation is asynchronous (a pageable
        # source would make the host wait for the launches in flight); a row
        # is reused once its query retires
        import numpy as np
        rows = (depth + 1) * MQ_MAX + max_active
        self.y_host = torch.empty((rows, self.limbs), dtype=torch.int32, pin_memory=True)
        self.y_np = self.y_host.numpy().view(np.uint32)
        self.y_free = list(range(rows))
        if mode == "fifo":
            # per-launch completion counters and stamps, reused once retired;
            # each set has a pinned host copy of its stamps that the stream
            # fills right after the launch, so a retire reads host memory and
            # never waits for the device (a blocking read of device memory
            # would stall the host behind the launches still in flight)
            self.clock_hz = ctx.wall_clock_hz()
            with torch.cuda.stream(self.stream):
                self.stamp_sets = [(torch.zeros(MQ_MAX + 1, dtype=torch.int32, device=self.device),
                                    torch.zeros(MQ_MAX, dtype=torch.int64, device=self.device),
                                    torch.zeros(MQ_MAX, dtype=torch.int64, pin_memory=True))
                                   for _ in range(depth + 1)]

    def _anchor(self) -> None:
        """Pair a device timestamp with the host clock. Completion times are
        device event times mapped to the host clock through this pair; the two
        clocks drift apart by tens of parts per million, so the pair is renewed
        whenever the server is idle before a trace (an error of a millisecond
        after a minute would otherwise enter every latency)."""
        self.base_event.record(self.stream)
        self.base_event.synchronize()
        self.t_base = self.clock()

    # ---- requests ----
    def submit(self, y: int, arrival: Optional[float] = None, keep: bool = False) -> Request:
        if not 0 <= y < (1 << (32 * self.limbs)):
            raise ValueError("query ciphertext does not fit the modulus width")
        req = Request(self._next_id, int(y), self.clock() if arrival is None else arrival, keep)
        self._next_id += 1
        self.sched.submit(req)
        return req

    def _on_admit(self, req: Request, how: str) -> None:
        from .limbs import ints_to_limbs
        torch = self.torch
        with torch.cuda.stream(self.stream):  # allocations belong to the stream the kernels use
            buf = self.pool.pop() if how == "pool" else torch.empty((self.limbs, self.stride), dtype=torch.int32,
                                                                     device=self.device)
            op = torch.empty((self.mod.operand_words,), dtype=torch.int32, device=self.device)
        # the query's operand is prepared on the library's stream, behind the
        # iterations in flight and ahead of the next launch: no synchronisation
        row = self.y_free.pop()
        self.y_np[row] = ints_to_limbs([req.y], self.limbs)[0]
        self.mod.operands(self.y_np[row:row + 1], op)
        req._buf, req._op, req._row = buf, op, row

    # ---- loop ----
    def _retire(self, block: bool) -> int:
        n = 0
        while self.inflight:
            ev, finished, stamps = self.inflight[0]
            if not block and not ev.query():
                break
            ev.synchronize()
            self.inflight.popleft()
            t = self.t_base + self.base_event.elapsed_time(ev) / 1e3
            if stamps is not None:  # fifo batch: each query completed at its own stamp, the last one at the event
                self.stamp_sets.append(stamps)
                st = stamps[2][:len(finished)].tolist()
                last = max(st)
                for r, s_q in zip(finished, st):
                    r.done = t - (last - s_q) / self.clock_hz
            if self.mode == "fifo":
                self.sched.complete(finished)
            for r in finished:
                if stamps is None:
                    r.done = t
                if r.keep:
                    r.result = r._buf
                else:
                    self.pool.append(r._buf)
                    self.admission.release()
                self.y_free.append(r._row)
                r._buf = r._op = r._row = None
                self.completed.append(r)
            n += 1
            block = False
        return n

    def release(self, req: Request) -> None:
        """Return a kept result buffer to the pool."""
        if req.result is not None:
            self.pool.append(req.result)
            req.result = None
            self.admission.release()

    def pump(self) -> bool:
        """Retire finished iterations, admit, and launch one iteration if the
        queue depth allows; returns whether anything was launched
        (``self.prepared``: how many queries this call admitted without a
        launch, which a caller keeps pumping for before it waits)."""
        self._retire(block=False)
        self.prepared = 0
        if len(self.inflight) >= self.depth:
            if self.mode == "fifo":
                # prepare a few of the queries that will share the next launch,
                # then return, so the caller sees a finished launch soon
                self.prepared = self.sched.admit(self.admission, self.clock(), self._on_admit,
                                                 limit=self.prep_per_pump)
            return False
        if self.mode == "fifo":
            # the device queue has room: launch what is prepared now; with
            # nothing prepared, admit the next query and launch it alone
            # (queries admitted while the queue is full leave as one batch)
            if not self.sched.pending:
                self.sched.admit(self.admission, self.clock(), self._on_admit, limit=1)
        else:
            self.sched.admit(self.admission, self.clock(), self._on_admit)
        if not self.sched.active:
            return False
        if self.mode == "fifo":
            batch = finished = self.sched.next_batch()
            begin, cnt = 0, self.count
            if len(batch) > 1:
                stamps = self.stamp_sets.pop()
                self.mod.mulmod_multi_stamped_dev(self.index, self.stride, begin, cnt, [r._op for r in batch],
                                                  [r._buf for r in batch], self.stride, stamps[0], stamps[1])
                with self.torch.cuda.stream(self.stream):
                    stamps[2].copy_(stamps[1], non_blocking=True)  # complete when the launch's event is
            else:  # one query: the plain launch, complete at its event
                stamps = None
                self.mod.mulmod_multi_dev(self.index, self.stride, begin, cnt, [batch[0]._op], [batch[0]._buf],
                                          self.stride)
        else:
            chunk, batch, finished = self.sched.next_iteration()
            begin = chunk * self.chunk
            cnt = min(self.chunk, self.count - begin)
            stamps = None
            self.mod.mulmod_multi_dev(self.index, self.stride, begin, cnt, [r._op for r in batch],
                                      [r._buf for r in batch], self.stride)
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        now = self.clock()
        for r in batch:
            r.launched = now
        self.inflight.append((ev, finished, stamps))
        self.iterations += 1
        self.launched_pairs += len(batch) * cnt
        return True

    def drain(self) -> None:
        while not self.sched.idle() or self.inflight:
            if not self.pump() and not self.prepared:
                if self.inflight:
                    self._retire(block=True)
                elif self.sched.waiting and not self.sched.active:
                    raise RuntimeError("admission cannot make progress: not enough device memory for one query")

    def run_trace(self, arrivals: list[tuple[float, int]], keep: Callable[[int], bool] = lambda i: False) -> list:
        """Replay (offset seconds, query ciphertext) arrivals against the host
        clock; returns the completed requests in arrival order."""
        if self.sched.idle() and not self.inflight:
            self._anchor()
        t0 = self.clock()
        reqs = []
        i = 0
        while i < len(arrivals) or not self.sched.idle() or self.inflight:
            now = self.clock()
            while i < len(arrivals) and t0 + arrivals[i][0] <= now:
                reqs.append(self.submit(arrivals[i][1], t0 + arrivals[i][0], keep(i)))
                i += 1
            launched = self.pump()
            if not launched and not self.prepared:
                if self.inflight:
                    # the device is busy: wait for the oldest iteration unless an arrival is due first
                    nxt = t0 + arrivals[i][0] if i < len(arrivals) else math.inf
                    if nxt > self.clock():
                        ev = self.inflight[0][0]
                        while not ev.query() and self.clock() < nxt:
                            time.sleep(20e-6)
                elif i < len(arrivals):
                    wait = t0 + arrivals[i][0] - self.clock()
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
        return reqs


def poisson_arrivals(rate: float, n: int, rng) -> list[float]:
    """n arrival offsets (seconds) of a Poisson process of `rate` per second."""
    t, out = 0.0, []
    for _ in range(n):
        t += rng.expovariate(rate)
        out.append(t)
    return out


def latency_stats(reqs: list[Request]) -> dict:
    lat = sorted(r.latency for r in reqs if r.latency is not None)
    if not lat:
        return {"n": 0}

    def pct(p):
        k = min(len(lat) - 1, max(0, math.ceil(p / 100 * len(lat)) - 1))
        return lat[k]

    return {"n": len(lat), "p50_ms": pct(50) * 1e3, "p99_ms": pct(99) * 1e3, "mean_ms": sum(lat) / len(lat) * 1e3,
            "max_ms": lat[-1] * 1e3}
