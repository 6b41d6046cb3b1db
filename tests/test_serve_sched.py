"""Served-kNN host logic on the CPU: FIFO and cyclic-scan continuous batching
and memory-based admission (no device)."""

import random

import pytest

from cuhepy_amd.serve import (FIFO_MAX_ACTIVE, MQ_MAX, FifoBatcher, MemoryAdmission, Request, ScanScheduler, latency_stats,
                               poisson_arrivals)


def _drive(sched, adm, arrivals_per_iter, iters, release_finished=True):
    """Submit arrivals_per_iter[t] requests before iteration t; return finished."""
    done, rid = [], 0
    admitted = []
    for t in range(iters):
        for _ in range(arrivals_per_iter[t] if t < len(arrivals_per_iter) else 0):
            sched.submit(Request(rid, 0, float(t)))
            rid += 1
        sched.admit(adm, float(t), lambda r, how: admitted.append((r.rid, how)))
        if not sched.active:
            continue
        _, batch, fin = sched.next_iteration()
        assert len(batch) <= sched.max_active
        for r in fin:
            done.append(r)
            if release_finished:
                adm.release()
    return done, admitted


def test_every_request_sees_every_range_once():
    rng = random.Random(1)
    sched = ScanScheduler(nchunks=5, max_active=4)
    adm = MemoryAdmission(lambda: 1 << 40, request_bytes=100)
    arrivals = [rng.randrange(3) for _ in range(40)]
    done, _ = _drive(sched, adm, arrivals, 400)
    assert len(done) == sum(arrivals)
    for r in done:
        assert sorted(r.chunks_seen) == list(range(5))
        # consecutive ranges of the cyclic scan starting where the request joined
        assert r.chunks_seen == [(r.start_chunk + k) % 5 for k in range(5)]
    assert sched.idle()


def test_fifo_admission_and_slot_cap():
    sched = ScanScheduler(nchunks=3, max_active=2)
    adm = MemoryAdmission(lambda: 1 << 40, request_bytes=1)
    for i in range(5):
        sched.submit(Request(i, 0, 0.0))
    order = []
    sched.admit(adm, 0.0, lambda r, how: order.append(r.rid))
    assert order == [0, 1] and len(sched.active) == 2
    done, _ = _drive(sched, adm, [], 30)
    assert [r.rid for r in done][:2] == [0, 1]
    assert sorted(order + [r.rid for r in done[2:]]) == [0, 1, 2, 3, 4]


def test_memory_admission_limits_and_reuses_buffers():
    free = {"b": 350}
    sched = ScanScheduler(nchunks=2, max_active=MQ_MAX)
    # each request needs 100 bytes over a 50-byte reserve: 3 fit in 350 bytes
    adm = MemoryAdmission(lambda: free["b"], request_bytes=100, reserve=50)

    def on_admit(r, how):
        if how == "new":
            free["b"] -= 100

    for i in range(6):
        sched.submit(Request(i, 0, 0.0))
    sched.admit(adm, 0.0, on_admit)
    assert len(sched.active) == 3 and len(sched.waiting) == 3
    assert adm.held == 3 and adm.allocated == 3
    # finish the first three; their buffers are pooled and reused, nothing new is allocated
    for _ in range(2):
        _, _, fin = sched.next_iteration()
    for _ in range(3):
        adm.release()
    hows = []
    sched.admit(adm, 1.0, lambda r, how: hows.append(how))
    assert hows == ["pool"] * 3 and adm.allocated == 3 and free["b"] == 50


def test_admission_waits_when_memory_is_short():
    sched = ScanScheduler(nchunks=1, max_active=4)
    adm = MemoryAdmission(lambda: 10, request_bytes=100)
    sched.submit(Request(0, 0, 0.0))
    assert sched.admit(adm, 0.0, lambda r, how: None) == 0
    assert len(sched.waiting) == 1 and not sched.active


def test_bad_parameters():
    with pytest.raises(ValueError):
        ScanScheduler(0)
    with pytest.raises(ValueError):
        ScanScheduler(4, max_active=MQ_MAX + 1)
    with pytest.raises(RuntimeError):
        ScanScheduler(2).next_iteration()


def test_poisson_and_percentiles():
    rng = random.Random(7)
    ts = poisson_arrivals(100.0, 20000, rng)
    assert all(b > a for a, b in zip(ts, ts[1:]))
    assert abs(len(ts) / ts[-1] - 100.0) < 3.0  # mean rate within 3%
    reqs = []
    for i in range(100):
        r = Request(i, 0, 0.0)
        r.done = (i + 1) / 1000.0
        reqs.append(r)
    s = latency_stats(reqs)
    assert s["n"] == 100 and abs(s["p50_ms"] - 50.0) < 1e-9 and abs(s["p99_ms"] - 99.0) < 1e-9


def test_fifo_batcher_launches_in_arrival_order():
    """Every request is launched exactly once, in arrival order; a launch takes
    every admitted request not yet launched; at most max_active are active."""
    rng = random.Random(3)
    sched = FifoBatcher(max_active=5)
    adm = MemoryAdmission(lambda: 1 << 40, request_bytes=1)
    rid, launched, inflight = 0, [], []
    for step in range(300):
        for _ in range(rng.randrange(4) if step < 200 else 0):
            sched.submit(Request(rid, 0, float(step)))
            rid += 1
        sched.admit(adm, float(step), lambda r, how: None)
        assert len(sched.pending) + sched.running <= 5
        if len(inflight) < 2 and sched.pending:
            batch = sched.next_batch()
            assert batch and not sched.pending
            launched += [r.rid for r in batch]
            inflight.append(batch)
        if inflight and rng.random() < 0.6:  # the oldest launch completes
            done = inflight.pop(0)
            sched.complete(done)
            for _ in done:
                adm.release()
    while inflight:
        sched.complete(inflight.pop(0))
    while not sched.idle():  # drain
        sched.admit(adm, 0.0, lambda r, how: None)
        b = sched.next_batch()
        launched += [r.rid for r in b]
        sched.complete(b)
        for _ in b:
            adm.release()
    assert launched == list(range(rid))
    assert sched.idle() and adm.held == 0


def test_fifo_batcher_memory_and_bad_parameters():
    free = {"b": 250}
    sched = FifoBatcher(max_active=MQ_MAX)
    adm = MemoryAdmission(lambda: free["b"], request_bytes=100, reserve=0)

    def on_admit(r, how):
        if how == "new":
            free["b"] -= 100

    for i in range(4):
        sched.submit(Request(i, 0, 0.0))
    assert sched.admit(adm, 0.0, on_admit) == 2 and len(sched.waiting) == 2
    b = sched.next_batch()
    assert [r.rid for r in b] == [0, 1] and sched.running == 2
    sched.complete(b)
    adm.release()
    adm.release()
    hows = []
    sched.admit(adm, 1.0, lambda r, how: hows.append(how))
    assert hows == ["pool", "pool"] and adm.allocated == 2
    with pytest.raises(ValueError):
        FifoBatcher(0)
    with pytest.raises(ValueError):
        FifoBatcher(FIFO_MAX_ACTIVE + 1)


def test_fifo_batcher_caps_a_launch_and_prepares_the_next():
    """More admitted requests than one launch holds: each launch takes the
    oldest MQ_MAX, the rest stay admitted (prepared) for the next launch."""
    sched = FifoBatcher(max_active=3 * MQ_MAX)
    adm = MemoryAdmission(lambda: 1 << 40, request_bytes=1)
    for i in range(3 * MQ_MAX + 5):
        sched.submit(Request(i, 0, 0.0))
    assert sched.admit(adm, 0.0, lambda r, how: None) == 3 * MQ_MAX and len(sched.waiting) == 5
    b1 = sched.next_batch()
    b2 = sched.next_batch()
    assert [r.rid for r in b1] == list(range(MQ_MAX)) and [r.rid for r in b2] == list(range(MQ_MAX, 2 * MQ_MAX))
    assert len(sched.pending) == MQ_MAX and sched.running == 2 * MQ_MAX
    assert sched.admit(adm, 0.0, lambda r, how: None) == 0  # at the cap: the rest waits
    sched.complete(b1)
    assert sched.admit(adm, 0.0, lambda r, how: None) == 5
    b3 = sched.next_batch()
    assert [r.rid for r in b3] == list(range(2 * MQ_MAX, 3 * MQ_MAX))
    assert [r.rid for r in sched.next_batch()] == list(range(3 * MQ_MAX, 3 * MQ_MAX + 5))
