"""Served encrypted kNN on the GPU: the multi-query mulmod kernel against exact
integers, and the request loop (continuous batching, memory admission):
each query's result is the same alone and in the mix, bit for bit."""

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _index(ctx, N, limbs, xs):
    torch = _torch()
    from cuhepy_amd.engine import stride_for
    from cuhepy_amd.limbs import ints_to_limbs
    count = len(xs)
    stride = stride_for(count)
    dev = torch.device("cuda:0")
    src = torch.from_numpy(ints_to_limbs(xs, limbs).view(np.int32).copy()).to(dev)
    il = torch.zeros((limbs, stride), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ctx.aos_to_il(src, il, count, limbs, stride)
    ctx.sync()
    return il, stride


def _read(ctx, il, count, limbs, stride):
    torch = _torch()
    from cuhepy_amd.limbs import limbs_to_ints
    aos = torch.empty((count, limbs), dtype=torch.int32, device=il.device)
    torch.cuda.synchronize()
    ctx.il_to_aos(il, aos, count, limbs, stride)
    ctx.sync()
    return limbs_to_ints(aos.cpu().numpy().view(np.uint32))


def _operands(ctx, mod, ys):
    torch = _torch()
    from cuhepy_amd.limbs import ints_to_limbs
    dev = torch.device("cuda:0")
    y = torch.from_numpy(ints_to_limbs(ys, mod.limbs).view(np.int32).copy()).to(dev)
    ops = torch.zeros((len(ys), mod.operand_words), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    mod.operands_dev(y, ops, len(ys))
    ctx.sync()
    return [ops[i] for i in range(len(ys))]


def _modulus(bits, seed):
    rng = random.Random(seed)
    n = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
    return n * n


@pytest.mark.parametrize("nbits,limbs", [(512, 32), (1024, 64), (2048, 128)])
@pytest.mark.parametrize("nq", [1, 3, 32])
def test_mulmod_multi_exact(ctx, nbits, limbs, nq):
    torch = _torch()
    from cuhepy_amd.engine import Modulus
    N = _modulus(nbits, nbits + nq)
    mod = Modulus(ctx, N, limbs)
    rng = random.Random(nq * 7 + limbs)
    top = (1 << (32 * limbs)) - 1
    xs = [0, 1, N - 1, N, N + 1, top] + [rng.randrange(N) for _ in range(291)] + [rng.randrange(top) for _ in range(3)]
    ys = [rng.randrange(N) for _ in range(nq - 1)] + [top if nq > 1 else N - 1]
    il, stride = _index(ctx, N, limbs, xs)
    ops = _operands(ctx, mod, ys)
    begin, count = 5, len(xs) - 9  # a ragged range inside the index
    outs = [torch.full((limbs, stride), -1, dtype=torch.int32, device=il.device) for _ in ys]
    torch.cuda.synchronize()
    mod.mulmod_multi_dev(il, stride, begin, count, ops, outs, stride)
    ctx.sync()
    sentinel = (1 << (32 * limbs)) - 1
    for q, y in enumerate(ys):
        got = _read(ctx, outs[q], len(xs), limbs, stride)
        for i, x in enumerate(xs):
            want = x * y % N if begin <= i < begin + count else sentinel
            assert got[i] == want, (q, i)


def test_mulmod_multi_rejects_bad_arguments(ctx):
    torch = _torch()
    from cuhepy_amd.engine import Modulus
    N = _modulus(1024, 3)
    mod = Modulus(ctx, N, 64)
    il, stride = _index(ctx, N, 64, list(range(1, 100)))
    ops = _operands(ctx, mod, [5])
    out = torch.zeros((64, stride), dtype=torch.int32, device=il.device)
    with pytest.raises(ValueError):
        mod.mulmod_multi_dev(il, stride, 0, 10, [], [], stride)
    with pytest.raises(ValueError):
        mod.mulmod_multi_dev(il, stride, 0, 10, ops * 33, [out] * 33, stride)
    with pytest.raises(ValueError):
        mod.mulmod_multi_dev(il, stride, stride - 5, 10, ops, [out], stride)


def _server(ctx, mod, il, count, **kw):
    from cuhepy_amd.serve import KnnServer
    return KnnServer(ctx, mod, il, count, **kw)


def _active(srv):
    return len(srv.sched.active) + getattr(srv.sched, "running", 0)


def test_mulmod_multi_stamped(ctx):
    """The stamped launch gives the unstamped launch's products; every query
    gets a stamp of this launch and a full count of its tiles."""
    import torch
    from cuhepy_amd.engine import Modulus
    from cuhepy_amd.limbs import ints_to_limbs
    N = _modulus(2048, 21)
    mod = Modulus(ctx, N, 128)
    rng = random.Random(21)
    count = 60000
    xs = [rng.randrange(N) for _ in range(count)]
    il, stride = _index(ctx, N, 128, xs)
    ys = [rng.randrange(N) for _ in range(3)]
    dev = torch.device("cuda:0")
    ops = [torch.empty(mod.operand_words, dtype=torch.int32, device=dev) for _ in ys]
    for y, op in zip(ys, ops):
        mod.operands(ints_to_limbs([y], 128), op)
    outs_a = [torch.full((128, stride), -1, dtype=torch.int32, device=dev) for _ in ys]
    outs_b = [torch.full((128, stride), -1, dtype=torch.int32, device=dev) for _ in ys]
    done = torch.full((32,), 7, dtype=torch.int32, device=dev)
    stamps = torch.full((32,), 5, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()  # the fills above ran on torch's stream, the library uses its own
    ctx.sync()
    mod.mulmod_multi_dev(il, stride, 0, count, ops, outs_a, stride)
    mod.mulmod_multi_stamped_dev(il, stride, 0, count, ops, outs_b, stride, done, stamps)
    ctx.sync()
    for a, b in zip(outs_a, outs_b):
        assert torch.equal(a[:, :count], b[:, :count])
    tiles = (count + 63) // 64
    # counters complete; the queue head passed every unit; the words beyond untouched
    assert done[:3].tolist() == [tiles] * 3 and done[3].item() >= 3 * tiles and done[4:].tolist() == [7] * 28
    st = stamps[:3].tolist()
    # every query is stamped by this launch (the clock of its last unit, all
    # within the launch: well under 10 ms at 100 MHz); the words beyond are
    # untouched. At this size the queries' units overlap (each workgroup holds
    # two units at a time), so their order is not checked here; the bench's
    # full-index launches show it.
    assert all(v > 5 for v in st) and max(st) - min(st) < 1_000_000 and stamps[3:].tolist() == [5] * 29
    assert ctx.wall_clock_hz() > 1e6
    got = _read(ctx, outs_b[1], 64, 128, stride)
    assert got == [x * ys[1] % N for x in xs[:64]]


@pytest.mark.parametrize("mode", ["scan", "fifo"])
def test_server_alone_equals_mix(ctx, mode):
    """A query's result is identical served alone and admitted into a running mix."""
    from cuhepy_amd.engine import Modulus
    N = _modulus(2048, 11)
    mod = Modulus(ctx, N, 128)
    rng = random.Random(5)
    count = 2000
    xs = [rng.randrange(N) for _ in range(count)]
    il, stride = _index(ctx, N, 128, xs)
    ys = [rng.randrange(N) for _ in range(6)]
    srv = _server(ctx, mod, il, count, chunk=512, max_active=4, reserve_bytes=0, mode=mode)
    alone = srv.submit(ys[0], keep=True)
    srv.drain()
    ref = _read(ctx, alone.result, count, 128, stride)
    assert ref == [x * ys[0] % N for x in xs]
    # the mix: two queries start, the tested query joins while they run, more arrive after it
    reqs = [srv.submit(ys[1], keep=True), srv.submit(ys[2], keep=True)]
    srv.pump()
    srv.pump()
    mixed = srv.submit(ys[0], keep=True)
    srv.pump()
    reqs += [srv.submit(y, keep=True) for y in ys[3:]]
    srv.drain()
    if mode == "scan":
        assert mixed.start_chunk != 0  # joined mid-scan
        assert sorted(mixed.chunks_seen) == list(range(srv.nchunks))
    assert _read(ctx, mixed.result, count, 128, stride) == ref
    for r in reqs:
        got = _read(ctx, r.result, count, 128, stride)
        assert got == [x * r.y % N for x in xs]
    assert all(r.latency is not None and r.latency >= 0 for r in reqs + [mixed])


@pytest.mark.parametrize("mode", ["scan", "fifo"])
def test_server_memory_admission(ctx, mode):
    """With room for two result buffers, five queries still all complete,
    never more than two at once, reusing the pooled buffers."""
    from cuhepy_amd.engine import Modulus
    N = _modulus(1024, 12)
    mod = Modulus(ctx, N, 64)
    rng = random.Random(6)
    count = 700
    xs = [rng.randrange(N) for _ in range(count)]
    il, stride = _index(ctx, N, 64, xs)
    per = 64 * stride * 4
    state = {"free": 2 * per}
    srv = _server(ctx, mod, il, count, chunk=256, max_active=8, reserve_bytes=0, free_bytes=lambda: state["free"],
                  mode=mode)
    orig = srv._on_admit

    def on_admit(req, how):
        if how == "new":
            state["free"] -= per
        orig(req, how)

    srv._on_admit = on_admit
    ys = [rng.randrange(N) for _ in range(5)]
    reqs = [srv.submit(y) for y in ys]
    peak = 0
    while not srv.sched.idle() or srv.inflight:
        if not srv.pump() and srv.inflight:
            srv._retire(block=True)
        peak = max(peak, _active(srv))
    assert peak <= 2 and srv.admission.allocated == 2
    assert all(r.done is not None for r in reqs)


def test_server_hamming_end_to_end(ctx):
    """Index of encrypted binary vectors, encrypted query, served evaluation,
    client decode = plaintext Hamming distances (hamming/paillier_lookup.py flow)."""
    import json

    from cuhepy_amd.engine import Modulus
    from cuhepy_amd.limbs import limbs_to_ints
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    client = PaillierGPULookupClient(embed_len=64, key_len=512, rng=random.Random(31))
    n2 = int(json.loads(client.stringify_pk())["n_squared"])
    mod = Modulus(ctx, n2, 64)
    rng = np.random.default_rng(4)
    vecs = rng.integers(0, 2, size=(300, 64), dtype=np.uint8)
    cts = [int(c[0]) for c in client.encrypt(vecs.tolist())]
    il, stride = _index(ctx, n2, 64, cts)
    q = rng.integers(0, 2, size=64, dtype=np.uint8)
    y = int(client.encrypt([q.tolist()])[0][0])
    srv = _server(ctx, mod, il, len(cts), chunk=128, reserve_bytes=0)
    r = srv.submit(y, keep=True)
    srv.drain()
    enc = _read(ctx, r.result, len(cts), 64, stride)
    dist = client.decode_hamming_client([[c] for c in enc])
    assert dist == [int((v != q).sum()) for v in vecs]
