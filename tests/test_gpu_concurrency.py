"""The C ABI's threading contract (include/chx.h): calls on one context are
serialised by its lock, different contexts run concurrently, and ctypes
releases the GIL around every call. Four Python threads encrypt and decrypt
their own batches on one shared context, then on two contexts; every result
equals the serial run's. A key used with another context is refused.
One functional run, not a stress test."""

import random
import threading

import pytest

from oracle import paillier_oracle as O

pytestmark = pytest.mark.gpu


def _batches(n, seed, count=4, size=300):
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        ms = [rng.randrange(n) for _ in range(size)]
        rs = [O.random_r(n, rng) for _ in range(size)]
        out.append((ms, rs))
    return out


def _run_threads(jobs):
    results = [None] * len(jobs)
    errors = []

    def work(i, fn):
        try:
            results[i] = fn()
        except Exception as e:  # pragma: no cover - reported below
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=work, args=(i, fn)) for i, fn in enumerate(jobs)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    return results


def test_threads_on_one_and_two_contexts():
    from cuhepy_amd.engine import Context, PaillierKey
    from cuhepy_amd.limbs import limbs_to_ints
    keys = O.key_gen(512, 909)
    n, phi, inv, p, q = keys["pk"]["n"], keys["sk"]["phi"], keys["sk"]["inv"], keys["p"], keys["q"]
    ctx_a = Context.default(0)
    ctx_b = Context(0)
    key_b = None
    try:
        key_a = PaillierKey(ctx_a, n, phi, inv, p, q)
        key_a_plain = PaillierKey(ctx_a, n, phi, inv)
        key_b = PaillierKey(ctx_b, n, phi, inv, p, q)
        batches = _batches(n, 11)

        def job(key, ms, rs):
            def fn():
                out = []
                for _ in range(3):
                    c = limbs_to_ints(key.encrypt(ms, rs))
                    out.append((c, limbs_to_ints(key.decrypt(c))))
                return out
            return fn

        serial = [job(key_a, ms, rs)() for ms, rs in batches]
        for (ms, rs), runs in zip(batches, serial):
            want = [O.encrypt(m, keys["pk"], r) for m, r in zip(ms[:8], rs[:8])]
            assert runs[0][0][:8] == want and runs[0][1] == ms
        # one context, four threads (two with the factorless key: other kernels interleave)
        got = _run_threads([job(key_a if i % 2 == 0 else key_a_plain, ms, rs) for i, (ms, rs) in enumerate(batches)])
        assert got == serial
        # two contexts, two threads each
        got = _run_threads([job(key_a if i < 2 else key_b, ms, rs) for i, (ms, rs) in enumerate(batches)])
        assert got == serial
        # a key belongs to the context it was created on
        with pytest.raises(ValueError, match="another context"):
            import numpy as np
            from cuhepy_amd import _native as N
            from cuhepy_amd.engine import _p32
            from cuhepy_amd.limbs import ints_to_limbs
            c = ints_to_limbs(serial[0][0][0][:4], key_b.width)
            out = np.empty_like(c)
            N.check(N.lib.chx_paillier_decrypt(ctx_a.handle, key_b.handle, _p32(c), _p32(out), 4))
    finally:
        if key_b is not None:
            key_b.close()  # before its context
        ctx_b.close()
