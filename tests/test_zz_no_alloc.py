"""chx_paillier_decrypt_crt_noalloc against chx_paillier_decrypt_crt_dev,
through the C ABI: equal outputs (and equal to the oracle's plaintexts) on
two calls in a row, the second with a smaller batch; the context's scratch
slots do not grow; a batch beyond the fixed array is refused."""

import ctypes
import random

import numpy as np
import pytest

from oracle import paillier_oracle as O

pytestmark = pytest.mark.gpu


def _dev(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("key_len,counts", [(1024, (2000, 301)), (512, (4000, 77))])
def test_noalloc_equals_dev_entry_point(key_len, counts):
    import torch
    from cuhepy_amd import _native as N
    from cuhepy_amd.engine import Context, PaillierKey
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    keys = O.key_gen(key_len, 4100 + key_len)
    n, phi, inv, p, q = keys["pk"]["n"], keys["sk"]["phi"], keys["sk"]["inv"], keys["p"], keys["q"]
    ctx = Context(0)  # fresh: its scratch slots start empty
    key = None
    try:
        key = PaillierKey(ctx, n, phi, inv, p, q)
        W, Hl = key.width, key.half
        dev = torch.device("cuda:0")
        rng = random.Random(key_len)
        for count in counts:
            print(f"noalloc: key_len {key_len} count {count}", flush=True)
            ms = [rng.randrange(n) for _ in range(count)]
            rs = [O.random_r(n, rng) for _ in range(count)]
            cs = limbs_to_ints(key.encrypt(ms, rs))
            assert cs[:4] == [O.encrypt(m, keys["pk"], r) for m, r in zip(ms[:4], rs[:4])]
            stride = (count + 63) // 64 * 64
            il = np.zeros((W, stride), dtype=np.uint32)
            il[:, :count] = ints_to_limbs(cs, W).T
            c_d = torch.from_numpy(il.view(np.int32)).to(dev)
            want_d = torch.full((Hl, stride), -1, dtype=torch.int32, device=dev)
            got_d = torch.full((Hl, stride), -1, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            N.check(N.lib.chx_paillier_decrypt_crt_dev(ctx.handle, key.handle, _dev(c_d), _dev(want_d), count, stride))
            ctx.sync()
            slots = ctx.scratch_bytes()
            N.check(N.lib.chx_paillier_decrypt_crt_noalloc(ctx.handle, key.handle, _dev(c_d), _dev(got_d), count,
                                                           stride))
            ctx.sync()
            assert ctx.scratch_bytes() == slots  # no slot of the context grew
            want = want_d.cpu().numpy().view(np.uint32)
            got = got_d.cpu().numpy().view(np.uint32)
            assert np.array_equal(got[:, :count], want[:, :count])
            assert limbs_to_ints(np.ascontiguousarray(got[:, :count].T)) == ms
        # a batch whose scratch passes the 64 MiB array is refused before any launch
        big = 20000 if key_len == 1024 else 60000
        stride = (big + 63) // 64 * 64
        c_big = torch.zeros((W, stride), dtype=torch.int32, device=dev)
        o_big = torch.zeros((Hl, stride), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        with pytest.raises(ValueError, match="allocation-free"):
            N.check(N.lib.chx_paillier_decrypt_crt_noalloc(ctx.handle, key.handle, _dev(c_big), _dev(o_big), big,
                                                           stride))
    finally:
        if key is not None:
            key.close()
        ctx.close()
