"""Client-level bit-exactness at the benchmark's own shapes.

* The lookup client at C3's configuration (embed_len 512, key_len 1024, so
  n^2 of 4096 bits, alpha 280, 256 message chunks of which packed 512-bit
  embeddings use rows 0-127 with digits in the packing alphabet): 256
  Bernoulli(0.5) embeddings through the row-quint (with noise triples),
  row-quad, row-triple, row-pair and single-row tables, each ciphertext
  equal to lookup.py:213-252 (the oracle) for the same tables and noise
  indices; quints at every noise count's remainder, and packed digits above
  the rows the client's group tables cover.
* Multi-chunk embeddings (embed_len > key_len, hamming/paillier.py:26-32):
  chunk_num 2 and 3 for both clients; every chunk's ciphertext equal to the
  oracle's encryption of the oracle's packing with the same randomness, and
  encrypt -> encode_hamming_server -> decode_hamming_client equal to the
  oracle's decode of the decrypted sums. A chunk of 2 key_len bits is as wide
  as n, so m_a + m_b may pass n and wrap (the reference's documented PL-05,
  docs/research/paillier-security.md:104-112: no embed_len < key_len guard);
  where no chunk sum wraps, the decode is the plaintext Hamming distance.
"""

import json
import random

import numpy as np
import pytest

from oracle import paillier_oracle as O

pytestmark = pytest.mark.gpu

OP_LOOKUP = 2


def _lookup_oracle_keys(client):
    pk, sk = json.loads(client.stringify_pk()), json.loads(client.stringify_sk())
    n = int(pk["n"])
    return {"pk": {"g": int(pk["g"]), "n": n, "n_squared": n * n, "g_n": int(pk["g_n"])},
            "sk": {"phi": int(sk["phi"]), "a": int(sk["a"]), "g_a_inv": int(sk["g_a_inv"])}}


def _tables(client):
    t = client.dump_tables()
    gt = {int(i): [int(v) for v in row] for i, row in t["g_table"].items()}
    return gt, [int(v) for v in t["noise_table"]]


def _expected_decodes(embs, embed_len, key_len, n):
    """(oracle decode of the decrypted chunk sums, pairs whose sums stay below n)."""
    chunk_len, _ = O.chunk_geometry(embed_len, key_len)
    want, exact = [], []
    for i in range(len(embs) - 1):
        ma = O.embed_to_plaintexts(embs[i], embed_len, key_len)
        mb = O.embed_to_plaintexts(embs[i + 1], embed_len, key_len)
        want.append(O.decode_hamming([(x + y) % n for x, y in zip(ma, mb)], chunk_len))
        if all(x + y < n for x, y in zip(ma, mb)):
            exact.append(i)
    return want, exact


def test_lookup_client_c3_shape_every_row_group(ctx):
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    client = PaillierGPULookupClient(embed_len=512, key_len=1024, alpha_len=280, rng=random.Random(31))
    assert client.chunk_num == 1
    keys = _lookup_oracle_keys(client)
    n = keys["pk"]["n"]
    gt, nt = _tables(client)
    mc = json.loads(client.stringify_config())["message_chunks"]
    assert mc == 256 and keys["sk"]["a"].bit_length() >= 278
    rng = random.Random(5)
    embs = [[rng.getrandbits(1) for _ in range(512)] for _ in range(254)] + [[0] * 512, [1] * 512]
    ms = [O.embed_to_plaintexts(e, 512, 1024)[0] for e in embs]
    idx = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
    idx[0] = 0        # every noise factor from entry 0
    idx[1] = 255      # ... and from the last entry
    want = [O.lookup_encrypt(m, keys["pk"], gt, nt, mc, ix.tolist()) for m, ix in zip(ms, idx)]
    m_limbs = ints_to_limbs(ms, 128)
    W = 128
    try:
        for rows in (0, 5, 4, 3, 2, 1):  # 0: the default (quads); 5 (with noise triples), 4, 3, 2, 1 forced
            ctx.set_tuning(OP_LOOKUP, W, 0, rows)
            got = limbs_to_ints(client.encrypt_limbs(m_limbs, noise_idx=idx))
            bad = [i for i, (a, b) in enumerate(zip(got, want)) if a != b]
            assert not bad, f"rows {rows}: {len(bad)} mismatches, first at {bad[:4]}"
        # quints: noise counts where the triples leave 1, 2 or 3 factors, or none are taken
        ctx.set_tuning(OP_LOOKUP, W, 0, 5)
        n2 = keys["pk"]["n_squared"]
        for nm in (13, 12, 4, 3, 2, 1, 0):
            sub = idx[:64, :nm]
            got = limbs_to_ints(client.encrypt_limbs(m_limbs[:64], noise_idx=sub))
            # the oracle takes 14 factors (lookup.py:213-252): divide out the omitted ones
            exp = []
            for c, ix in zip(want[:64], idx[:64]):
                for j in ix[nm:]:
                    c = c * pow(nt[int(j)], -1, n2) % n2
                exp.append(c)
            assert got == exp, f"quints with {nm} noise factors"
        # digits in the packing alphabet above the rows the client's tables
        # cover (rows 127-254: a packed value shifted up 127 bytes) go singly
        hi = [m << (8 * 127) for m in ms[:64]]
        assert max(hi) < n
        got = limbs_to_ints(client.encrypt_limbs(ints_to_limbs(hi, 128), noise_idx=idx[:64]))
        assert got == [O.lookup_encrypt(m, keys["pk"], gt, nt, mc, ix.tolist()) for m, ix in zip(hi, idx[:64])]
    finally:
        ctx.set_tuning(OP_LOOKUP, W, 0, 0)
    # the reference surface (bits in, injected noise through the client's rng)
    client._rng = random.Random(8)
    clone = random.Random(8)
    got = [int(c[0]) for c in client.encrypt(embs[:16])]
    idx16 = np.frombuffer(clone.randbytes(16 * 14), dtype=np.uint8).reshape(16, 14).tolist()
    assert got == [O.lookup_encrypt(m, keys["pk"], gt, nt, mc, ix) for m, ix in zip(ms[:16], idx16)]
    assert client.decrypt_cts(got) == ms[:16]


@pytest.mark.parametrize("embed_len,key_len", [(400, 256), (600, 256), (1500, 512)])
def test_plain_client_multi_chunk(embed_len, key_len):
    from cuhepy_amd.paillier._gpu_ext import PaillierGPUClient
    client = PaillierGPUClient(embed_len=embed_len, key_len=key_len, rng=random.Random(embed_len))
    chunk_len, chunk_num = O.chunk_geometry(embed_len, key_len)
    assert (client.chunk_len, client.chunk_num) == (chunk_len, chunk_num) and chunk_num >= 2
    pk = json.loads(client.stringify_pk())
    n = int(pk["n"])
    keys = {"pk": {"g": n + 1, "n": n, "n_squared": n * n}}
    rng = random.Random(embed_len + 1)
    embs = [[rng.getrandbits(1) for _ in range(embed_len)] for _ in range(9)]
    embs.append([1] * embed_len)
    client._rng = random.Random(77)
    clone = random.Random(77)
    cts = client.encrypt(embs)
    assert len(cts) == len(embs) and all(len(row) == chunk_num for row in cts)
    for e, row in zip(embs, cts):
        ms = O.embed_to_plaintexts(e, embed_len, key_len)
        want = [O.encrypt(m, keys["pk"], O.random_r(n, clone)) for m in ms]
        assert [int(c, 16) for c in row] == want
    encs = [PaillierGPUClient.encode_hamming_server(cts[i], cts[i + 1], {"n_squared": format(n * n, "x")})
            for i in range(len(embs) - 1)]
    got = client.decode_hamming_client(encs)
    want, exact = _expected_decodes(embs, embed_len, key_len, n)
    assert got == want and len(exact) >= 4
    assert [got[i] for i in exact] == [O.hamming(embs[i], embs[i + 1]) for i in exact]
    assert client.decode_hamming_client(cts[-1:]) == [embed_len]  # weight of the all-ones embedding


@pytest.mark.parametrize("embed_len,key_len", [(400, 256), (600, 256), (1500, 512)])
def test_lookup_client_multi_chunk(embed_len, key_len):
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    client = PaillierGPULookupClient(embed_len=embed_len, key_len=key_len, rng=random.Random(embed_len))
    chunk_len, chunk_num = O.chunk_geometry(embed_len, key_len)
    assert (client.chunk_len, client.chunk_num) == (chunk_len, chunk_num) and chunk_num >= 2
    keys = _lookup_oracle_keys(client)
    gt, nt = _tables(client)
    mc = json.loads(client.stringify_config())["message_chunks"]
    rng = random.Random(embed_len + 2)
    embs = [[rng.getrandbits(1) for _ in range(embed_len)] for _ in range(9)]
    embs += [[0] * embed_len, [1] * embed_len]  # the pairs next to zeros never wrap
    client._rng = random.Random(78)
    clone = random.Random(78)
    cts = client.encrypt(embs)
    assert len(cts) == len(embs) and all(len(row) == chunk_num for row in cts)
    # the client draws the batch's noise indices in one call, row-major (ciphertext, factor)
    idx = iter(np.frombuffer(clone.randbytes(len(embs) * chunk_num * 14), dtype=np.uint8).reshape(-1, 14).tolist())
    for e, row in zip(embs, cts):
        ms = O.embed_to_plaintexts(e, embed_len, key_len)
        want = [O.lookup_encrypt(m, keys["pk"], gt, nt, mc, next(idx)) for m in ms]
        assert [int(c) for c in row] == want
    pk = json.loads(client.stringify_pk())
    encs = [PaillierGPULookupClient.encode_hamming_server(cts[i], cts[i + 1], pk) for i in range(len(embs) - 1)]
    got = client.decode_hamming_client(encs)
    want, exact = _expected_decodes(embs, embed_len, key_len, keys["pk"]["n"])
    assert got == want and {8, 9} <= set(exact)
    assert [got[i] for i in exact] == [O.hamming(embs[i], embs[i + 1]) for i in exact]
    assert client.decode_hamming_client(cts[-1:]) == [embed_len]
