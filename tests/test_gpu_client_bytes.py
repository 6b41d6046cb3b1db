"""The little-endian bytes forms of both clients (the convention of
hamming/paillier.py:101, 162 and of the msgpack packets of
benchmarks/bfv_client_matrix.py:49-69: minimal LE bytes per ciphertext) and
the batched host conversions behind every client method:

* encrypt_bytes -> encode_hamming_server_bytes -> decode_hamming_client over
  bytes gives the plaintext Hamming distance, and every ciphertext equals the
  integer/hex form's and the oracle's (scheme.py:63-83, lookup.py:213-252)
  under the same injected randomness;
* encrypt_cts_bytes / decrypt_cts_bytes round trips, out-of-range inputs
  refused as by the integer forms;
* numpy embeddings give the same ciphertexts as nested lists;
* the batched r draw (no injected generator): r in [1, n-1], gcd(r, n) = 1
  checked through the product of the batch mod n (``Modulus.product``,
  against Python's product), and a batch holding a multiple of p is repaired.
"""

import json
import math
import random

import numpy as np
import pytest

from oracle import paillier_oracle as O

pytestmark = pytest.mark.gpu

E, KL = 96, 512


def _le(v: int) -> bytes:
    return v.to_bytes(max(1, (v.bit_length() + 7) // 8), "little")


def _embs(seed, count):
    rng = random.Random(seed)
    out = [[rng.getrandbits(1) for _ in range(E)] for _ in range(count)]
    return out + [[0] * E, [1] * E]


def test_plain_client_bytes_path():
    from cuhepy_amd.paillier._gpu_ext import PaillierGPUClient
    client = PaillierGPUClient(embed_len=E, key_len=KL, rng=random.Random(4))
    n = int(json.loads(client.stringify_pk())["n"])
    pk = {"n_squared": format(n * n, "x")}  # the class takes hex, as hamming/paillier.py:224-228 passes it
    keys = {"pk": {"g": n + 1, "n": n, "n_squared": n * n}}
    embs = _embs(6, 14)
    client._rng = random.Random(90)
    by = client.encrypt_bytes(embs)
    client._rng = random.Random(90)
    hx = client.encrypt(np.array(embs, dtype=np.uint8))
    clone = random.Random(90)
    want = [O.encrypt(O.embed_to_plaintexts(e, E, KL)[0], keys["pk"], O.random_r(n, clone)) for e in embs]
    assert [[int.from_bytes(c, "little") for c in row] for row in by] == [[w] for w in want]
    assert [[int(c, 16) for c in row] for row in hx] == [[w] for w in want]
    assert all(c == _le(int.from_bytes(c, "little")) for row in by for c in row)  # minimal length
    encs = [PaillierGPUClient.encode_hamming_server_bytes(by[i], by[i + 1], pk) for i in range(len(embs) - 1)]
    encs_hex = [PaillierGPUClient.encode_hamming_server(hx[i], hx[i + 1], pk) for i in range(len(embs) - 1)]
    assert [[int.from_bytes(c, "little") for c in r] for r in encs] == [[int(c, 16) for c in r] for r in encs_hex]
    dist = [O.hamming(embs[i], embs[i + 1]) for i in range(len(embs) - 1)]
    assert client.decode_hamming_client(encs) == dist
    assert client.decode_hamming_client(encs_hex) == dist
    assert client.decode_hamming_client([[int.from_bytes(c, "little") for c in r] for r in encs]) == dist
    # scheme-level bytes round trip, zero and n - 1 included
    rng = random.Random(8)
    ms = [0, 1, n - 1] + [rng.randrange(n) for _ in range(40)]
    cts = client.encrypt_cts_bytes([_le(m) for m in ms])
    assert client.decrypt_cts_bytes(cts) == [_le(m) for m in ms]
    assert client.decrypt_cts([int.from_bytes(c, "little") for c in cts]) == ms
    with pytest.raises(ValueError, match="plaintext"):
        client.encrypt_cts_bytes([_le(n)])
    with pytest.raises(ValueError, match="ciphertext"):
        client.decrypt_cts_bytes([_le(n * n)])
    with pytest.raises(ValueError, match="ciphertext"):
        client.decrypt_cts_bytes([b"\x01" * 600])
    assert client.encrypt_bytes([]) == [] and client.decrypt_cts_bytes([]) == []


def test_plain_client_batched_r_draws(ctx):
    from cuhepy_amd.engine import Modulus
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    from cuhepy_amd.paillier import keygen as kg
    from cuhepy_amd.paillier._gpu_ext import PaillierGPUClient
    client = PaillierGPUClient(embed_len=E, key_len=KL, rng=random.Random(5))
    sk = json.loads(client.stringify_sk())
    n = int(json.loads(client.stringify_pk())["n"])
    p, q = kg.factor_from_phi(n, int(sk["phi"]))
    client._rng = None  # the batched draw
    rng = random.Random(3)
    ms = [rng.randrange(n) for _ in range(300)]
    c1, c2 = client.encrypt_cts(ms), client.encrypt_cts(ms)
    assert client.decrypt_cts(c1) == ms and client.decrypt_cts(c2) == ms
    assert len(set(c1) | set(c2)) == 600  # fresh randomness per ciphertext
    # the product of a batch mod n, against Python's
    mod = Modulus(ctx, n)
    vals = [rng.randrange(1, n) for _ in range(37)]
    assert mod.product(ints_to_limbs(vals, mod.limbs)) == math.prod(vals) % n
    assert mod.product(ints_to_limbs(vals[:1], mod.limbs)) == vals[0]
    # a batch holding multiples of p and q: the product is no unit and those rows are redrawn
    bad = [rng.randrange(1, n) for _ in range(50)]
    bad[7], bad[31] = p * 3, q * 5
    W = 2 * mod.limbs
    assert math.gcd(mod.product(ints_to_limbs(bad, mod.limbs)), n) != 1
    real = kg.random_below_limbs
    try:
        kg.random_below_limbs = lambda n_, count, limbs: ints_to_limbs(bad[:count], limbs)
        R = limbs_to_ints(client._draw_r(50, W))
    finally:
        kg.random_below_limbs = real
    assert all(math.gcd(r, n) == 1 and 1 <= r < n for r in R)
    assert [r for i, r in enumerate(R) if i not in (7, 31)] == [b for i, b in enumerate(bad) if i not in (7, 31)]


def test_lookup_client_bytes_path():
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    client = PaillierGPULookupClient(embed_len=E, key_len=KL, rng=random.Random(11))
    pk = json.loads(client.stringify_pk())
    n = int(pk["n"])
    embs = _embs(12, 10)
    client.dump_tables()  # the tables are built on first use, drawing from the client's generator
    client._rng = random.Random(5)
    by = client.encrypt_bytes(embs)
    client._rng = random.Random(5)
    ints = client.encrypt(np.array(embs, dtype=bool))
    assert [[int.from_bytes(c, "little") for c in r] for r in by] == [[int(c) for c in r] for r in ints]
    encs = [PaillierGPULookupClient.encode_hamming_server_bytes(by[i], by[i + 1], pk) for i in range(len(embs) - 1)]
    encs_int = [PaillierGPULookupClient.encode_hamming_server(ints[i], ints[i + 1], pk) for i in range(len(embs) - 1)]
    assert [[int.from_bytes(c, "little") for c in r] for r in encs] == [[int(c) for c in r] for r in encs_int]
    dist = [O.hamming(embs[i], embs[i + 1]) for i in range(len(embs) - 1)]
    assert client.decode_hamming_client(encs) == dist
    assert client.decode_hamming_client(encs_int) == dist
    assert client.decode_hamming_client([[str(int(c)) for c in r] for r in encs_int]) == dist  # decimal strings
    rng = random.Random(9)
    ms = [0, n - 1] + [rng.randrange(n) for _ in range(30)]
    idx = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
    cb = client.encrypt_cts_bytes([_le(m) for m in ms], noise_idx=idx)
    assert [int.from_bytes(c, "little") for c in cb] == [int(c) for c in client.encrypt_cts(ms, noise_idx=idx)]
    assert client.decrypt_cts_bytes(cb) == [_le(m) for m in ms]
    with pytest.raises(RuntimeError, match="out of range"):
        client.encrypt_cts_bytes([_le(n)])
    with pytest.raises(ValueError, match="ciphertext"):
        client.decrypt_cts_bytes([_le(n * n + 1)])
