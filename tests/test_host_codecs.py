"""Batched host conversions of the client path (no GPU): limb <-> LE bytes /
hex codecs against Python's int.to_bytes / format / int(s, 16), the row-wise
range check against integer comparison, the embedding validation fast paths
(numpy arrays, lists through bytes(row)) with the reference client's errors
(cu:885-896), the bit packing against the oracle at ragged lengths, and the
batched r draws (uniform in [1, n-1])."""

import collections
import random

import numpy as np
import pytest

from oracle import paillier_oracle as O
from cuhepy_amd.limbs import (DoesNotFit, bytes_to_limbs, hex_to_limbs, ints_to_limbs, limbs_below, limbs_to_bytes,
                              limbs_to_hex, limbs_to_ints)
from cuhepy_amd.paillier import keygen as kg
from cuhepy_amd.paillier.packing import embeddings_to_plaintexts, validate_embeddings


def _values(rng, W):
    top = 32 * W
    edges = [0, 1, 255, 256, (1 << 32) - 1, 1 << 32, (1 << top) - 1, 1 << (top - 1)]
    return [v for v in edges if v < 1 << top] + [rng.getrandbits(rng.randrange(1, top + 1)) for _ in range(300)]


def _le(v):
    return v.to_bytes(max(1, (v.bit_length() + 7) // 8), "little")


@pytest.mark.parametrize("W", [1, 32, 128])
def test_bytes_and_hex_codecs_match_python_ints(W):
    rng = random.Random(W)
    vals = _values(rng, W)
    A = ints_to_limbs(vals, W)
    assert limbs_to_bytes(A) == [_le(v) for v in vals]
    assert limbs_to_hex(A) == [format(v, "x") for v in vals]
    assert limbs_to_ints(bytes_to_limbs([_le(v) for v in vals], W)) == vals
    # fixed-width and zero-padded forms, bytearray and memoryview items
    assert limbs_to_ints(bytes_to_limbs([v.to_bytes(4 * W + 3, "little") for v in vals], W)) == vals
    assert limbs_to_ints(bytes_to_limbs([bytearray(_le(v)) for v in vals[:20]], W)) == vals[:20]
    assert limbs_to_ints(bytes_to_limbs([memoryview(_le(v)) for v in vals[:20]], W)) == vals[:20]
    hx = [format(v, "x") for v in vals]
    assert limbs_to_ints(hex_to_limbs(hx, W)) == vals
    assert limbs_to_ints(hex_to_limbs([" 0x" + h.upper() + "\n" for h in hx], W)) == vals
    assert limbs_to_ints(hex_to_limbs(["000" + h for h in hx], W)) == vals
    assert bytes_to_limbs([], W).shape == (0, W) and hex_to_limbs([], W).shape == (0, W)
    for arr in (A, bytes_to_limbs([_le(v) for v in vals], W), hex_to_limbs(hx, W)):  # callers may update rows
        assert arr.flags.writeable and arr.dtype == np.uint32 and arr.flags.c_contiguous


def test_codec_errors():
    with pytest.raises(DoesNotFit):
        bytes_to_limbs([b"\x01" * 129], 32)
    with pytest.raises(DoesNotFit):
        hex_to_limbs(["1" + "0" * 1024], 128)
    with pytest.raises(DoesNotFit):
        ints_to_limbs([-1], 4)
    with pytest.raises(ValueError, match="hex"):
        hex_to_limbs(["12g4"], 4)
    with pytest.raises(ValueError):
        bytes_to_limbs([5], 4)  # an int is not a byte string


@pytest.mark.parametrize("W", [1, 4, 64])
def test_limbs_below_matches_integer_comparison(W):
    rng = random.Random(10 + W)
    vals = _values(rng, W)
    A = ints_to_limbs(vals, W)
    bounds = [0, 1, 1 << 32, (1 << (32 * W)) - 1, 1 << (32 * W), 1 << (32 * W + 5)] + rng.sample(vals, 10)
    bounds += [v + 1 for v in rng.sample(vals, 5)]
    for b in bounds:
        assert limbs_below(A, b).tolist() == [v < b for v in vals], b
    # a planted mistake: comparing the least significant differing limb fails somewhere
    b = vals[20]
    ne = A != ints_to_limbs([b], W)[0]
    low = np.argmax(ne, axis=1)
    wrong = ne.any(axis=1) & (A[np.arange(len(A)), low] < ints_to_limbs([b], W)[0][low])
    assert W == 1 or wrong.tolist() != [v < b for v in vals]


def test_validate_embeddings_fast_paths_and_errors():
    rng = random.Random(3)
    E = 37
    rows = [[rng.getrandbits(1) for _ in range(E)] for _ in range(9)]
    want = np.array(rows, dtype=np.uint8)
    for form in (rows, [tuple(r) for r in rows], want, want.astype(np.int64), want.astype(bool),
                 [np.array(r, dtype=np.uint8) for r in rows], [np.array(r, dtype=np.int64) for r in rows],
                 [[bool(b) for b in r] for r in rows]):
        got = validate_embeddings(form, E)
        assert got.dtype == np.uint8 and got.shape == (9, E) and (got == want).all()
    assert validate_embeddings([], E).shape == (0, E)
    assert validate_embeddings(np.zeros((0, E)), E).shape == (0, E)
    for bad in ([r[:-1] for r in rows], rows[:3] + [rows[3] + [0]], want[:, :-1]):
        with pytest.raises(RuntimeError, match="length"):
            validate_embeddings(bad, E)
    for v in (2, -1, 300, 0.5):
        bad = [list(r) for r in rows]
        bad[4][7] = v
        with pytest.raises(RuntimeError, match="binary"):
            validate_embeddings(bad, E)
        with pytest.raises(RuntimeError, match="binary"):
            validate_embeddings(np.array(bad), E)


@pytest.mark.parametrize("embed_len,key_len", [(1, 16), (7, 16), (17, 16), (33, 16), (250, 128), (511, 512),
                                               (1023, 512), (1025, 512)])
def test_packing_ragged_lengths_match_oracle(embed_len, key_len):
    from cuhepy_amd.limbs import width_for_bits
    rng = random.Random(embed_len * 7 + key_len)
    bits = np.array([[rng.getrandbits(1) for _ in range(embed_len)] for _ in range(5)] +
                    [[1] * embed_len, [0] * embed_len], dtype=np.uint8)
    W = width_for_bits(4 * key_len)
    got = limbs_to_ints(embeddings_to_plaintexts(bits, key_len, W))
    assert got == [m for row in bits.tolist() for m in O.embed_to_plaintexts(row, embed_len, key_len)]


def test_random_r_batch_is_in_range_and_uniform():
    for n in (7, (1 << 64) + 13, (1 << 1023) + 1155):
        W = max(1, (n.bit_length() + 31) // 32)
        v = limbs_to_ints(kg.random_below_limbs(n, 4000, W))
        assert all(1 <= x < n for x in v)
    counts = collections.Counter(limbs_to_ints(kg.random_below_limbs(7, 12000, 1)))
    assert set(counts) == {1, 2, 3, 4, 5, 6}
    assert all(1700 < c < 2300 for c in counts.values()), counts  # 2000 expected each, 6+ sigma margin
    # n just above a power of two: half the draws are rejected, the rest stay uniform
    n = (1 << 40) + 1
    v = limbs_to_ints(kg.random_below_limbs(n, 4000, 2))
    assert max(v) < n and sum(x >= (1 << 39) for x in v) > 1800
