"""Writes the golden vectors under tests/golden/ from the exact-integer oracle.

Run from the repository root:  python tests/golden/make_golden.py

The reference ships no ciphertext golden vectors (its tests are round trips;
ciphertexts there depend on OS randomness). These fixtures pin the oracle and
the GPU path to fixed seeded keys, messages, r values and noise indices, and
carry the reference's own plaintext-level known answers (the three _VECTORS of
tests/unit/test_GPU_paillier_clients.py:12-16 and the README pair,
README.md:59-76).
"""

from __future__ import annotations

import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import paillier_oracle as O  # noqa: E402


def hx(v: int) -> str:
    return format(v, "x")


def paillier_case(key_len: int, seed: int, count: int) -> dict:
    keys = O.key_gen(key_len, seed)
    n, n2 = keys["pk"]["n"], keys["pk"]["n_squared"]
    rng = random.Random(seed + 1)
    ms = [0, 1, n - 1, n // 2] + [rng.randrange(n) for _ in range(count - 4)]
    rs = [1, n - 1, 2, O.random_r(n, rng)] + [O.random_r(n, rng) for _ in range(count - 4)]
    cs = [O.encrypt(m, keys["pk"], r) for m, r in zip(ms, rs)]
    dec = [O.decrypt(c, keys) for c in cs]
    assert dec == ms
    crt = [O.decrypt_crt(c, keys) for c in cs]
    assert crt == ms
    # edge ciphertexts: c = 1 (m = 0), c = 0 and c = n (not coprime; floor-division semantics)
    edge_c = [1, 0, n, n2 - 1]
    edge_dec = [O.decrypt(c, keys) for c in edge_c]
    edge_crt = [O.decrypt_crt(c, keys) for c in edge_c]
    # mulmod pairs (homomorphic add) incl. operands >= n^2 as the evaluator accepts
    pairs = [(cs[i], cs[(i + 1) % count]) for i in range(count)]
    mul = [O.add(a, b, keys["pk"]) for a, b in pairs]
    # generic modexp mod n^2 with full-width exponents
    bases = [rng.randrange(n2) for _ in range(8)] + [0, 1, n2 - 1]
    exps = [rng.getrandbits(n2.bit_length()) | (1 << (n2.bit_length() - 1)) for _ in range(3)] + [1, 2, 65537]
    modexp = [[hx(pow(b, e, n2)) for b in bases] for e in exps]
    return {
        "key_len": key_len,
        "seed": seed,
        "p": hx(keys["p"]), "q": hx(keys["q"]), "n": hx(n), "n_squared": hx(n2),
        "phi": hx(keys["sk"]["phi"]), "inv": hx(keys["sk"]["inv"]),
        "m": [hx(v) for v in ms], "r": [hx(v) for v in rs], "c": [hx(v) for v in cs],
        "edge_c": [hx(v) for v in edge_c], "edge_dec": [hx(v) for v in edge_dec],
        "edge_crt": [hx(v) for v in edge_crt],
        "mul_a": [hx(a) for a, _ in pairs], "mul_b": [hx(b) for _, b in pairs], "mul": [hx(v) for v in mul],
        "modexp_bases": [hx(b) for b in bases], "modexp_exps": [hx(e) for e in exps], "modexp": modexp,
    }


def lookup_case(key_len: int, alpha_len: int, seed: int, count: int) -> dict:
    keys = O.lookup_key_gen(key_len, alpha_len, seed, with_tables=False)
    n, n2, g = keys["pk"]["n"], keys["pk"]["n_squared"], keys["pk"]["g"]
    rng = random.Random(seed + 7)
    noise_rs = [O.lookup_random_r(n, rng) for _ in range(O.NOISE_TABLE_SIZE)]
    table = O.precompute_g_table(g, n, key_len)
    noise = O.precompute_noise_table(g, n, noise_rs)
    mc = keys["message_chunks"]
    ms = [0, 1, n - 1] + [rng.randrange(n) for _ in range(count - 3)]
    idx = [[rng.randrange(256) for _ in range(O.NOISE_MULTIPLES)] for _ in ms]
    cs = [O.lookup_encrypt(m, keys["pk"], table, noise, mc, ix) for m, ix in zip(ms, idx)]
    dec = [O.lookup_decrypt(c, keys) for c in cs]
    assert dec == ms
    return {
        "key_len": key_len, "alpha_len": alpha_len, "seed": seed,
        "n": hx(n), "g": hx(g), "a": hx(keys["sk"]["a"]), "g_a_inv": hx(keys["sk"]["g_a_inv"]),
        "phi": hx(keys["sk"]["phi"]), "g_n": hx(keys["pk"]["g_n"]), "message_chunks": mc,
        "noise_r": [hx(v) for v in noise_rs],
        "m": [hx(v) for v in ms], "noise_idx": idx, "c": [hx(v) for v in cs],
        # spot entries of the g table: (i, j, value)
        "g_table_spots": [[i, j, hx(table[i][j])] for i, j in ((0, 0), (0, 1), (1, 255), (mc - 1, 128))],
    }


def hamming_known_answers() -> dict:
    vectors = [[0, 1, 0, 1, 1, 0, 1, 0], [1, 1, 0, 0, 1, 0, 0, 1], [1, 0, 1, 0, 1, 1, 0, 0]]
    a = [i % 2 for i in range(512)]
    b = [(i // 2) % 2 for i in range(512)]
    return {
        "vectors_8": vectors,
        "dist_01": O.hamming(vectors[0], vectors[1]),
        "dist_12": O.hamming(vectors[1], vectors[2]),
        "readme_a": a, "readme_b": b, "readme_dist": O.hamming(a, b),
    }


def main() -> None:
    data = {
        "paillier_w64": paillier_case(512, 1001, 24),
        "paillier_w128": paillier_case(1024, 1002, 12),
        "lookup_w32": lookup_case(256, 280, 2001, 12),
        "hamming": hamming_known_answers(),
    }
    ka = data["hamming"]
    assert ka["dist_01"] == 4 and ka["dist_12"] == 4 and ka["readme_dist"] == 256
    with open(os.path.join(HERE, "golden_paillier.json"), "w") as f:
        json.dump(data, f, indent=0)


if __name__ == "__main__":
    main()
