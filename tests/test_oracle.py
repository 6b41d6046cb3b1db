"""Pins the exact-integer oracle (CPU only).

Every oracle function is checked against something other than its own
formula: a closed form or identity, an exhaustive tiny-modulus brute force, the
reference's plaintext known answers, an independent engine (GMP mpz_powm via
ctypes vs CPython pow), or a planted mistake that must be caught.
"""

import math
import random

import pytest

from oracle import gmp_engine as G
from oracle import paillier_oracle as O
from conftest import H


# --------------------------------------------------------------------------
# reference known answers (plaintext level)
# --------------------------------------------------------------------------

def test_reference_known_answers(golden):
    ka = golden["hamming"]
    # tests/unit/test_GPU_paillier_clients.py:12-16 and README.md:59-76
    assert O.hamming(ka["vectors_8"][0], ka["vectors_8"][1]) == 4
    assert O.hamming(ka["vectors_8"][1], ka["vectors_8"][2]) == 4
    assert O.hamming(ka["readme_a"], ka["readme_b"]) == 256


@pytest.mark.parametrize("embed_len,key_len", [(8, 1024), (512, 1024), (64, 512), (600, 256)])
def test_bit_packing_decode_is_hamming(embed_len, key_len):
    """decode(encode(a) + encode(b)) = Hamming(a, b) -- the protocol identity the
    reference's GPU tests assert (test_GPU_paillier_clients.py:59-95)."""
    rng = random.Random(embed_len * 7 + key_len)
    chunk_len, chunk_num = O.chunk_geometry(embed_len, key_len)
    for _ in range(20):
        a = [rng.getrandbits(1) for _ in range(embed_len)]
        b = [rng.getrandbits(1) for _ in range(embed_len)]
        ma = O.embed_to_plaintexts(a, embed_len, key_len)
        mb = O.embed_to_plaintexts(b, embed_len, key_len)
        assert len(ma) == chunk_num
        sums = [x + y for x, y in zip(ma, mb)]
        assert O.decode_hamming(sums, chunk_len) == O.hamming(a, b)
        # the mask form used on the GPU counts the same bits
        mask = O.odd_mask(chunk_len)
        assert sum(bin(s & mask).count("1") for s in sums) == O.hamming(a, b)


def test_chunk_geometry_matches_reference_formula():
    # hamming/paillier.py:26-32 and cu:739-746
    assert O.chunk_geometry(512, 1024) == (2048, 1)
    assert O.chunk_geometry(8, 1024) == (2048, 1)
    assert O.chunk_geometry(1500, 1024) == (2048, 2)
    assert O.chunk_geometry(2048, 1024) == (2048, 2)
    assert O.chunk_geometry(2049, 1024) == (2048, 3)


# --------------------------------------------------------------------------
# Paillier: closed forms, identities, brute force
# --------------------------------------------------------------------------

def test_binomial_identity_and_encrypt_closed_form():
    keys = O.key_gen(256, 5)
    n, n2 = keys["pk"]["n"], keys["pk"]["n_squared"]
    rng = random.Random(9)
    for _ in range(10):
        m, r = rng.randrange(n), O.random_r(n, rng)
        # (1+n)^m = 1 + m n (mod n^2): the form the reference GPU kernel uses (cu:522-530)
        assert pow(n + 1, m, n2) == (1 + m * n) % n2
        c = O.encrypt(m, keys["pk"], r)
        assert c == (1 + m * n) * pow(r, n, n2) % n2
        assert O.decrypt(c, keys) == m
        assert O.decrypt_crt(c, keys) == m


def test_homomorphic_addition():
    keys = O.key_gen(256, 6)
    n = keys["pk"]["n"]
    rng = random.Random(10)
    for _ in range(10):
        m1, m2 = rng.randrange(n), rng.randrange(n)
        c1 = O.encrypt(m1, keys["pk"], O.random_r(n, rng))
        c2 = O.encrypt(m2, keys["pk"], O.random_r(n, rng))
        assert O.decrypt(O.add(c1, c2, keys["pk"]), keys) == (m1 + m2) % n


def test_exhaustive_tiny_modulus():
    """n = 11*13: every m in [0, n) and every r coprime to n round-trips;
    r^(n*lambda) = 1 and the ciphertexts of one m are exactly the n-th residues
    times (1 + m n)."""
    p, q = 11, 13
    n, n2 = p * q, (p * q) ** 2
    phi = (p - 1) * (q - 1)
    keys = {"pk": {"g": n + 1, "n": n, "n_squared": n2}, "sk": {"phi": phi, "inv": pow(phi, -1, n)},
            "p": p, "q": q}
    lam = phi // math.gcd(p - 1, q - 1)
    seen = set()
    for m in range(n):
        for r in range(1, n):
            if math.gcd(r, n) != 1:
                continue
            c = O.encrypt(m, keys["pk"], r)
            assert pow(r, n * lam, n2) == 1
            assert O.decrypt(c, keys) == m
            assert O.decrypt_crt(c, keys) == m
            seen.add(c)
    # encryption is injective on (m, r) classes: n * phi distinct ciphertexts = |Z*_{n^2}|
    assert len(seen) == n * phi


def test_decrypt_floor_semantics_on_non_units():
    """c not coprime to n: the reference computes Python floor division; the
    oracle must too (c = 0 gives (-1)//n = -1)."""
    keys = O.key_gen(128, 8)
    n, inv = keys["pk"]["n"], keys["sk"]["inv"]
    assert O.decrypt(0, keys) == (-1 * inv) % n
    x = pow(n, keys["sk"]["phi"], n * n) - 1
    assert O.decrypt(n, keys) == ((x // n) * inv) % n


def test_factor_from_phi():
    keys = O.key_gen(256, 11)
    assert O.factor_from_phi(keys["pk"]["n"], keys["sk"]["phi"]) == tuple(sorted((keys["p"], keys["q"])))


@pytest.mark.parametrize("bits", [64, 257, 512])
def test_prime_generation_is_prime_and_sized(bits):
    rng = random.Random(bits)
    p = O.get_prime(bits, rng)
    assert p.bit_length() == bits
    # independent check: Fermat with several bases + trial division
    assert all(pow(a, p - 1, p) == 1 for a in (2, 3, 5, 7, 11, 13))
    assert all(p % d for d in range(3, 1000, 2))


# --------------------------------------------------------------------------
# GMP cross-check (independent exact engine)
# --------------------------------------------------------------------------

@pytest.mark.skipif(not G.available(), reason="libgmp not loadable")
def test_gmp_powm_agrees_with_cpython_pow(golden):
    for case in ("paillier_w64", "paillier_w128"):
        g = golden[case]
        n2 = H(g["n_squared"])
        for e_hex, row in zip(g["modexp_exps"], g["modexp"]):
            e = H(e_hex)
            for b_hex, want in zip(g["modexp_bases"][:4], row[:4]):
                assert G.powm(H(b_hex), e, n2) == H(want) == pow(H(b_hex), e, n2)


@pytest.mark.skipif(not G.available(), reason="libgmp not loadable")
def test_gmp_crt_decryptor_matches_oracle(golden):
    g = golden["paillier_w64"]
    p, q = H(g["p"]), H(g["q"])
    dec = G.CrtDecryptor(**O.crt_params(p, q))
    for c, m in zip(g["c"], g["m"]):
        assert dec(H(c)) == H(m)


# --------------------------------------------------------------------------
# golden vectors are reproduced by the oracle
# --------------------------------------------------------------------------

@pytest.mark.parametrize("case", ["paillier_w64", "paillier_w128"])
def test_golden_paillier_reproduces(golden, case):
    g = golden[case]
    keys = O.key_gen(g["key_len"], g["seed"])
    assert keys["pk"]["n"] == H(g["n"]) and keys["p"] == H(g["p"])
    for m, r, c in zip(g["m"], g["r"], g["c"]):
        assert O.encrypt(H(m), keys["pk"], H(r)) == H(c)
        assert O.decrypt(H(c), keys) == H(m)
        assert O.decrypt_crt(H(c), keys) == H(m)
    for c, d, dc in zip(g["edge_c"], g["edge_dec"], g["edge_crt"]):
        assert O.decrypt(H(c), keys) == H(d)
        assert O.decrypt_crt(H(c), keys) == H(dc)


def test_golden_lookup_reproduces(golden):
    g = golden["lookup_w32"]
    keys = O.lookup_key_gen(g["key_len"], g["alpha_len"], g["seed"], with_tables=False)
    n, n2, gg = keys["pk"]["n"], keys["pk"]["n_squared"], keys["pk"]["g"]
    assert n == H(g["n"]) and gg == H(g["g"])
    a = keys["sk"]["a"]
    # lookup.py:78-81 sanity conditions
    assert pow(gg, n * a, n2) == 1 and pow(gg, n, n2) != 1
    assert a.bit_length() >= min(g["key_len"], g["alpha_len"]) - 2  # q1*q2 bits (lookup.py:52-54)
    table = O.precompute_g_table(gg, n, g["key_len"])
    for i, j, v in g["g_table_spots"]:
        assert table[i][j] == H(v) == pow(gg, j * 2 ** (8 * i), n2)
    noise = O.precompute_noise_table(gg, n, [H(x) for x in g["noise_r"]])
    for m, ix, c in zip(g["m"], g["noise_idx"], g["c"]):
        got = O.lookup_encrypt(H(m), keys["pk"], table, noise, g["message_chunks"], ix)
        assert got == H(c)
        # closed form: g^m * prod g_n^r_k
        assert got == pow(gg, H(m), n2) * math.prod(noise[k] for k in ix) % n2
        assert O.lookup_decrypt(got, keys) == H(m)


# --------------------------------------------------------------------------
# planted mistakes: each must be caught by one of the checks above
# --------------------------------------------------------------------------

def test_planted_mistakes_are_caught(golden, monkeypatch):
    g = golden["paillier_w64"]
    keys = O.key_gen(g["key_len"], g["seed"])
    m, r, c = H(g["m"][5]), H(g["r"][5]), H(g["c"][5])

    # dropped r^n term
    def enc_no_r(m_, pk, r_):
        return pow(pk["g"], m_, pk["n_squared"])
    assert enc_no_r(m, keys["pk"], r) != c
    # wrong sign in the CRT recombination
    cp = O.crt_params(keys["p"], keys["q"])
    p, q = cp["p"], cp["q"]
    mp = ((pow(c % (p * p), p - 1, p * p) - 1) // p) * cp["hp"] % p
    mq = ((pow(c % (q * q), q - 1, q * q) - 1) // q) * cp["hq"] % q
    bad = mq + q * ((mq - mp) * cp["q_inv_p"] % p)
    assert bad != H(g["m"][5])
    # transposed digit order in lookup encryption
    digits = O.lookup_digits(12345678, 8)
    assert sum(d << (8 * i) for i, d in enumerate(digits)) == 12345678
    assert sum(d << (8 * i) for i, d in enumerate(reversed(digits))) != 12345678
    # off-by-one odd mask (even positions) miscounts Hamming distance
    a = [1, 0] * 8
    b = [0, 0] * 8
    ma = O.embed_to_plaintexts(a, 16, 1024)[0]
    mb = O.embed_to_plaintexts(b, 16, 1024)[0]
    even_mask = O.odd_mask(2048) << 1
    assert bin((ma + mb) & even_mask).count("1") != O.hamming(a, b)


# --------------------------------------------------------------------------
# bit packing and decode pinned to their closed forms
# --------------------------------------------------------------------------
# hamming/paillier.py:86-94 writes the pair "0 b_j" for every bit, cuts the
# string into chunks of chunk_len = 2 key_len characters and parses each chunk
# as a binary numeral, most significant character first. Chunk c therefore
# holds the bits j in [c key_len, min((c+1) key_len, embed_len)); with E_c of
# them, bit u of the chunk (u = j - c key_len) sits at string index 2u + 1 of
# a numeral of 2 E_c digits, i.e. at weight 2^(2 E_c - 2 - 2u) = 4^(E_c-1-u).
# The last chunk may be shorter: its numeral is just its own 2 E_c digits.

def packing_closed_form(bits, key_len):
    out = []
    for c0 in range(0, len(bits), key_len):
        chunk = bits[c0:c0 + key_len]
        e = len(chunk)
        out.append(sum(b * 4 ** (e - 1 - u) for u, b in enumerate(chunk)))
    return out


def packing_reversed_pairs(bits, key_len):
    """A plausible mistake: the pairs of a chunk emitted in reversed bit order."""
    return [sum(b * 4 ** u for u, b in enumerate(bits[c0:c0 + key_len])) for c0 in range(0, len(bits), key_len)]


def test_bit_packing_known_answers():
    # literal numerals of the "0 b" string: [0,1,0,1,1,0,1,0] -> "0001000101000100"
    assert O.embed_to_plaintexts([0, 1, 0, 1, 1, 0, 1, 0], 8, 1024) == [0b0001000101000100]
    assert O.embed_to_plaintexts([1, 1, 0, 0, 1, 0, 0, 1], 8, 1024) == [0b0101000001000001]
    assert O.embed_to_plaintexts([1] + [0] * 7, 8, 1024) == [4 ** 7]
    assert O.embed_to_plaintexts([0] * 7 + [1], 8, 1024) == [1]
    # two chunks at key_len 4 (chunk_len 8): bits 0..3 and 4..5 (short last chunk)
    assert O.embed_to_plaintexts([1, 0, 0, 1, 1, 1], 6, 4) == [0b01000001, 0b0101]


@pytest.mark.parametrize("embed_len,key_len", [(8, 1024), (512, 1024), (1024, 1024), (1500, 1024),
                                               (2048, 1024), (2049, 1024), (600, 256), (64, 512)])
def test_bit_packing_matches_closed_form(embed_len, key_len):
    rng = random.Random(embed_len * 31 + key_len)
    chunk_len, chunk_num = O.chunk_geometry(embed_len, key_len)
    for t in range(8):
        bits = [rng.getrandbits(1) for _ in range(embed_len)]
        if t == 0:
            bits = [1] * embed_len  # every pair set: the widest numeral of each chunk
        want = packing_closed_form(bits, key_len)
        assert len(want) == chunk_num
        assert O.embed_to_plaintexts(bits, embed_len, key_len) == want
        # every chunk's numeral fits a plaintext of n (2 key_len bits, top two set)
        assert all(m < 2 ** (chunk_len - 1) for m in want)
        # decode of a single encoding is the weight of the embedding
        assert O.decode_hamming(want, chunk_len) == sum(bits)


def test_decode_matches_closed_form():
    """decode = popcount at even LSB indices of each plaintext (the odd string
    indices after left-padding to chunk_len), for any plaintext below 2^chunk_len."""
    rng = random.Random(77)
    for chunk_len in (8, 512, 2048):
        even = (4 ** (chunk_len // 2) - 1) // 3
        for _ in range(20):
            vals = [rng.getrandbits(rng.randrange(1, chunk_len + 1)) for _ in range(rng.randrange(1, 4))]
            want = sum(bin(v & even).count("1") for v in vals)
            assert O.decode_hamming(vals, chunk_len) == want
        assert O.decode_hamming([0], chunk_len) == 0
        assert O.decode_hamming([1], chunk_len) == 1  # "1" padded: index chunk_len - 1 is odd
        assert O.decode_hamming([2], chunk_len) == 0
        assert O.odd_mask(chunk_len) == even


def test_planted_reversed_pairs_are_caught():
    """Reversed pair order keeps every order-free identity but fails the closed form."""
    rng = random.Random(5)
    a = [rng.getrandbits(1) for _ in range(512)]
    b = [rng.getrandbits(1) for _ in range(512)]
    ra, rb = packing_reversed_pairs(a, 1024), packing_reversed_pairs(b, 1024)
    # the identity decode(enc(a) + enc(b)) = Hamming(a, b) still holds ...
    assert O.decode_hamming([x + y for x, y in zip(ra, rb)], 2048) == O.hamming(a, b)
    # ... but the closed form (and so the oracle) disagrees
    assert ra != packing_closed_form(a, 1024)
    assert ra != O.embed_to_plaintexts(a, 512, 1024)


# --------------------------------------------------------------------------
# randomness and noise: what the scheme requires of them
# --------------------------------------------------------------------------

def test_noise_entries_encrypt_zero(golden):
    """lookup.py:113-117: every noise entry is an encryption of 0 (an n-th power
    residue g^(n r)), so it is killed by a and decrypts to 0; the mistake of
    dropping the exponent n (g^r) is caught."""
    g = golden["lookup_w32"]
    keys = O.lookup_key_gen(g["key_len"], g["alpha_len"], g["seed"], with_tables=False)
    n, n2, gg, a = keys["pk"]["n"], keys["pk"]["n_squared"], keys["pk"]["g"], keys["sk"]["a"]
    rs = [H(x) for x in g["noise_r"]]
    noise = O.precompute_noise_table(gg, n, rs)
    assert len(noise) == len(rs)
    for v in noise:
        assert pow(v, a, n2) == 1 and O.lookup_decrypt(v, keys) == 0 and v != 1
    # a message times noise decrypts to the message
    for m in (0, 1, n - 1, 12345):
        c = pow(gg, m, n2) * noise[m % len(noise)] % n2
        assert O.lookup_decrypt(c, keys) == m
    wrong = [pow(gg, r, n2) for r in rs[:8]]  # exponent n dropped
    assert any(O.lookup_decrypt(v, keys) != 0 for v in wrong)


def test_random_r_draws_units_in_range():
    """scheme.py:47-60 and lookup.py:198-210: r is a unit of Z_n in the stated
    range, drawn afresh (no fixed fallback); checked on a modulus with small
    factors, where non-units are common and must be rejected."""
    n = 3 * 5 * 7 * 11 * 13 * 17  # 255,255: about 62% of [0, n) are non-units
    rng = random.Random(12)
    plain = [O.random_r(n, rng) for _ in range(400)]
    look = [O.lookup_random_r(n, rng) for _ in range(400)]
    assert all(0 < r < n and math.gcd(r, n) == 1 for r in plain)
    assert all(1 <= r <= n and math.gcd(r, n) == 1 for r in look)
    assert len(set(plain)) > 350 and len(set(look)) > 350
