"""CPU model of the CRT finish (k_crt_finish in cuhepy_amd/csrc/kernels.hpp)
against the oracle's CRT decryption (oracle/paillier_oracle.py decrypt_crt).

The kernel relies on u_p = c^(p-1) mod p^2 being either 0 (p | c) or 1 mod p
(Fermat), so that L_p(u_p) = floor((u_p - 1) / p) is the exact quotient of
u_p - 1 by p, below p: the Hensel division of the low H/2 limbs of u_p - 1
(multiplication by p^-1 mod 2^(16 H)) gives it, and u_p = 0 gives -1, which
is p - 1 mod p. Then m_p = L_p h_p mod p, m_q likewise, and Garner
m = m_q + q ((m_p - m_q) q^-1 mod p). Checked on units and non-units of both
factors at 1024- and 2048-bit n; a planted mistake (the zero case taken as a
quotient of 0) must fail on a non-unit.
"""

import random

import pytest

from cuhepy_amd.limbs import width_for_bits
from oracle import paillier_oracle as O


def finish_model(c, keys, half_bits, zero_as_minus_one=True):
    p, q = keys["p"], keys["q"]
    R = 1 << half_bits
    assert p < R and q < R  # the factors fit the half width (the host checks it)
    cp = O.crt_params(p, q)

    def half(f, h):
        u = pow(c % (f * f), f - 1, f * f)
        assert u == 0 or u % f == 1
        if u == 0:
            lq = f - 1 if zero_as_minus_one else 0
        else:
            lq = ((u - 1) % R) * pow(f, -1, R) % R  # Hensel quotient of the low half
            assert lq * f == u - 1
        return lq * h % f

    mp, mq = half(p, cp["hp"]), half(q, cp["hq"])
    h = (mp - mq) * cp["q_inv_p"] % p
    return mq + q * h


@pytest.mark.parametrize("key_len", [512, 1024])
def test_crt_finish_model_matches_oracle(key_len):
    keys = O.key_gen(key_len, 900 + key_len)
    n, p, q = keys["pk"]["n"], keys["p"], keys["q"]
    n2 = n * n
    half_bits = 8 * width_for_bits(n2.bit_length())  # H/2 limbs of 32 bits, H = W/2
    assert half_bits == {512: 512, 1024: 1024}[key_len]
    rng = random.Random(key_len)
    cs = [0, 1, p, q, n, p * p, q * q, 3 * p, 5 * q, n2 - 1, n2 - p]
    cs += [rng.randrange(n2) for _ in range(10)]
    cs += [p * rng.randrange(1, q * n) % n2 for _ in range(3)]  # non-units in p
    for c in cs:
        assert finish_model(c, keys, half_bits) == O.decrypt_crt(c, keys), hex(c)[:20]


def test_crt_finish_model_zero_case_is_minus_one():
    keys = O.key_gen(512, 901)
    p, q = keys["p"], keys["q"]
    c = p * p  # u_p = 0
    assert finish_model(c, keys, 512) == O.decrypt_crt(c, keys)
    assert finish_model(c, keys, 512, zero_as_minus_one=False) != O.decrypt_crt(c, keys)


def lookup_crt_model(c, keys, swap_idempotents=False):
    """k_lookup_crt_finish's recombination: x = x_p e_p + x_q e_q mod n^2 from
    x_p = c^a mod p^2, x_q = c^a mod q^2, then L and g_a_inv as lookup.py:257-278."""
    p, q, n = keys["p"], keys["q"], keys["pk"]["n"]
    n2, a = n * n, keys["sk"]["a"]
    p2, q2 = p * p, q * q
    ep = q2 * pow(q2, -1, p2) % n2
    eq = p2 * pow(p2, -1, q2) % n2
    if swap_idempotents:
        ep, eq = eq, ep
    xp, xq = pow(c % p2, a, p2), pow(c % q2, a, q2)
    x = (xp * ep + xq * eq) % n2
    return O.L(x, n) * keys["sk"]["g_a_inv"] % n


@pytest.mark.parametrize("key_len", [512, 1024])
def test_lookup_crt_model_matches_oracle(key_len):
    """The half-width lookup decryption equals the oracle's full-width formula
    for valid ciphertexts and for non-units (multiples of p^2, q^2, p, 0); a
    planted mistake (idempotents swapped) fails."""
    keys = O.lookup_key_gen(key_len, 280, 900 + key_len, with_tables=False)
    n, p, q = keys["pk"]["n"], keys["p"], keys["q"]
    n2, g, g_n = n * n, keys["pk"]["g"], keys["pk"]["g_n"]
    assert max(p, q) < 1 << (16 * width_for_bits(n2.bit_length()))  # p^2, q^2 fit the half width
    rng = random.Random(key_len)
    ms = [rng.randrange(n) for _ in range(4)]
    cs = [pow(g, m, n2) * pow(g_n, O.lookup_random_r(n, rng), n2) % n2 for m in ms]
    cs += [0, 1, p * p, 3 * q * q, 7 * p, n2 - 1, rng.randrange(n2)]
    want = [O.lookup_decrypt(c, keys) for c in cs]
    assert want[:4] == ms
    assert [lookup_crt_model(c, keys) for c in cs] == want
    assert [lookup_crt_model(c, keys, swap_idempotents=True) for c in cs] != want


@pytest.mark.parametrize("key_len", [512, 1024])
def test_paillier_exact_crt_model_matches_oracle(key_len):
    """k_encrypt_crt_finish / k_lookup_crt_finish for plain keys: r^n and
    c^phi recombined from their residues mod p^2 and q^2 equal the full-width
    powers for units and non-units, so encryption equals g^m r^n mod n^2 and
    decryption equals scheme.py:86-104's formula (the oracle) for every c; a
    planted mistake (x_q taken as x_p) fails."""
    from cuhepy_amd.engine import crt_idempotents
    keys = O.key_gen(key_len, 700 + key_len)
    n, p, q = keys["pk"]["n"], keys["p"], keys["q"]
    n2, phi, inv = n * n, keys["sk"]["phi"], keys["sk"]["inv"]
    p2, q2 = p * p, q * q
    ep, eq = crt_idempotents(p, q)

    def power(x, e, planted=False):
        xp, xq = pow(x % p2, e, p2), pow(x % q2, e, q2)
        return (xp * ep + (xp if planted else xq) * eq) % n2

    rng = random.Random(key_len)
    xs = [0, 1, p, 3 * q, p2, q2, n, n2 - 1] + [rng.randrange(n2) for _ in range(6)]
    for x in xs:
        assert power(x, n) == pow(x, n, n2)
        assert (power(x, phi) - 1) // n * inv % n == O.decrypt(x, keys)
    assert any(power(x, phi, planted=True) != pow(x, phi, n2) for x in xs)
    m, r = rng.randrange(n), O.random_r(n, rng)
    assert (1 + m * n) * power(r, n) % n2 == O.encrypt(m, keys["pk"], r)


def lookup_fast_model(c, keys, quarter_bits, drop_gai=False):
    """k_lookup_crt_fast: x_p = c^a mod p^2 passes when x_p R = lo R + hi R^2
    = R (mod p) for x_p = lo + hi R (R = 2^quarter_bits), likewise for q; then
    L_p = (x_p - 1)/p (the Hensel quotient of the low quarter), m_p = L_p h_p
    mod p with h_p = q^-1 g_a_inv, m_q likewise, and Garner. Returns (passed,
    m)."""
    n, p, q, a = keys["pk"]["n"], keys["p"], keys["q"], keys["sk"]["a"]
    gai = 1 if drop_gai else keys["sk"]["g_a_inv"]
    R = 1 << quarter_bits
    assert p < R and q < R

    def half(f, g):
        x = pow(c, a, f * f)
        lo, hi = x % R, x // R
        ok = (lo * R + hi * R * R) % f == R % f
        lq = ((x - 1) % R) * pow(f, -1, R) % R
        return ok, lq * (pow(g, -1, f) * gai % f) % f

    okp, mp = half(p, q)
    okq, mq = half(q, p)
    h = (mp - mq) * pow(q, -1, p) % p
    return okp and okq, mq + q * h


@pytest.mark.parametrize("key_len", [512, 1024])
def test_lookup_fast_model(key_len):
    """The textbook CRT route of lookup decryption passes every valid
    ciphertext with the oracle's plaintext (lookup.py:257-278), and fails its
    test exactly where x = c^a mod n^2 is not 1 mod n (random c in [0, n^2),
    non-units, 0), which the exact recombination then recomputes; a planted
    mistake (h_p without g_a_inv) gives wrong plaintexts."""
    keys = O.lookup_key_gen(key_len, 280, 900 + key_len, with_tables=False)
    n, p, q, a = keys["pk"]["n"], keys["p"], keys["q"], keys["sk"]["a"]
    n2, g, g_n = n * n, keys["pk"]["g"], keys["pk"]["g_n"]
    qb = 8 * width_for_bits(n2.bit_length())  # quarter width in bits: p, q at H/2 limbs
    rng = random.Random(key_len + 7)
    ms = [0, 1, n - 1] + [rng.randrange(n) for _ in range(6)]
    valid = [pow(g, m, n2) * pow(g_n, O.lookup_random_r(n, rng), n2) % n2 for m in ms]
    for m, c in zip(ms, valid):
        ok, got = lookup_fast_model(c, keys, qb)
        assert ok and got == m == O.lookup_decrypt(c, keys)
    odd = [0, 1, p * p, 3 * q * q, 7 * p, n2 - 1] + [rng.randrange(n2) for _ in range(6)]
    for c in odd:
        ok, got = lookup_fast_model(c, keys, qb)
        assert ok == (pow(c, a, n2) % n == 1)
        if ok:
            assert got == O.lookup_decrypt(c, keys)
    assert any(not lookup_fast_model(c, keys, qb)[0] for c in odd)
    assert [lookup_fast_model(c, keys, qb, drop_gai=True)[1] for c in valid[3:]] != ms[3:]
