"""Host-side key generation for the Paillier and Paillier-lookup schemes.

Key generation is not on the batched hot path: the reference runs it on the
host with GMP (``PaillierCPU::key_gen``, paillier_GPU_client.cu:183-216;
lookup keygen paillier_GPU_lookup_client.cu:248-429) or with gmpy2 +
pycryptodome (``paillier/scheme.py:24-44``, ``paillier/lookup.py:44-100``).
Here it is plain Python integers. Randomness comes from ``secrets`` unless a
``random.Random`` is passed (seeded, reproducible keys for tests and benches).
"""

from __future__ import annotations

import math
import os
import random
import secrets
from typing import Optional

import numpy as np

from ..limbs import limbs_below

MSG_BITS = 8                 # lookup.py:11
NOISE_TABLE_SIZE = 2 ** 8    # lookup.py:12
NOISE_MULTIPLES = 14         # lookup.py:13
MIN_ALPHA_LEN = 256          # lookup.py:26
DEFAULT_ALPHA_LEN = 280      # lookup.py:27

_SMALL = [p for p in range(3, 3000) if all(p % d for d in range(2, int(p ** 0.5) + 1))]


def _rng(rng: Optional[random.Random]) -> random.Random:
    return rng if rng is not None else secrets.SystemRandom()


def is_probable_prime(n: int, rounds: int = 40, rng: Optional[random.Random] = None) -> bool:
    """Miller-Rabin (gmpy2.is_prime(p, 40) in lookup.py:181)."""
    if n < 2:
        return False
    for p in [2] + _SMALL:
        if n == p:
            return True
        if n % p == 0:
            return False
    r = _rng(rng)
    d, s = n - 1, 0
    while not d & 1:
        d >>= 1
        s += 1
    for _ in range(rounds):
        a = r.randrange(2, n - 1)
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def random_prime(bits: int, rng: Optional[random.Random] = None, top_two: bool = True) -> int:
    """A `bits`-bit probable prime (number.getPrime in scheme.py:33-35)."""
    r = _rng(rng)
    while True:
        c = r.getrandbits(bits) | (1 << (bits - 1)) | 1
        if top_two and bits >= 2:
            c |= 1 << (bits - 2)
        if is_probable_prime(c, 40, r):
            return c


def paillier_key_gen(key_len: int, rng: Optional[random.Random] = None) -> dict:
    """scheme.py:24-44: two key_len-bit primes, g = n + 1, sk = {phi, phi^-1 mod n}.

    The factors p, q are kept (needed for CRT decryption; the reference key dict
    holds only phi, from which they are recoverable).
    """
    r = _rng(rng)
    p = random_prime(key_len, r)
    q = random_prime(key_len, r)
    while q == p:
        q = random_prime(key_len, r)
    n = p * q
    phi = (p - 1) * (q - 1)
    return {"sk": {"phi": phi, "inv": pow(phi, -1, n)}, "pk": {"g": n + 1, "n": n, "n_squared": n * n},
            "p": p, "q": q}


def factor_from_phi(n: int, phi: int) -> Optional[tuple[int, int]]:
    """(p, q) from (n, phi) with p + q = n - phi + 1; None if phi is not (p-1)(q-1)."""
    s = n - phi + 1
    disc = s * s - 4 * n
    if disc < 0:
        return None
    d = math.isqrt(disc)
    if d * d != disc:
        return None
    p, q = (s - d) // 2, (s + d) // 2
    if p * q != n or p <= 1:
        return None
    return p, q


def random_r(n: int, rng: Optional[random.Random] = None) -> int:
    """scheme.py:47-60: r uniform in [1, n-1] with gcd(r, n) = 1."""
    r = _rng(rng)
    while True:
        v = r.randrange(n - 1) + 1
        if math.gcd(v, n) == 1:
            return v


def random_below_limbs(n: int, count: int, limbs: int) -> np.ndarray:
    """`count` values uniform in [1, n-1] as (count, limbs) uint32, from the
    operating system's random bytes: draws of bit_length(n - 1) bits,
    rejecting 0 and values >= n (at most half of them). Coprimality with n
    is the caller's check (``Modulus.product`` over the batch)."""
    kb = (n - 1).bit_length()
    words = (kb + 31) // 32
    if words > limbs:
        raise ValueError("n does not fit the width")
    top = np.uint32((1 << (kb - 32 * (words - 1))) - 1)
    out = np.zeros((count, limbs), dtype=np.uint32)
    have = 0
    while have < count:
        draw = (count - have) * 2 + 8
        x = np.zeros((draw, limbs), dtype=np.uint32)
        x[:, :words] = np.frombuffer(os.urandom(draw * words * 4), dtype="<u4").reshape(draw, words)
        x[:, words - 1] &= top
        x = x[limbs_below(x, n) & x.any(axis=1)][: count - have]
        out[have:have + len(x)] = x
        have += len(x)
    return out


# ---------------------------------------------------------------------------
# lookup variant
# ---------------------------------------------------------------------------

def _dsa_params(p_bits: int, q_bits: int, r: random.Random) -> tuple[int, int, int]:
    """lookup.py:150-195."""
    if q_bits >= p_bits:
        raise ValueError("alpha_len too large: q_bits must be < p_bits")
    q = random_prime(q_bits, r, top_two=False)
    k_lo = (1 << (p_bits - 1)) // q
    k_hi = ((1 << p_bits) - 1) // q
    if k_lo % 2 == 1:
        k_lo += 1
    while True:
        span = (k_hi - k_lo) // 2 + 1
        k = k_lo + 2 * r.randrange(0, span)
        p = k * q + 1
        if p.bit_length() != p_bits:
            continue
        if is_probable_prime(p, 40, r):
            break
    e = (p - 1) // q
    while True:
        h = r.randrange(2, p - 1)
        g = pow(h, e, p)
        if g != 1:
            return p, q, g


def lookup_key_gen(key_len: int, alpha_len: int, rng: Optional[random.Random] = None) -> dict:
    """lookup.py:44-100 (tables are built separately, see ``build_tables``)."""
    r = _rng(rng)
    while True:
        bound = min(key_len, alpha_len)
        q1_bits = bound // 2
        q2_bits = bound - q1_bits
        p, q1, g1 = _dsa_params(key_len, q1_bits, r)
        q, q2, g2 = _dsa_params(key_len, q2_bits, r)
        if p == q:
            continue
        n = p * q
        n_sq = n * n
        g1_p2 = (g1 % p) * (1 + p) % (p * p)
        g2_q2 = (g2 % q) * (1 + q) % (q * q)
        m1, m2 = p * p, q * q
        t = ((g2_q2 - g1_p2) % m2) * pow(m1 % m2, -1, m2) % m2
        g = (g1_p2 + m1 * t) % (m1 * m2)
        a = q1 * q2 // math.gcd(q1, q2)
        if pow(g, n * a, n_sq) != 1 or pow(g, n, n_sq) == 1:
            continue
        break
    phi = (p - 1) * (q - 1)
    g_n = pow(g, n, n_sq)
    g_a_inv = pow((pow(g, a, n_sq) - 1) // n, -1, n)
    return {
        "sk": {"phi": phi, "a": a, "g_a_inv": g_a_inv},
        "pk": {"g": g, "n": n, "n_squared": n_sq, "g_n": g_n},
        "key_len": key_len,
        "message_chunks": (key_len * 2 + MSG_BITS - 1) // MSG_BITS,
    }


def lookup_random_r(n: int, rng: Optional[random.Random] = None) -> int:
    """lookup.py:198-210: r in [1, n] with gcd(r, n) = 1."""
    r = _rng(rng)
    while True:
        v = r.randrange(n) + 1
        if math.gcd(v, n) == 1:
            return v


def build_g_table(g: int, n: int, key_len: int) -> list[list[int]]:
    """g_table[i][j] = (g^(2^(8i)))^j mod n^2 (lookup.py:104-110, lcu:550-588)."""
    n2 = n * n
    rows = (key_len * 2 + MSG_BITS - 1) // MSG_BITS
    table = []
    base = g % n2
    for _ in range(rows):
        row = [1, base]
        for _ in range(2, 2 ** MSG_BITS):
            row.append(row[-1] * base % n2)
        table.append(row)
        base = pow(base, 2 ** MSG_BITS, n2)
    return table


def build_noise_table(g: int, n: int, rng: Optional[random.Random] = None,
                      size: int = NOISE_TABLE_SIZE) -> list[int]:
    """noise[k] = (g^n)^(r_k) mod n^2 (lookup.py:113-117, lcu:592-626)."""
    n2 = n * n
    g_n = pow(g, n, n2)
    r = _rng(rng)
    return [pow(g_n, lookup_random_r(n, r), n2) for _ in range(size)]
