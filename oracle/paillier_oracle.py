"""Exact-integer CPU oracle for the Paillier batched-modexp path.

TEST INFRASTRUCTURE ONLY. This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it. The product package ``cuhepy_amd`` never imports anything under
``oracle/``.

It restates, in plain Python integers, the reference scheme of XTraceAI/cuhepy:

* ``src/cuhepy/paillier/scheme.py``   -- textbook Paillier, g = n + 1
* ``src/cuhepy/paillier/lookup.py``   -- the alpha-subgroup "lookup" variant
* ``src/cuhepy/hamming/paillier.py`` and ``hamming/paillier_lookup.py``
  -- bit packing of binary embeddings and the odd-bit Hamming decode

The reference computes with gmpy2 2.3.0 (GMP 6.3.0, ``uv.lock:297-298``) and
draws primes with pycryptodome 3.23.0 ``number.getPrime`` (``uv.lock:731-732``);
neither is installed here. ``gmpy2.powmod(b, e, m)``, ``a * b % m`` and
``gmpy2.invert`` are exact integer functions, so CPython's ``pow`` and ``*``/``%``
give the same integers for the same inputs. The only non-reproducible parts of
the reference are its random draws (``secrets``, ``/dev/urandom``,
``getPrime``); here every random quantity is an explicit argument or comes from
a seeded ``random.Random``, so parity is defined for identical injected keys,
messages, r values and noise indices.

Pinning (see ``tests/test_oracle.py``): closed forms, the (1+n)^m = 1+mn
identity, decrypt(encrypt(m)) = m, homomorphic addition, exhaustive checks at
tiny moduli, the reference's plaintext known answers (``_VECTORS`` distances,
README pair distance 256), agreement with GMP's ``mpz_powm`` through ctypes,
and planted-mistake tests that must fail.
"""

from __future__ import annotations

import math
import random
from typing import Iterable

# lookup.py:11-27
MSG_BITS = 8
NOISE_TABLE_SIZE = 2 ** 8
NOISE_MULTIPLES = 14
MIN_ALPHA_LEN = 256
DEFAULT_ALPHA_LEN = 280

_SMALL_PRIMES = [p for p in range(3, 2000) if all(p % d for d in range(2, int(p ** 0.5) + 1))]


# --------------------------------------------------------------------------
# Seeded primality / prime generation (replaces getPrime / gmpy2.is_prime)
# --------------------------------------------------------------------------

def is_probable_prime(n: int, rng: random.Random, rounds: int = 40) -> bool:
    """Miller-Rabin with ``rounds`` seeded bases (gmpy2.is_prime(p, 40) at lookup.py:181)."""
    if n < 2:
        return False
    for p in [2] + _SMALL_PRIMES:
        if n == p:
            return True
        if n % p == 0:
            return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for _ in range(rounds):
        a = rng.randrange(2, n - 1)
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


def get_prime(bits: int, rng: random.Random, top_two: bool = True) -> int:
    """Seeded stand-in for ``number.getPrime(bits)`` (scheme.py:33-35).

    Top bit set so the prime has exactly ``bits`` bits; with ``top_two`` the
    second bit is set too, so p*q has exactly 2*bits bits.
    """
    while True:
        c = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if top_two:
            c |= 1 << (bits - 2)
        if is_probable_prime(c, rng, 40):
            return c


# --------------------------------------------------------------------------
# Textbook Paillier -- scheme.py
# --------------------------------------------------------------------------

def key_gen(key_len: int, seed: int) -> dict:
    """scheme.py:24-44 with seeded primes: g = n + 1, sk = {phi, phi^-1 mod n}."""
    rng = random.Random(seed)
    p = get_prime(key_len, rng)
    q = get_prime(key_len, rng)
    while q == p:
        q = get_prime(key_len, rng)
    n = p * q
    phi = (p - 1) * (q - 1)
    return {
        "sk": {"phi": phi, "inv": pow(phi, -1, n)},
        "pk": {"g": n + 1, "n": n, "n_squared": n * n},
        # CRT factors (not in the reference key dict; used by decrypt_crt only)
        "p": p,
        "q": q,
    }


def random_r(n: int, rng: random.Random) -> int:
    """scheme.py:47-60: r uniform in [1, n-1] with gcd(r, n) = 1."""
    while True:
        r = rng.randrange(n - 1) + 1
        if math.gcd(r, n) == 1:
            return r


def encrypt(m: int, pk: dict, r: int) -> int:
    """scheme.py:63-83 with an injected r: c = g^m * r^n mod n^2."""
    if not (0 <= m < pk["n"]):
        raise ValueError("plaintext must be in range [0, n)")
    n, n2 = pk["n"], pk["n_squared"]
    if math.gcd(r, n) != 1:
        raise ValueError("random r is not coprime with n")
    a = pow(pk["g"], m, n2)
    b = pow(r, n, n2)
    return (a * b) % n2


def decrypt(c: int, keys: dict) -> int:
    """scheme.py:86-104: m = ((c^phi mod n^2 - 1) // n * inv) mod n."""
    n, n2 = keys["pk"]["n"], keys["pk"]["n_squared"]
    if not (0 <= c < n2):
        raise ValueError("ciphertext must be in range [0, n^2)")
    x = pow(c, keys["sk"]["phi"], n2) - 1
    return ((x // n) * keys["sk"]["inv"]) % n


def add(c1: int, c2: int, pk: dict) -> int:
    """scheme.py:107-119: homomorphic addition = c1 * c2 mod n^2."""
    return (c1 * c2) % pk["n_squared"]


def crt_params(p: int, q: int) -> dict:
    """Constants of the CRT decryption (Paillier 1999, sec. 7), g = n + 1.

    h_p = L_p(g^(p-1) mod p^2)^-1 mod p, L_p(x) = (x - 1) // p; same for q.
    """
    n = p * q
    g = n + 1
    hp = pow((pow(g, p - 1, p * p) - 1) // p, -1, p)
    hq = pow((pow(g, q - 1, q * q) - 1) // q, -1, q)
    return {"p": p, "q": q, "hp": hp, "hq": hq, "q_inv_p": pow(q, -1, p)}


def decrypt_crt(c: int, keys: dict) -> int:
    """CRT decryption, derived from scheme.py:86-104 (no reference counterpart).

    m_p = L_p(c^(p-1) mod p^2) * h_p mod p, m_q likewise, m = CRT(m_p, m_q)
    (Garner: m = m_q + q * ((m_p - m_q) * q^-1 mod p)).
    """
    n2 = keys["pk"]["n_squared"]
    if not (0 <= c < n2):
        raise ValueError("ciphertext must be in range [0, n^2)")
    cp = crt_params(keys["p"], keys["q"])
    p, q = cp["p"], cp["q"]
    mp = ((pow(c % (p * p), p - 1, p * p) - 1) // p) * cp["hp"] % p
    mq = ((pow(c % (q * q), q - 1, q * q) - 1) // q) * cp["hq"] % q
    h = (mp - mq) * cp["q_inv_p"] % p
    return mq + q * h


def factor_from_phi(n: int, phi: int) -> tuple[int, int]:
    """Recover p < q from (n, phi): p + q = n - phi + 1, p*q = n."""
    s = n - phi + 1
    d = math.isqrt(s * s - 4 * n)
    assert d * d == s * s - 4 * n
    return (s - d) // 2, (s + d) // 2


def modexp(b: int, e: int, N: int) -> int:
    """gmpy2.powmod(b, e, N) for e >= 0 (exact)."""
    return pow(b, e, N)


def mulmod(a: int, b: int, N: int) -> int:
    """int(mpz(a) * b % N), the reference evaluator (bfv_client_matrix.py:166-169)."""
    return (a * b) % N


# --------------------------------------------------------------------------
# Lookup variant -- lookup.py
# --------------------------------------------------------------------------

def _bitlen(x: int) -> int:
    """lookup.py:315-320 (bitlen)"""
    return x.bit_length() if x > 0 else 0


def gen_dsa_params_custom(p_bits: int, q_bits: int, rng: random.Random) -> tuple[int, int, int]:
    """lookup.py:150-195 with seeded draws."""
    if q_bits >= p_bits:
        raise ValueError("alpha_len too large: q_bits must be < p_bits")
    q = get_prime(q_bits, rng, top_two=False)
    two_pow_p_1 = 1 << (p_bits - 1)
    two_pow_p = 1 << p_bits
    k_lo = two_pow_p_1 // q
    k_hi = (two_pow_p - 1) // q
    if k_lo % 2 == 1:
        k_lo += 1
    while True:
        span = (k_hi - k_lo) // 2 + 1
        t = rng.randrange(0, span)
        k = k_lo + 2 * t
        p = k * q + 1
        if _bitlen(p) != p_bits:
            continue
        if is_probable_prime(p, rng, 40):
            break
    e = (p - 1) // q
    while True:
        h = rng.randrange(2, p - 1)
        g = pow(h, e, p)
        if g != 1:
            break
    return p, q, g


def crt_pair(a1: int, m1: int, a2: int, m2: int) -> tuple[int, int]:
    """lookup.py:323-329"""
    M = m1 * m2
    inv = pow(m1 % m2, -1, m2)
    t = ((a2 - a1) % m2) * inv % m2
    return (a1 + m1 * t) % M, M


def lift_to_p2_with_p_component(g_mod_p: int, p: int) -> int:
    """lookup.py:332-335"""
    return (g_mod_p * (1 + p)) % (p * p)


def L(x: int, n: int) -> int:
    """lookup.py:309-312"""
    return (x - 1) // n


def lookup_key_gen(key_len: int, alpha_len: int, seed: int, with_tables: bool = True) -> dict:
    """lookup.py:44-100 with seeded draws. The noise table uses seeded r values."""
    rng = random.Random(seed)
    while True:
        q_bits_bound = min(key_len, alpha_len)
        q_bits_1 = q_bits_bound // 2
        q_bits_2 = q_bits_bound - q_bits_1
        p1, q1, g1 = gen_dsa_params_custom(key_len, q_bits_1, rng)
        p2, q2, g2 = gen_dsa_params_custom(key_len, q_bits_2, rng)
        p, q = p1, p2
        n = p * q
        n_sq = n * n
        g1_p2 = lift_to_p2_with_p_component(g1 % p, p)
        g2_q2 = lift_to_p2_with_p_component(g2 % q, q)
        g, _ = crt_pair(g1_p2, p * p, g2_q2, q * q)
        a = q1 * q2 // math.gcd(q1, q2)
        if pow(g, n * a, n_sq) != 1 or pow(g, n, n_sq) == 1:
            continue  # lookup.py:78-81 regenerates
        break
    phi = (p - 1) * (q - 1)
    g_n = pow(g, n, n_sq)
    g_a_inv = pow(L(pow(g, a, n_sq), n), -1, n)
    keys = {
        "sk": {"phi": phi, "a": a, "g_a_inv": g_a_inv},
        "pk": {"g": g, "n": n, "n_squared": n_sq, "g_n": g_n},
        "key_len": key_len,
        "message_chunks": (key_len * 2 + MSG_BITS - 1) // MSG_BITS,
        "p": p,
        "q": q,
    }
    if with_tables:
        keys["g_table"] = precompute_g_table(g, n, key_len)
        rs = [lookup_random_r(n, rng) for _ in range(NOISE_TABLE_SIZE)]
        keys["noise_table"] = precompute_noise_table(g, n, rs)
    return keys


def precompute_g_table(g: int, n: int, key_len: int, msg_bits: int = MSG_BITS) -> dict[int, list[int]]:
    """lookup.py:104-110: g_table[i][j] = (g^(2^(8i)))^j mod n^2."""
    n2 = n * n
    table = {}
    for i in range(0, (key_len * 2 + MSG_BITS - 1) // MSG_BITS):
        base = pow(g, 2 ** (msg_bits * i), n2)
        row = [1]
        for _ in range(1, 2 ** msg_bits):
            row.append(row[-1] * base % n2)
        table[i] = row
    return table


def lookup_random_r(n: int, rng: random.Random) -> int:
    """lookup.py:198-210: r in [1, n] with gcd(r, n) = 1."""
    while True:
        r = rng.randrange(n) + 1
        if math.gcd(r, n) == 1:
            return r


def precompute_noise_table(g: int, n: int, rs: Iterable[int]) -> list[int]:
    """lookup.py:113-117 with injected r values: noise[k] = (g^n)^(r_k) mod n^2."""
    n2 = n * n
    g_n = pow(g, n, n2)
    return [pow(g_n, r, n2) for r in rs]


def lookup_digits(m: int, message_chunks: int) -> list[int]:
    """lookup.py:239: base-2^8 digits of m, little-endian."""
    return [(m >> (MSG_BITS * i)) & (2 ** MSG_BITS - 1) for i in range(message_chunks)]


def lookup_encrypt(m: int, pk: dict, g_table: dict, noise_table: list[int],
                   message_chunks: int, noise_idx: list[int]) -> int:
    """lookup.py:213-252 with injected noise indices (secrets.randbelow(256) there)."""
    if not (0 <= m < pk["n"]):
        raise ValueError("plaintext must be in range [0, n)")
    n_sq = pk["n_squared"]
    g_m = 1
    for i, d in enumerate(lookup_digits(m, message_chunks)):
        g_m = (g_m * g_table[i][d]) % n_sq
    noise = 1
    assert len(noise_idx) == NOISE_MULTIPLES
    for k in noise_idx:
        noise = (noise * noise_table[k]) % n_sq
    return (g_m * noise) % n_sq


def lookup_decrypt(c: int, keys: dict) -> int:
    """lookup.py:257-278: m = L(c^a mod n^2) * g_a_inv mod n."""
    n, n_sq = keys["pk"]["n"], keys["pk"]["n_squared"]
    if not (0 <= c < n_sq):
        raise ValueError("ciphertext must be in range [0, n^2)")
    c_a = L(pow(c, keys["sk"]["a"], n_sq), n)
    return (c_a * keys["sk"]["g_a_inv"]) % n


# --------------------------------------------------------------------------
# Hamming bit packing / decode -- hamming/paillier.py, hamming/paillier_lookup.py
# --------------------------------------------------------------------------

def chunk_geometry(embed_len: int, key_len: int) -> tuple[int, int]:
    """hamming/paillier.py:26-32: (chunk_len, chunk_num)."""
    chunk_len = key_len * 2
    if embed_len > key_len:
        chunk_num = 2 * embed_len // chunk_len + int((2 * embed_len) % chunk_len != 0)
    else:
        chunk_num = 1
    return chunk_len, chunk_num


def embed_to_plaintexts(embd: list[int], embed_len: int, key_len: int) -> list[int]:
    """hamming/paillier.py:81-95: bits b_j -> "0 b_0 0 b_1 ..." -> chunk_len-bit MSB-first ints."""
    if len(embd) != embed_len:
        raise ValueError("embedding length mismatch")
    chunk_len, chunk_num = chunk_geometry(embed_len, key_len)
    padded: list[str] = []
    for b in embd:
        if b not in (0, 1):
            raise ValueError("Embedding vector must be binary")
        padded += ["0", str(b)]
    return [int("".join(padded[i * chunk_len:(i + 1) * chunk_len]), 2) for i in range(chunk_num)]


def decode_hamming(plaintexts: list[int], chunk_len: int) -> int:
    """hamming/paillier.py:97-110: left-pad each plaintext to chunk_len bits, count '1's at odd indices."""
    bits = ""
    for v in plaintexts:
        s = f"{v:b}"
        if len(s) != chunk_len and len(s) != 0:
            s = "0" * (chunk_len - len(s)) + s
        bits += s
    return sum(int(bits[i]) for i in range(1, len(bits), 2))


def odd_mask(chunk_len: int) -> int:
    """cu:410-424: bits at MSB-first odd positions t -> LSB index chunk_len-1-t."""
    mask = 0
    for t in range(1, chunk_len, 2):
        mask |= 1 << (chunk_len - 1 - t)
    return mask


def hamming(a: list[int], b: list[int]) -> int:
    """test_GPU_paillier_clients.py:24-25"""
    return sum(int(x != y) for x, y in zip(a, b))


# --------------------------------------------------------------------------
# Limb codecs (cu:235-279; lcu:526-545): little-endian 32-bit limbs
# --------------------------------------------------------------------------

def to_limbs(x: int, nlimbs: int) -> list[int]:
    if x < 0 or x >> (32 * nlimbs):
        raise ValueError("value does not fit")
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(nlimbs)]


def from_limbs(limbs: Iterable[int]) -> int:
    v = 0
    for i, w in enumerate(limbs):
        v |= int(w) << (32 * i)
    return v
