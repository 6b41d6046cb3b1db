"""Host-side logic and the C ABI surface (CPU only: nothing here launches a kernel).

* the library loads and exports every function include/chx.h declares;
* the host-built exponent schedules, evaluated by the C oracle's Montgomery
  product, give x^e (CPython pow);
* the Montgomery constants and the bit/digit packing agree with the oracle.
"""

import ctypes
import os
import random
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import paillier_oracle as O


@pytest.fixture(scope="module")
def montref():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libmontref.so"))
    P = ctypes.POINTER(ctypes.c_uint32)
    lib.mr_mont_mul.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_int, P]
    lib.mr_modexp.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, P]
    lib.mr_apply_schedule.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P]
    lib.mr_exact_div.argtypes = [P, P, ctypes.c_int, P]
    lib.mr_n0inv.argtypes = [ctypes.c_uint32]
    lib.mr_n0inv.restype = ctypes.c_uint32
    return lib


def arr(x: int, limbs: int) -> np.ndarray:
    return np.array(O.to_limbs(x, limbs), dtype=np.uint32)


def p32(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def test_library_exports_every_declared_symbol():
    from cuhepy_amd import _native as N
    header = open(os.path.join(ROOT, "include", "chx.h")).read()
    declared = set(re.findall(r"\b(chx_\w+)\s*\(", header))
    assert len(declared) > 30
    missing = [s for s in declared if not hasattr(N.lib, s)]
    assert not missing, missing
    # the Python binding covers the same surface
    assert declared == set(N.SIGNATURES), declared ^ set(N.SIGNATURES)
    assert b"gfx950" in N.lib.chx_version()


def test_context_creation_fails_loudly_without_device():
    from cuhepy_amd import _native as N
    h = ctypes.c_void_p()
    rc = N.lib.chx_ctx_create(0, ctypes.byref(h))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert rc != 0 and N.last_error()
    else:
        assert rc == 0
        N.lib.chx_ctx_destroy(h)


# --------------------------------------------------------------------------
# C oracle pinned against CPython integers
# --------------------------------------------------------------------------

@pytest.mark.parametrize("L", [4, 32, 64])
def test_c_oracle_montgomery(montref, L):
    rng = random.Random(L)
    R = 1 << (32 * L)
    for _ in range(20):
        N = rng.getrandbits(32 * L) | 1 | (1 << (32 * L - 1))
        a, b = rng.randrange(N), rng.randrange(N)
        out = np.zeros(L, dtype=np.uint32)
        montref.mr_mont_mul(p32(arr(a, L)), p32(arr(b, L)), p32(arr(N, L)), montref.mr_n0inv(N & 0xFFFFFFFF), L,
                            p32(out))
        assert O.from_limbs(out) == a * b * pow(R, -1, N) % N
    # carries that ripple through every limb: N = R - 1 (odd), operands N-1
    N = R - 1
    out = np.zeros(L, dtype=np.uint32)
    montref.mr_mont_mul(p32(arr(N - 1, L)), p32(arr(N - 1, L)), p32(arr(N, L)), montref.mr_n0inv(N & 0xFFFFFFFF),
                        L, p32(out))
    assert O.from_limbs(out) == (N - 1) ** 2 * pow(R, -1, N) % N


def test_c_oracle_modexp_and_exact_div(montref):
    rng = random.Random(3)
    for L in (8, 32):
        for _ in range(5):
            N = rng.getrandbits(32 * L) | 1 | (1 << (32 * L - 1))
            b, e = rng.randrange(N), rng.getrandbits(32 * L)
            out = np.zeros(L, dtype=np.uint32)
            ea = arr(e, L)
            assert montref.mr_modexp(p32(arr(b, L)), p32(ea), L, p32(arr(N, L)), L, p32(out)) == 0
            assert O.from_limbs(out) == pow(b, e, N)
            # exact division X / M for X = q*M < R (half-width M, as in L(u) = (u-1)/n)
            M = rng.getrandbits(16 * L) | 1 | (1 << (16 * L - 1))
            q = rng.getrandbits(16 * L - 1)
            got = np.zeros(L, dtype=np.uint32)
            montref.mr_exact_div(p32(arr(q * M, L)), p32(arr(M, L)), L, p32(got))
            assert O.from_limbs(got) == q


# --------------------------------------------------------------------------
# product host logic
# --------------------------------------------------------------------------

@pytest.mark.parametrize("window", [1, 2, 3, 4, 5, 6])
def test_host_schedule_evaluates_to_power(montref, window):
    from cuhepy_amd import _native as N
    rng = random.Random(window)
    L = 16
    Nmod = rng.getrandbits(32 * L) | 1 | (1 << (32 * L - 1))
    exps = [1, 2, 3, 5, 0x10000, (1 << 300) + 1, rng.getrandbits(512), rng.getrandbits(511) | (1 << 510)]
    # long zero runs (> 65535 squarings) exercise the split encoding
    exps.append((1 << 70000) + 3)
    for e in exps:
        el = max(1, (e.bit_length() + 31) // 32)
        ea = arr(e, el)
        cap = 8192
        out = np.zeros(cap, dtype=np.uint32)
        n = N.lib.chx_host_schedule(p32(ea), el, window, p32(out), cap)
        assert 0 < n <= cap
        b = rng.randrange(Nmod)
        res = np.zeros(L, dtype=np.uint32)
        rc = montref.mr_apply_schedule(p32(arr(b, L)), p32(out[:n].copy()), n, window, p32(arr(Nmod, L)), L,
                                       p32(res))
        assert rc == 0
        assert O.from_limbs(res) == pow(b, e, Nmod), (window, e.bit_length())


def test_host_mont_consts():
    from cuhepy_amd import _native as N
    rng = random.Random(5)
    for L in (32, 64, 128):
        n = rng.getrandbits(32 * L - 3) | 1
        n0 = (ctypes.c_uint32 * 1)()
        r2 = np.zeros(L, dtype=np.uint32)
        assert N.lib.chx_host_mont_consts(p32(arr(n, L)), L, n0, p32(r2)) == 0
        assert (n0[0] * n) % (1 << 32) == (1 << 32) - 1
        assert O.from_limbs(r2) == pow(2, 64 * L, n)
    # even modulus is rejected
    assert N.lib.chx_host_mont_consts(p32(arr(10, 32)), 32, n0, p32(np.zeros(32, dtype=np.uint32))) != 0


@pytest.mark.parametrize("embed_len,key_len", [(8, 1024), (512, 1024), (64, 512), (1500, 256), (3, 16)])
def test_embedding_packing_matches_oracle(embed_len, key_len):
    from cuhepy_amd.limbs import limbs_to_ints, width_for_bits
    from cuhepy_amd.paillier.packing import embeddings_to_plaintexts, odd_mask
    rng = random.Random(embed_len)
    bits = np.array([[rng.getrandbits(1) for _ in range(embed_len)] for _ in range(7)], dtype=np.uint8)
    W = width_for_bits(4 * key_len)
    got = limbs_to_ints(embeddings_to_plaintexts(bits, key_len, W))
    want = [m for row in bits.tolist() for m in O.embed_to_plaintexts(row, embed_len, key_len)]
    assert got == want
    assert odd_mask(2 * key_len) == O.odd_mask(2 * key_len)


def test_digit_split_matches_oracle():
    from cuhepy_amd.limbs import ints_to_limbs
    from cuhepy_amd.paillier.packing import plaintexts_to_digits
    rng = random.Random(17)
    ms = [0, 1, 255, 256, (1 << 2047) + 12345] + [rng.getrandbits(2048) for _ in range(20)]
    d = plaintexts_to_digits(ints_to_limbs(ms, 128), 256)
    for m, row in zip(ms, d):
        assert row.tolist() == O.lookup_digits(m, 256)


def test_crt_constants_match_oracle():
    from cuhepy_amd.engine import crt_constants
    k = O.key_gen(256, 21)
    cp = O.crt_params(k["p"], k["q"])
    c = crt_constants(k["p"], k["q"])
    assert (c["hp"], c["hq"], c["qinv"]) == (cp["hp"], cp["hq"], cp["q_inv_p"])


def test_product_keygen_is_valid():
    from cuhepy_amd.paillier import keygen as kg
    k = kg.paillier_key_gen(256, random.Random(4))
    p, q, n = k["p"], k["q"], k["pk"]["n"]
    assert p * q == n and p.bit_length() == q.bit_length() == 256
    assert O.is_probable_prime(p, random.Random(1)) and O.is_probable_prime(q, random.Random(2))
    assert kg.factor_from_phi(n, k["sk"]["phi"]) == tuple(sorted((p, q)))
    lk = kg.lookup_key_gen(256, 280, random.Random(8))
    n, n2, g, a = lk["pk"]["n"], lk["pk"]["n_squared"], lk["pk"]["g"], lk["sk"]["a"]
    assert pow(g, n * a, n2) == 1 and pow(g, n, n2) != 1
    assert lk["sk"]["g_a_inv"] == pow(O.L(pow(g, a, n2), n), -1, n)


@pytest.mark.parametrize("key_len", [16, 256, 512])
def test_paillier_factor_check_accepts_valid_and_refuses_each_inconsistency(key_len):
    """chx_host_paillier_check (run by chx_paillier_key_create on keys given
    with factors): a valid key passes; each single inconsistency is refused
    with a message that names it."""
    import math
    from cuhepy_amd import _native as N
    from cuhepy_amd.engine import crt_constants
    k = O.key_gen(key_len, 31 + key_len)
    n, phi, inv, p, q = k["pk"]["n"], k["sk"]["phi"], k["sk"]["inv"], k["p"], k["q"]
    c = crt_constants(p, q)

    def run(**kw):
        v = dict(n=n, phi=phi, inv=inv, p=p, q=q, hp=c["hp"], hq=c["hq"], qinv=c["qinv"])
        v.update(kw)
        big = ("n", "phi", "inv")
        nl = max((v[x].bit_length() + 31) // 32 for x in big)
        hl = max((v[x].bit_length() + 31) // 32 for x in v if x not in big)
        a = {key: arr(val, nl if key in big else hl) for key, val in v.items()}
        rc = N.lib.chx_host_paillier_check(p32(a["n"]), nl, p32(a["phi"]), p32(a["inv"]), p32(a["p"]), p32(a["q"]),
                                           hl, p32(a["hp"]), p32(a["hq"]), p32(a["qinv"]))
        return rc, N.last_error()

    assert run()[0] == N.CHX_OK, run()[1]
    # phi = lambda = lcm(p-1, q-1) with its own inverse is a consistent key too
    lam = (p - 1) * (q - 1) // math.gcd(p - 1, q - 1)
    assert run(phi=lam, inv=pow(lam, -1, n))[0] == N.CHX_OK
    bad_phi = phi + 1 if math.gcd(phi + 1, n) == 1 else phi + 3  # odd: p - 1 does not divide it
    cases = {
        "h_p": dict(hp=(c["hp"] + 1) % p),
        "h_q": dict(hq=(c["hq"] * 2) % q),
        "q_inv": dict(qinv=(c["qinv"] + 1) % p),
        "phi * inv": dict(inv=(inv + 1) % n),
        "lcm": dict(phi=bad_phi, inv=pow(bad_phi, -1, n)),
        "p * q": dict(q=q + 2),
        "prime": dict(p=p * 3, n=p * 3 * q),
        "differ": dict(q=p, n=p * p),
    }
    for what, kw in cases.items():
        rc, msg = run(**kw)
        assert rc == N.CHX_EINVAL, (what, msg)
        assert what in msg, (what, msg)
