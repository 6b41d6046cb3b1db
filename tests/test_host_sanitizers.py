"""The host-only C ABI and the host bignum code under AddressSanitizer and
UndefinedBehaviorSanitizer (the reference runs its native bindings' tests
under the same sanitizers, tests/unit/native/run_binding_sanitizers.py).

cuhepy_amd/csrc/host_api.hpp (chx_host_schedule, chx_host_mont_consts,
chx_host_paillier_check, chx_last_error) and host_bn.hpp are plain C++:
tests/native/host_api_driver.cpp includes them and is built here with g++
``-fsanitize=address,undefined -fno-sanitize-recover=all``, so any invalid
access, leak, overflow of a signed int or bad shift ends the run with an
error. The driver is fed a batch of commands; every result is checked against
Python's integers:
* schedules at windows 1-7 decode to their exponent (the decoding is the
  format's definition: first odd window value, then squarings and odd
  multiplies), including a 70,000-bit exponent whose zero runs need the split
  encoding, and e = 0 (empty);
* Montgomery constants n0inv * n = -1 mod 2^32 and R2 = 2^(64 L) mod n at
  1-128 limbs; an even modulus refused;
* Montgomery modexp, product and reduction against pow, * and %;
* the probable-prime test on primes, Carmichael numbers and strong
  pseudoprimes to the small bases;
* the Paillier factor check accepting a valid key and refusing each single
  inconsistency with a message that names it;
* the entry points' argument errors (null pointers, bad sizes).
"""

import math
import os
import random
import shutil
import subprocess

import pytest

from oracle import paillier_oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
DRIVER = os.path.join(ROOT, "tests", "native", "host_api_driver.cpp")
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-Wall", "-Werror"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path_factory.mktemp("san") / "host_api_driver")
    subprocess.run([gxx, *FLAGS, "-o", exe, DRIVER], check=True, cwd=ROOT)

    def run(lines):
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
        env.pop("LD_PRELOAD", None)  # the sanitizer runtime must come first
        p = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, env=env,
                           timeout=300)
        assert p.returncode == 0 and "runtime error" not in p.stderr and "Sanitizer" not in p.stderr, p.stderr[-4000:]
        out = p.stdout.splitlines()
        assert len(out) == len(lines)
        return out
    return run


def _decode_schedule(words, window):
    """Exponent of a schedule: words[0] = (v0 - 1) / 2 of the first odd window
    value v0; every later word = (squarings << 16) | (index + 1), index + 1 = 0
    meaning no multiply, index = (v - 1) / 2 of the odd value v multiplied in."""
    if not words:
        return 0
    e = 2 * words[0] + 1
    assert e < (1 << window)
    for w in words[1:]:
        e <<= w >> 16
        if w & 0xFFFF:
            v = 2 * ((w & 0xFFFF) - 1) + 1
            assert v < (1 << window)
            e += v
    return e


def test_schedules_and_montgomery_constants(driver):
    rng = random.Random(1)
    exps = [0, 1, 2, 3, 0xFFFF, 1 << 64, (1 << 70000) + 3, (1 << 300) + (1 << 150)]
    exps += [rng.getrandbits(rng.choice([7, 33, 512, 1024, 2048, 4096])) | 1 for _ in range(12)]
    cmds, want = [], []
    for e in exps:
        for w in range(1, 8):
            cmds.append(f"S {e:x} {w}")
            want.append((e, w))
    mods = []
    for L in (1, 2, 3, 17, 32, 64, 128):
        for _ in range(3):
            n = rng.getrandbits(32 * L - rng.randrange(0, 3)) | 1
            mods.append((n, L))
    mods += [((1 << (32 * 64)) - 1, 64), (3, 1)]
    cmds += [f"M {n:x} {L}" for n, L in mods] + ["M 10 2"]
    out = driver(cmds)
    for (e, w), line in zip(want, out):
        tok = line.split()
        count, words = int(tok[0]), [int(t, 16) for t in tok[1:]]
        assert count == len(words)
        assert _decode_schedule(words, w) == e, (e.bit_length(), w)
    for (n, L), line in zip(mods, out[len(want):]):
        rc, n0, r2 = line.split()
        assert rc == "0"
        assert (int(n0, 16) * n) % (1 << 32) == (1 << 32) - 1
        assert int(r2, 16) == pow(2, 64 * L, n)
    assert out[-1].split()[0] == "1"  # an even modulus is refused


def test_host_bignum_arithmetic(driver):
    rng = random.Random(2)
    cmds, want = [], []
    for bits in (31, 32, 33, 64, 100, 1024, 2048, 4096):
        for _ in range(4):
            n = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
            b, e = rng.randrange(n), rng.getrandbits(rng.choice([1, 17, bits]))
            cmds.append(f"P {b:x} {e:x} {n:x}")
            want.append(pow(b, e, n))
            a, c = rng.getrandbits(bits), rng.getrandbits(bits)
            cmds += [f"X {a:x} {c:x}", f"D {a * c:x} {n:x}"]
            want += [a * c, a * c % n]
    n = (1 << 521) - 1
    cmds += [f"P 0 5 {n:x}", f"P 7 0 {n:x}", f"P {n - 1:x} 2 {n:x}", "X 0 ffffffff", f"D {n:x} {n:x}"]
    want += [0, 1, 1, 0, 0]
    out = driver(cmds)
    assert [int(x, 16) for x in out] == want


def test_probable_prime(driver):
    primes = [3, 5, 101, 65537, (1 << 61) - 1, (1 << 127) - 1, (1 << 521) - 1, (1 << 607) - 1]
    composites = [9, 561, 41041, 825265, 321197185, 3215031751, 3825123056546413051,  # Carmichael / strong pseudoprimes
                  ((1 << 61) - 1) * ((1 << 89) - 1), ((1 << 127) - 1) ** 2, (1 << 521) + 1]
    out = driver([f"Q {n:x}" for n in primes + composites])
    assert out == ["1"] * len(primes) + ["0"] * len(composites)


@pytest.mark.parametrize("key_len", [16, 256, 512])
def test_paillier_factor_check(driver, key_len):
    from cuhepy_amd.engine import crt_constants
    k = O.key_gen(key_len, 71 + key_len)
    n, phi, inv, p, q = k["pk"]["n"], k["sk"]["phi"], k["sk"]["inv"], k["p"], k["q"]
    c = crt_constants(p, q)
    bad_phi = phi + 1 if math.gcd(phi + 1, n) == 1 else phi + 3
    cases = {
        "": {},
        "h_p": dict(hp=(c["hp"] + 1) % p),
        "h_q": dict(hq=(c["hq"] * 2) % q),
        "q_inv": dict(qinv=(c["qinv"] + 1) % p),
        "phi * inv": dict(inv=(inv + 1) % n),
        "lcm": dict(phi=bad_phi, inv=pow(bad_phi, -1, n)),
        "p * q": dict(q=q + 2),
        "prime": dict(p=p * 3, n=p * 3 * q),
        "differ": dict(q=p, n=p * p),
    }
    cmds = []
    for kw in cases.values():
        v = dict(n=n, phi=phi, inv=inv, p=p, q=q, hp=c["hp"], hq=c["hq"], qinv=c["qinv"])
        v.update(kw)
        hl = max((v[x].bit_length() + 31) // 32 for x in ("p", "q", "hp", "hq", "qinv"))
        cmds.append(f"C {v['n']:x} {v['phi']:x} {v['inv']:x} {v['p']:x} {v['q']:x} {hl} {v['hp']:x} {v['hq']:x} "
                    f"{v['qinv']:x}")
    out = driver(cmds)
    for (what, _), line in zip(cases.items(), out):
        rc, msg = line.split(" ", 1)
        if not what:
            assert (rc, msg) == ("0", "ok")
        else:
            assert rc == "1" and what in msg, (what, line)


def test_argument_errors(driver):
    assert driver(["E"]) == ["-1 -1 -1 1 1 1 1"]
