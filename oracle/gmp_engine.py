"""GMP (libgmp.so.10) through ctypes: the CPU engine the reference's gmpy2 wraps.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline; see paillier_oracle.py).

The reference computes ``gmpy2.powmod`` (scheme.py:81-83, 101-104;
lookup.py:108-116, 243-252, 275-278), which is GMP's ``mpz_powm``; ``a * b % m``
is ``mpz_mul`` + ``mpz_mod``. gmpy2 is not installed here, but the GMP runtime
is, so this module calls the same GMP routines directly. It serves two roles:

* a second, independent exact engine that the pure-Python oracle is checked
  against (CPython ``pow`` vs GMP ``mpz_powm``);
* the CPU baseline timed beside the GPU in ``bench.py`` (single thread and one
  process per host core).

When libgmp cannot be loaded, ``available()`` is False and callers fall back to
CPython ``pow`` (and say so).
"""

from __future__ import annotations

import ctypes
import ctypes.util
import os
import types

_LIB = None


class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def _load():
    global _LIB
    if _LIB is not None:
        return _LIB
    for name in ("libgmp.so.10", ctypes.util.find_library("gmp") or ""):
        if not name:
            continue
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        _LIB = False
        return _LIB
    P = ctypes.POINTER(_Mpz)
    lib.__gmpz_init.argtypes = [P]
    lib.__gmpz_clear.argtypes = [P]
    lib.__gmpz_import.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t,
                                  ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
    lib.__gmpz_export.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int,
                                  ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, P]
    lib.__gmpz_export.restype = ctypes.c_void_p
    for fn in ("__gmpz_powm", ):
        getattr(lib, fn).argtypes = [P, P, P, P]
    for fn in ("__gmpz_mul", "__gmpz_mod", "__gmpz_divexact", "__gmpz_add", "__gmpz_sub"):
        getattr(lib, fn).argtypes = [P, P, P]
    lib.__gmpz_sub_ui.argtypes = [P, P, ctypes.c_ulong]
    lib.__gmpz_sizeinbase.argtypes = [P, ctypes.c_int]
    lib.__gmpz_sizeinbase.restype = ctypes.c_size_t
    _LIB = types.SimpleNamespace()
    for k in ("init", "clear", "import", "export", "powm", "mul", "mod", "divexact", "add", "sub",
              "sub_ui", "sizeinbase"):
        setattr(_LIB, k, getattr(lib, "__gmpz_" + k))
    _LIB.raw = lib
    return _LIB


def available() -> bool:
    return bool(_load())


def version() -> str:
    lib = _load()
    if not lib:
        return "unavailable"
    try:
        return ctypes.c_char_p.in_dll(lib.raw, "__gmp_version").value.decode()
    except ValueError:
        return "unknown"


class Mpz:
    """Owning wrapper of one mpz_t."""

    __slots__ = ("_z", "_lib")

    def __init__(self, value: int = 0):
        self._lib = _load()
        self._z = _Mpz()
        self._lib.init(ctypes.byref(self._z))
        if value:
            self.set(value)

    def __del__(self):
        if getattr(self, "_z", None) is not None and self._lib:
            self._lib.clear(ctypes.byref(self._z))

    @property
    def ref(self):
        return ctypes.byref(self._z)

    def set(self, value: int) -> "Mpz":
        if value < 0:
            raise ValueError("negative values are not used here")
        nbytes = max(1, (value.bit_length() + 7) // 8)
        buf = value.to_bytes(nbytes, "little")
        getattr(self._lib, 'import')(self.ref, nbytes, -1, 1, -1, 0, buf)
        return self

    def get(self) -> int:
        if self._z.size == 0:
            return 0
        nbytes = (self._lib.sizeinbase(self.ref, 2) + 7) // 8
        buf = ctypes.create_string_buffer(nbytes)
        cnt = ctypes.c_size_t(0)
        self._lib.export(buf, ctypes.byref(cnt), -1, 1, -1, 0, self.ref)
        return int.from_bytes(buf.raw[: cnt.value], "little")


def powm(b: int, e: int, m: int) -> int:
    """mpz_powm(b, e, m) -- gmpy2.powmod."""
    lib = _load()
    r, B, E, M = Mpz(), Mpz(b), Mpz(e), Mpz(m)
    lib.powm(r.ref, B.ref, E.ref, M.ref)
    return r.get()


def mulmod(a: int, b: int, m: int) -> int:
    lib = _load()
    r, A, B, M = Mpz(), Mpz(a), Mpz(b), Mpz(m)
    lib.mul(r.ref, A.ref, B.ref)
    lib.mod(r.ref, r.ref, M.ref)
    return r.get()


class CrtDecryptor:
    """CRT Paillier decryption with GMP (the a19 formula of paillier_oracle.decrypt_crt).

    Constants are converted to mpz once; ``__call__`` decrypts one ciphertext.
    """

    def __init__(self, p: int, q: int, hp: int, hq: int, q_inv_p: int):
        self.lib = _load()
        self.p, self.q = Mpz(p), Mpz(q)
        self.p2, self.q2 = Mpz(p * p), Mpz(q * q)
        self.pm1, self.qm1 = Mpz(p - 1), Mpz(q - 1)
        self.hp, self.hq, self.qinv = Mpz(hp), Mpz(hq), Mpz(q_inv_p)
        self.t = [Mpz() for _ in range(4)]

    def __call__(self, c: int) -> int:
        lib = self.lib
        C = Mpz(c)
        t0, t1, t2, t3 = self.t
        # m_p
        lib.mod(t0.ref, C.ref, self.p2.ref)
        lib.powm(t0.ref, t0.ref, self.pm1.ref, self.p2.ref)
        lib.sub_ui(t0.ref, t0.ref, 1)
        lib.divexact(t0.ref, t0.ref, self.p.ref)
        lib.mul(t0.ref, t0.ref, self.hp.ref)
        lib.mod(t0.ref, t0.ref, self.p.ref)
        # m_q
        lib.mod(t1.ref, C.ref, self.q2.ref)
        lib.powm(t1.ref, t1.ref, self.qm1.ref, self.q2.ref)
        lib.sub_ui(t1.ref, t1.ref, 1)
        lib.divexact(t1.ref, t1.ref, self.q.ref)
        lib.mul(t1.ref, t1.ref, self.hq.ref)
        lib.mod(t1.ref, t1.ref, self.q.ref)
        # Garner: m = m_q + q * ((m_p - m_q) * q^-1 mod p)
        lib.sub(t2.ref, t0.ref, t1.ref)
        lib.mul(t2.ref, t2.ref, self.qinv.ref)
        lib.mod(t2.ref, t2.ref, self.p.ref)
        lib.mul(t3.ref, t2.ref, self.q.ref)
        lib.add(t3.ref, t3.ref, t1.ref)
        return t3.get()


class Evaluator:
    """The reference's CPU evaluate loop, int(mpz(a) * y % n^2) per index
    entry (benchmarks/bfv_client_matrix.py:166-169), with GMP: y and n^2 are
    converted once; each call converts a, multiplies, reduces and returns an
    int, as that loop does."""

    def __init__(self, y: int, n_squared: int):
        self.lib = _load()
        self.y, self.m = Mpz(y), Mpz(n_squared)
        self.a, self.r = Mpz(), Mpz()

    def __call__(self, a: int) -> int:
        lib = self.lib
        self.a.set(a)
        lib.mul(self.r.ref, self.a.ref, self.y.ref)
        lib.mod(self.r.ref, self.r.ref, self.m.ref)
        return self.r.get()


class LookupEncryptor:
    """PaillierLookup.encrypt (lookup.py:213-252) with GMP: c = prod_i
    g_table[i][digit_i(m)] * prod_k noise[idx_k] mod n^2, one table mulmod per
    8-bit digit (zero digits included, as the reference loops over every
    chunk) and one per noise factor. Table entries are converted to mpz once
    and cached (the reference's tables are mpz already)."""

    def __init__(self, table, noise: list[int], n_squared: int, chunks: int):
        self.lib = _load()
        self.table, self.chunks = table, chunks
        self.noise = [Mpz(v) for v in noise]
        self.m = Mpz(n_squared)
        self.cache: dict = {}
        self.acc = Mpz()

    def entry(self, i: int, d: int) -> "Mpz":
        e = self.cache.get((i, d))
        if e is None:
            e = self.cache[(i, d)] = Mpz(self.table[i][d])
        return e

    def prepare(self, m: int) -> None:
        for i in range(self.chunks):
            self.entry(i, (m >> (8 * i)) & 0xFF)

    def __call__(self, m: int, noise_idx) -> int:
        lib, acc, mod = self.lib, self.acc, self.m.ref
        acc.set(1)
        for i in range(self.chunks):
            lib.mul(acc.ref, acc.ref, self.entry(i, (m >> (8 * i)) & 0xFF).ref)
            lib.mod(acc.ref, acc.ref, mod)
        for k in noise_idx:
            lib.mul(acc.ref, acc.ref, self.noise[k].ref)
            lib.mod(acc.ref, acc.ref, mod)
        return acc.get()


def host_cores() -> int:
    """Host cores this process may use (affinity-aware)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
