"""Typed Python handles over the C ABI: device context, moduli and keys.

Everything here calls ``libchx.so``; there is no CPU path. Host-buffer calls
take and return numpy ``uint32`` limb batches (see ``limbs.py``); the
``*_dev`` methods take torch tensors that already live on the GPU in the
limb-interleaved layout (limb j of item i at ``[j, i]`` of a
``(limbs, stride)`` tensor) and enqueue on the context's stream (its own
non-blocking stream unless ``set_stream``/``use_torch_stream`` chose another):
work PyTorch does on those tensors on another stream must be synchronised.
"""

from __future__ import annotations

import ctypes
import threading
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .limbs import ints_to_limbs, limb_count, limbs_below, limbs_to_ints, width_for_bits

_u32p = N.u32p


def _p32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _p8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(N.u8p)


def _batch(x, limbs: int) -> np.ndarray:
    """Accept a (count, limbs) uint32 array or a sequence of ints."""
    if isinstance(x, np.ndarray):
        a = np.ascontiguousarray(x, dtype=np.uint32)
        if a.ndim != 2 or a.shape[1] != limbs:
            raise ValueError(f"expected a (count, {limbs}) uint32 batch, got shape {a.shape}")
        return a
    return ints_to_limbs(x, limbs)


def _dptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(int(t.data_ptr()))


def stride_for(count: int) -> int:
    return max(64, (count + 63) // 64 * 64)


class Context:
    """One device, one stream, reusable device scratch (chx_ctx)."""

    _default: dict[int, "Context"] = {}
    _lock = threading.Lock()

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        N.check(N.lib.chx_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        with cls._lock:
            ctx = cls._default.get(device)
            if ctx is None:
                ctx = cls._default[device] = Context(device)
            return ctx

    def close(self) -> None:
        if getattr(self, "handle", None):
            N.lib.chx_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr: Optional[int]) -> None:
        N.check(N.lib.chx_ctx_set_stream(self.handle, ctypes.c_void_p(stream_ptr or 0)))

    @property
    def stream_ptr(self) -> int:
        """The hipStream_t the library enqueues on (0 = the null stream)."""
        return int(N.lib.chx_ctx_get_stream(self.handle) or 0)

    def use_torch_stream(self) -> None:
        """Put this context and PyTorch on one stream, so their work is ordered.

        A non-default PyTorch stream is handed to the library. PyTorch's default
        stream cannot be (its handle is NULL, which selects the context's own
        non-blocking stream), so in that case PyTorch is switched onto the
        context's stream instead (``torch.cuda.set_stream``)."""
        import torch
        cur = torch.cuda.current_stream(self.device)
        if cur.cuda_stream:
            self.set_stream(cur.cuda_stream)
        else:
            self.set_stream(None)
            torch.cuda.set_stream(torch.cuda.ExternalStream(self.stream_ptr, device=torch.device("cuda", self.device)))

    def sync(self) -> None:
        N.check(N.lib.chx_ctx_sync(self.handle))

    def wall_clock_hz(self) -> float:
        """Frequency of the device wall clock (the stamps of mulmod_multi_stamped_dev)."""
        hz = ctypes.c_double()
        N.check(N.lib.chx_ctx_wall_clock_hz(self.handle, ctypes.byref(hz)))
        return hz.value

    def set_tuning(self, op: int, limbs: int, tpi: int = 0, window: int = 0) -> None:
        N.check(N.lib.chx_ctx_set_tuning(self.handle, op, limbs, tpi, window))

    @staticmethod
    def supported_tpi(limbs: int, op: int = 1) -> list[int]:
        """Lanes per bignum built for `op` at `limbs` (op 0 = exponentiation)."""
        buf = (ctypes.c_int * 16)()
        n = N.lib.chx_supported_tpi_op(op, limbs, buf, 16)
        return list(buf[:n])

    def scratch_bytes(self) -> int:
        return int(N.lib.chx_ctx_scratch_bytes(self.handle))

    def set_timing(self, enable: bool) -> None:
        N.check(N.lib.chx_ctx_set_timing(self.handle, int(enable)))

    def kernel_time(self, name: str) -> tuple[float, int]:
        ms, n = ctypes.c_double(), ctypes.c_int()
        N.check(N.lib.chx_ctx_kernel_time(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def reset_timing(self) -> None:
        N.check(N.lib.chx_ctx_reset_timing(self.handle))

    def probe_mad(self, blocks: int, iters: int) -> tuple[float, float]:
        s, m = ctypes.c_double(), ctypes.c_double()
        N.check(N.lib.chx_probe_mad(self.handle, blocks, iters, ctypes.byref(s), ctypes.byref(m)))
        return s.value, m.value

    def probe_rate(self, kind: int, blocks: int, iters: int) -> tuple[float, float]:
        s, m = ctypes.c_double(), ctypes.c_double()
        N.check(N.lib.chx_probe_rate(self.handle, kind, blocks, iters, ctypes.byref(s), ctypes.byref(m)))
        return s.value, m.value

    # layout conversion on device
    def probe_copy(self, src, dst, count: int, stride: int, limbs: int, tpi: int) -> None:
        """Calibration copy with the lane-group access pattern (byte counters)."""
        N.check(N.lib.chx_probe_copy(self.handle, _dptr(src), _dptr(dst), count, stride, limbs, tpi))

    def aos_to_il(self, aos, il, count: int, limbs: int, stride: int) -> None:
        N.check(N.lib.chx_aos_to_il_dev(self.handle, _dptr(aos), _dptr(il), count, limbs, stride))

    def il_to_aos(self, il, aos, count: int, limbs: int, stride: int) -> None:
        N.check(N.lib.chx_il_to_aos_dev(self.handle, _dptr(il), _dptr(aos), count, limbs, stride))


class Modulus:
    """An odd modulus with its Montgomery constants resident on the device."""

    def __init__(self, ctx: Context, n: int, limbs: Optional[int] = None):
        if n <= 1 or n % 2 == 0:
            raise ValueError("modulus must be odd and > 1")
        self.ctx = ctx
        self.n = n
        self.limbs = limbs or width_for_bits(n.bit_length())
        arr = ints_to_limbs([n], self.limbs)
        h = ctypes.c_void_p()
        N.check(N.lib.chx_mod_create(ctx.handle, _p32(arr), self.limbs, ctypes.byref(h)))
        self.handle = h

    def __del__(self):  # pragma: no cover
        try:
            if getattr(self, "handle", None):
                N.lib.chx_mod_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def mulmod(self, a, b) -> np.ndarray:
        A, B = _batch(a, self.limbs), _batch(b, self.limbs)
        if A.shape != B.shape:
            raise ValueError("a and b must have the same length")
        out = np.empty_like(A)
        if len(A):
            N.check(N.lib.chx_mulmod(self.ctx.handle, self.handle, _p32(A), _p32(B), _p32(out), len(A)))
        return out

    def product(self, a) -> int:
        """The product of all rows mod n, as a tree of batched mulmods (log2
        count launches); 1 for an empty batch. A product is a unit mod n
        exactly when every factor is, so one gcd with n checks a batch."""
        A = _batch(a, self.limbs)
        if len(A) == 0:
            return 1
        if not limbs_below(A, self.n).all():
            raise ValueError("factors must be below the modulus")
        while len(A) > 1:
            h = len(A) // 2
            P = self.mulmod(A[:h], A[h:2 * h])
            A = np.concatenate([P, A[2 * h:]]) if len(A) % 2 else P
        return limbs_to_ints(A)[0]

    def modexp(self, bases, e: int) -> np.ndarray:
        if e < 0:
            raise ValueError("negative exponent")
        B = _batch(bases, self.limbs)
        E = ints_to_limbs([e], limb_count(e))
        out = np.empty_like(B)
        if len(B):
            N.check(N.lib.chx_modexp(self.ctx.handle, self.handle, _p32(B), len(B), _p32(E), E.shape[1],
                                     _p32(out)))
        return out

    def modexp_each(self, bases, exps) -> np.ndarray:
        """bases[i] ** exps[i] mod n (per-item exponents)."""
        B = _batch(bases, self.limbs)
        ex = [int(x) for x in exps]
        if len(ex) != len(B):
            raise ValueError("bases and exponents must have the same length")
        if any(x < 0 for x in ex):
            raise ValueError("negative exponent")
        el = max([limb_count(x) for x in ex] + [1])
        E = ints_to_limbs(ex, el)
        out = np.empty_like(B)
        if len(B):
            N.check(N.lib.chx_modexp_each(self.ctx.handle, self.handle, _p32(B), len(B), _p32(E), el, _p32(out)))
        return out

    def modexp_each_dev(self, base, exps, e_limbs: int, e_bits: int, out, count: int, stride: int) -> None:
        N.check(N.lib.chx_modexp_each_dev(self.ctx.handle, self.handle, _dptr(base), _dptr(exps), e_limbs, e_bits,
                                          _dptr(out), count, stride))

    # device-resident forms
    def mulmod_dev(self, a, b, out, count: int, stride: int) -> None:
        N.check(N.lib.chx_mulmod_dev(self.ctx.handle, self.handle, _dptr(a), _dptr(b), _dptr(out), count, stride))

    def mulmod_bcast_dev(self, x, y: int, out, count: int, stride: int) -> None:
        Y = ints_to_limbs([y], self.limbs)
        N.check(N.lib.chx_mulmod_bcast_dev(self.ctx.handle, self.handle, _dptr(x), _p32(Y), _dptr(out), count,
                                           stride))

    @property
    def operand_words(self) -> int:
        """Words of one multiplier operand for mulmod_multi_dev."""
        return int(N.lib.chx_mulmod_operand_words(self.handle))

    def operands_dev(self, y, ops, count: int) -> None:
        """ops[j] = operand form of y[j] (device; y: count x limbs, ops: count x operand_words)."""
        N.check(N.lib.chx_mulmod_operands_dev(self.ctx.handle, self.handle, _dptr(y), _dptr(ops), count))

    def operands(self, y: np.ndarray, ops) -> None:
        """ops[j] = operand form of host limbs y[j] (count x limbs), enqueued on the
        context's stream without synchronising."""
        Y = np.ascontiguousarray(y, dtype=np.uint32)
        if Y.ndim != 2 or Y.shape[1] != self.limbs:
            raise ValueError(f"y must be (count, {self.limbs}) limbs")
        N.check(N.lib.chx_mulmod_operands(self.ctx.handle, self.handle, _p32(Y), _dptr(ops), Y.shape[0]))

    def mulmod_multi_dev(self, x, stride: int, begin: int, count: int, ops: Sequence, outs: Sequence,
                         out_stride: int) -> None:
        """outs[q][:, i] = x[:, i] * y_q mod N for i in [begin, begin + count), one launch
        (x and every outs[q] interleaved; ops[q] from operands_dev)."""
        nq = len(ops)
        if nq != len(outs):
            raise ValueError("ops and outs differ in length")
        P = (ctypes.c_void_p * max(nq, 1))(*[int(o.data_ptr()) for o in ops])
        O = (ctypes.c_void_p * max(nq, 1))(*[int(o.data_ptr()) for o in outs])
        N.check(N.lib.chx_mulmod_multi_dev(self.ctx.handle, self.handle, _dptr(x), stride, begin, count, nq, P, O,
                                           out_stride))

    def mulmod_multi_stamped_dev(self, x, stride: int, begin: int, count: int, ops: Sequence, outs: Sequence,
                                 out_stride: int, done, stamps) -> None:
        """mulmod_multi_dev, and stamps[q] (int64 device tensor) = the device
        clock when query q's last tile finished (done: int32 device scratch of
        at least nq + 1 words: tile counters and the unit queue's head)."""
        if done.numel() < len(ops) + 1 or stamps.numel() < len(ops):
            raise ValueError("done needs nq + 1 words and stamps nq")
        nq = len(ops)
        if nq != len(outs):
            raise ValueError("ops and outs differ in length")
        P = (ctypes.c_void_p * max(nq, 1))(*[int(o.data_ptr()) for o in ops])
        O = (ctypes.c_void_p * max(nq, 1))(*[int(o.data_ptr()) for o in outs])
        N.check(N.lib.chx_mulmod_multi_stamped_dev(self.ctx.handle, self.handle, _dptr(x), stride, begin, count, nq,
                                                   P, O, out_stride, _dptr(done), _dptr(stamps)))

    def modexp_dev(self, base, out, count: int, stride: int, e: int) -> None:
        E = ints_to_limbs([e], limb_count(e))
        N.check(N.lib.chx_modexp_dev(self.ctx.handle, self.handle, _dptr(base), _dptr(out), count, stride,
                                     _p32(E), E.shape[1]))


def crt_constants(p: int, q: int) -> dict:
    """h_p = L_p(g^(p-1) mod p^2)^-1 mod p for g = n + 1 (and h_q), q^-1 mod p.

    For g = n + 1, g^(p-1) = 1 + (p-1) n (mod p^2), so L_p(g^(p-1)) =
    (p-1) q mod p = -q mod p and h_p = (-q)^-1 mod p; likewise h_q = (-p)^-1 mod q.
    """
    return {"hp": pow(-q % p, -1, p), "hq": pow(-p % q, -1, q), "qinv": pow(q, -1, p)}


def crt_idempotents(p: int, q: int) -> tuple[int, int]:
    """e_p = q^2 (q^-2 mod p^2) and e_q = p^2 (p^-2 mod q^2): e_p is 1 mod p^2 and
    0 mod q^2, e_q the converse, so x_p e_p + x_q e_q mod n^2 is the x with
    those residues."""
    p2, q2 = p * p, q * q
    n2 = p2 * q2
    return q2 * pow(q2, -1, p2) % n2, p2 * pow(p2, -1, q2) % n2


class PaillierKey:
    """Device-resident Paillier key (g = n + 1): encrypt, decrypt, CRT decrypt."""

    def __init__(self, ctx: Context, n: int, phi: int, inv: int, p: Optional[int] = None,
                 q: Optional[int] = None):
        self.ctx = ctx
        self.n, self.phi, self.inv = n, phi, inv
        self.n2 = n * n
        nl = limb_count(n)
        self.width = width_for_bits(self.n2.bit_length())
        self.half = self.width // 2
        nn = ints_to_limbs([n], nl)
        ph = ints_to_limbs([phi], nl)
        iv = ints_to_limbs([inv], nl)
        h = ctypes.c_void_p()
        # CRT kernels are built for 32- and 64-limb halves (n^2 of 2048 or 4096 bits)
        self.crt = p is not None and q is not None and self.half in (32, 64)
        if self.crt:
            if p * q != n:
                raise ValueError("p * q != n")
            if p < q:  # Garner recombination is symmetric; keep the given order
                pass
            c = crt_constants(p, q)
            hl = max(limb_count(p), limb_count(q), limb_count(c["hp"]), limb_count(c["hq"]))
            arrs = [ints_to_limbs([v], hl) for v in (p, q, c["hp"], c["hq"], c["qinv"])]
            N.check(N.lib.chx_paillier_key_create(ctx.handle, _p32(nn), nl, _p32(ph), _p32(iv), _p32(arrs[0]),
                                                  _p32(arrs[1]), hl, _p32(arrs[2]), _p32(arrs[3]), _p32(arrs[4]),
                                                  ctypes.byref(h)))
        else:
            N.check(N.lib.chx_paillier_key_create(ctx.handle, _p32(nn), nl, _p32(ph), _p32(iv), None, None, 0,
                                                  None, None, None, ctypes.byref(h)))
        self.handle = h
        assert N.lib.chx_paillier_width(h) == self.width
        if self.crt:  # exact half-width decrypt / decode (same results as the full-width form)
            E = ints_to_limbs(list(crt_idempotents(p, q)), self.width)
            N.check(N.lib.chx_paillier_key_set_idempotents(ctx.handle, h, _p32(E[0:1]), _p32(E[1:2])))

    @property
    def exact_crt(self) -> bool:
        """decrypt / decode_hamming run c^phi at half width (recombined exactly)."""
        return bool(N.lib.chx_paillier_key_has_idempotents(self.handle))

    def close(self) -> None:
        """Free the key's device memory (before its context is closed)."""
        if getattr(self, "handle", None):
            N.lib.chx_paillier_key_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def encrypt(self, m, r) -> np.ndarray:
        M, R = _batch(m, self.width), _batch(r, self.width)
        if M.shape != R.shape:
            raise ValueError("m and r must have the same length")
        out = np.empty_like(M)
        if len(M):
            N.check(N.lib.chx_paillier_encrypt(self.ctx.handle, self.handle, _p32(M), _p32(R), _p32(out), len(M)))
        return out

    def decrypt(self, c) -> np.ndarray:
        C = _batch(c, self.width)
        out = np.empty_like(C)
        if len(C):
            N.check(N.lib.chx_paillier_decrypt(self.ctx.handle, self.handle, _p32(C), _p32(out), len(C)))
        return out

    def decrypt_crt(self, c) -> np.ndarray:
        if not self.crt:
            raise ValueError("key has no CRT factors")
        C = _batch(c, self.width)
        out = np.empty((len(C), self.half), dtype=np.uint32)
        if len(C):
            N.check(N.lib.chx_paillier_decrypt_crt(self.ctx.handle, self.handle, _p32(C), _p32(out), len(C)))
        return out

    def decode_hamming(self, c, mask: int) -> np.ndarray:
        C = _batch(c, self.width)
        Mk = ints_to_limbs([mask], self.width)
        out = np.empty(len(C), dtype=np.uint32)
        if len(C):
            N.check(N.lib.chx_paillier_decode_hamming(self.ctx.handle, self.handle, _p32(C), _p32(Mk), _p32(out),
                                                      len(C)))
        return out

    # device-resident forms (interleaved torch tensors)
    def encrypt_dev(self, m, r, out, count: int, stride: int) -> None:
        N.check(N.lib.chx_paillier_encrypt_dev(self.ctx.handle, self.handle, _dptr(m), _dptr(r), _dptr(out), count,
                                               stride))

    def decrypt_dev(self, c, out, count: int, stride: int) -> None:
        N.check(N.lib.chx_paillier_decrypt_dev(self.ctx.handle, self.handle, _dptr(c), _dptr(out), count, stride))

    def decrypt_crt_dev(self, c, out, count: int, stride: int) -> None:
        N.check(N.lib.chx_paillier_decrypt_crt_dev(self.ctx.handle, self.handle, _dptr(c), _dptr(out), count,
                                                   stride))


class LookupKey:
    """Device-resident Paillier-lookup key: tables in HBM (Montgomery form).

    Empty tables give a decrypt-only key.
    """

    def __init__(self, ctx: Context, n: int, a: int, g_a_inv: int, g_table: Sequence[Sequence[int]],
                 noise_table: Sequence[int]):
        self.ctx = ctx
        self.n, self.a, self.g_a_inv = n, a, g_a_inv
        self.n2 = n * n
        nl = limb_count(n)
        self.width = width_for_bits(self.n2.bit_length())
        if isinstance(g_table, np.ndarray):  # (chunks, 256, width) limb array
            if g_table.ndim != 3 or g_table.shape[1:] != (256, self.width):
                raise ValueError(f"g_table array must be (chunks, 256, {self.width})")
            self.chunks = g_table.shape[0]
            G = np.ascontiguousarray(g_table.reshape(-1, self.width), dtype=np.uint32)
        else:
            self.chunks = len(g_table)
            if any(len(row) != 256 for row in g_table):
                raise ValueError("every g_table row must have 256 entries")
            G = ints_to_limbs([v for row in g_table for v in row], self.width)
        NT = _batch(noise_table, self.width) if isinstance(noise_table, np.ndarray) else \
            ints_to_limbs(list(noise_table), self.width)
        if (self.chunks == 0) != (len(NT) == 0):
            raise ValueError("give both tables or neither (decrypt-only key)")
        nn = ints_to_limbs([n], nl)
        A = ints_to_limbs([a], limb_count(a))
        GA = ints_to_limbs([g_a_inv], nl)
        h = ctypes.c_void_p()
        N.check(N.lib.chx_lookup_key_create(ctx.handle, _p32(nn), nl, _p32(A), A.shape[1], _p32(GA),
                                            _p32(G) if self.chunks else None, self.chunks,
                                            _p32(NT) if len(NT) else None, len(NT), ctypes.byref(h)))
        self.handle = h
        self.noise_entries = len(NT)

    def __del__(self):  # pragma: no cover
        try:
            if getattr(self, "handle", None):
                N.lib.chx_lookup_key_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_crt(self, p: int, q: int) -> None:
        """Decrypt at half width from now on: c^a mod p^2 and mod q^2,
        recombined to c^a mod n^2 with the idempotents e_p, e_q, so every
        c in [0, n^2) decrypts as the full-width form (lookup.py:257-278).
        Raises (key unchanged) when p q != n or the width has no CRT form."""
        if p * q != self.n:
            raise ValueError("p * q != n")
        ep, eq = crt_idempotents(p, q)
        hl = max(limb_count(p), limb_count(q))
        P, Q = ints_to_limbs([p], hl), ints_to_limbs([q], hl)
        E = ints_to_limbs([ep, eq], self.width)
        N.check(N.lib.chx_lookup_key_set_crt(self.ctx.handle, self.handle, _p32(P), _p32(Q), hl, _p32(E[0:1]),
                                             _p32(E[1:2])))

    @property
    def crt(self) -> bool:
        return bool(N.lib.chx_lookup_key_has_crt(self.handle))

    def set_table_rows(self, rows: int) -> None:
        """Build the row-group tables over the first `rows` table rows only
        (0 = every row): plaintexts with zero digits above them (packed
        embeddings of chunk_len bits use ceil(chunk_len / 8) rows) need no
        more. Larger digits stay exact, one product per row."""
        N.check(N.lib.chx_lookup_key_set_table_rows(self.ctx.handle, self.handle, int(rows)))

    def encrypt_digits(self, digits: np.ndarray, noise_idx: np.ndarray) -> np.ndarray:
        D = np.ascontiguousarray(digits, dtype=np.uint8)
        NI = np.ascontiguousarray(noise_idx, dtype=np.uint8)
        if D.ndim != 2 or D.shape[1] != self.chunks:
            raise ValueError(f"digits must be (count, {self.chunks}) bytes")
        if NI.ndim != 2 or NI.shape[0] != D.shape[0]:
            raise ValueError("noise_idx must be (count, nmult) bytes")
        out = np.empty((len(D), self.width), dtype=np.uint32)
        if len(D):
            N.check(N.lib.chx_lookup_encrypt(self.ctx.handle, self.handle, _p8(D), _p8(NI), NI.shape[1], _p32(out),
                                             len(D)))
        return out

    def encrypt_digits_dev(self, digits, noise_idx, nmult: int, out, count: int, stride: int) -> None:
        N.check(N.lib.chx_lookup_encrypt_dev(self.ctx.handle, self.handle, _dptr(digits), _dptr(noise_idx), nmult,
                                             _dptr(out), count, stride))

    def decrypt(self, c) -> np.ndarray:
        C = _batch(c, self.width)
        out = np.empty_like(C)
        if len(C):
            N.check(N.lib.chx_lookup_decrypt(self.ctx.handle, self.handle, _p32(C), _p32(out), len(C)))
        return out

    def decode_hamming(self, c, mask: int) -> np.ndarray:
        C = _batch(c, self.width)
        Mk = ints_to_limbs([mask], self.width)
        out = np.empty(len(C), dtype=np.uint32)
        if len(C):
            N.check(N.lib.chx_lookup_decode_hamming(self.ctx.handle, self.handle, _p32(C), _p32(Mk), _p32(out),
                                                    len(C)))
        return out

    def reencrypt(self, c, target: "LookupKey", noise_idx: np.ndarray) -> np.ndarray:
        """Decrypt under this key and encrypt under `target` (its tables) with
        the given noise indices, the digits staying on the device
        (lcu:1053-1135, lcu:2134-2285)."""
        C = _batch(c, self.width)
        NI = np.ascontiguousarray(noise_idx, dtype=np.uint8)
        if NI.ndim != 2 or NI.shape[0] != len(C):
            raise ValueError("noise_idx must be (count, nmult) bytes")
        out = np.empty((len(C), target.width), dtype=np.uint32)
        if len(C):
            N.check(N.lib.chx_lookup_reencrypt(self.ctx.handle, self.handle, target.handle, _p32(C), _p8(NI),
                                               NI.shape[1], _p32(out), len(C)))
        return out

    def reencrypt_dev(self, c, target: "LookupKey", noise_idx, nmult: int, out, count: int, stride: int) -> None:
        N.check(N.lib.chx_lookup_reencrypt_dev(self.ctx.handle, self.handle, target.handle, _dptr(c),
                                               _dptr(noise_idx), nmult, _dptr(out), count, stride))


class RaggedPlan:
    """Mixed-width, per-item-exponent modexp over many sub-batches in one
    persistent launch captured into a hipGraph (chx_ragged_*).

    `subs`: dicts with keys mod (Modulus of 64 or 128 limbs), base, exps, out
    (device tensors, interleaved), count, stride, e_limbs, e_bits. The tensors
    must stay alive while the plan exists (the plan keeps references).
    """

    def __init__(self, ctx: Context, subs: list, variant: int = 0, window: int = 0, waves_per_cu: int = 0):
        self.ctx = ctx
        self._keep = subs
        arr = (N.RaggedSubStruct * len(subs))()
        for i, s in enumerate(subs):
            arr[i].mod = s["mod"].handle.value
            arr[i].base = int(s["base"].data_ptr())
            arr[i].exps = int(s["exps"].data_ptr())
            arr[i].out = int(s["out"].data_ptr())
            arr[i].count, arr[i].stride = s["count"], s["stride"]
            arr[i].e_limbs, arr[i].e_bits = s["e_limbs"], s["e_bits"]
        h = ctypes.c_void_p()
        N.check(N.lib.chx_ragged_create(ctx.handle, arr, len(subs), variant, window, waves_per_cu,
                                        ctypes.byref(h)))
        self.handle = h
        self.variant, self.window, self.waves_per_cu = variant, window, waves_per_cu

    def run(self, graph: bool = True) -> None:
        N.check(N.lib.chx_ragged_run(self.handle, int(graph)))

    def run_plain(self) -> None:
        N.check(N.lib.chx_ragged_run_plain(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            N.lib.chx_ragged_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
