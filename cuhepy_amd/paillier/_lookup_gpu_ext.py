"""``PaillierGPULookupClient`` -- drop-in for ``cuhepy.paillier._lookup_gpu_ext``.

Same constructor, methods and return types as the pybind11 class of
``src/cuhepy/paillier/_lookup_gpu_ext/paillier_GPU_lookup_client.cu:1564-2721``
(alias ``PaillierGPUClient`` as in lcu:2669, 2720). Ciphertexts are returned
as ``gmpy2.mpz`` when gmpy2 is importable, else as ``int`` (the reference
returns mpz, lcu:1331-1334; its Python wrappers normalise to int,
``hamming/paillier_lookup.py:304-308``).

Work placement:
  * table precompute (lcu:550-626): g_table rows by device mulmods (row j =
    row j-1 * base, one batched call per column), noise table by one batched
    per-item-exponent modexp; both converted to Montgomery form and kept
    resident in HBM (lcu:2625-2652);
  * encrypt (lcu:1873-2002, 2006-2131): digits and noise indices on the host,
    the table-product chain on the GPU;
  * decrypt / decode (lcu:2338-2499): c^a on the GPU, L(.) * g_a_inv and the
    odd-bit popcount fused after it;
  * decrypt_then_reencrypt (lcu:2134-2285): decrypt under key A and table
    re-encrypt under key B, both on the GPU.
Noise indices are uniform bytes (the reference takes u32 words mod 256,
lcu:933-934, which is uniform too); a ``random.Random`` can be injected.

Host conversions are batched numpy work (bit packing, range checks, limb
codecs). Added (no reference counterpart): the little-endian bytes forms
``encrypt_bytes``, ``encrypt_cts_bytes``, ``decrypt_cts_bytes`` and
``encode_hamming_server_bytes`` (minimal LE bytes per ciphertext, the
convention of hamming/paillier.py:101, 162), which never build Python
integers; every method that takes ciphertexts also takes LE bytes
(big_from_py, lcu:1301-1320, takes int, decimal and 0x-hex strings).
"""

from __future__ import annotations

import json
import os
import random
from typing import Optional, Sequence

import numpy as np

from ..engine import Context, LookupKey, Modulus
from ..limbs import (DoesNotFit, bytes_to_limbs, ints_to_limbs, limbs_below, limbs_to_bytes, limbs_to_ints,
                     width_for_bits)
from . import keygen as kg
from ._gpu_ext import KEY_BITS, pointwise_mulmod_limbs
from .packing import (chunk_geometry, embeddings_to_plaintexts, odd_mask, plaintexts_below, plaintexts_to_digits,
                      validate_embeddings)

ALPHA_LEN = kg.DEFAULT_ALPHA_LEN

try:  # the reference returns gmpy2.mpz
    import gmpy2 as _gmpy2  # type: ignore

    def _out(v: int):
        return _gmpy2.mpz(v)
except ImportError:  # pragma: no cover - gmpy2 is absent in this image
    def _out(v: int):
        return int(v)


def _to_int(v) -> int:
    """big_from_py (lcu:1301-1320): int, decimal string, 0x-hex string or LE bytes."""
    if isinstance(v, (bytes, bytearray)):
        return int.from_bytes(v, "little")
    if isinstance(v, str):
        t = v.strip()
        return int(t, 16) if t[:2].lower() == "0x" else int(t, 10)
    return int(v)


def _value_limbs(values: Sequence, limbs: int) -> np.ndarray:
    """Values as LE bytes, ints, decimal or 0x-hex strings (big_from_py,
    lcu:1301-1320) -> (count, limbs) uint32; bytes and int batches in bulk."""
    if values and all(isinstance(v, (bytes, bytearray, memoryview)) for v in values):
        return bytes_to_limbs(values, limbs)
    if values and all(type(v) is int for v in values):
        return ints_to_limbs(values, limbs)
    return ints_to_limbs([_to_int(v) for v in values], limbs)


def _with_crt(key: LookupKey, phi) -> LookupKey:
    """Give `key` half-width decryption when phi = (p-1)(q-1) yields n's
    factors and the width has a CRT form (n^2 of 2048 or 4096 bits); the
    results are those of the full-width form for every ciphertext."""
    if phi is None or key.width not in (64, 128):
        return key
    f = kg.factor_from_phi(key.n, _to_int(phi))
    if f is not None and max(f).bit_length() <= 8 * key.width:  # p^2, q^2 fit half the width
        key.set_crt(*f)
    return key


def build_g_table_device(ctx: Context, g: int, n: int, message_chunks: int) -> np.ndarray:
    """g_table[i][j] = (g^(2^(8i)))^j mod n^2 as a (chunks, 256, W) limb array.

    Column j is the pointwise product of column j-1 with the row bases on the
    GPU (lookup.py:104-110; lcu:550-588).
    """
    n2 = n * n
    W = width_for_bits(n2.bit_length())
    bases = []
    b = g % n2
    for _ in range(message_chunks):
        bases.append(b)
        b = pow(b, 256, n2)
    B = ints_to_limbs(bases, W)
    table = np.zeros((message_chunks, 256, W), dtype=np.uint32)
    table[:, 0, 0] = 1
    table[:, 1, :] = B
    mod = Modulus(ctx, n2, W)
    cur = B
    for j in range(2, 256):
        cur = mod.mulmod(cur, B)
        table[:, j, :] = cur
    return table


def build_noise_table_device(ctx: Context, g: int, n: int, rng: Optional[random.Random] = None,
                             size: int = kg.NOISE_TABLE_SIZE) -> np.ndarray:
    """noise[k] = (g^n)^(r_k) mod n^2 with r_k drawn as lookup.py:198-210 (one
    batched per-item-exponent modexp on the GPU)."""
    n2 = n * n
    W = width_for_bits(n2.bit_length())
    mod = Modulus(ctx, n2, W)
    g_n = limbs_to_ints(mod.modexp([g % n2], n))[0]
    rs = [kg.lookup_random_r(n, rng) for _ in range(size)]
    return mod.modexp_each([g_n] * size, rs)


class PaillierGPULookupClient:
    """Paillier-lookup client whose batched arithmetic runs on the GPU."""

    @staticmethod
    def max_key_len() -> int:
        return KEY_BITS

    def __init__(self, embed_len: int = 512, key_len: int = KEY_BITS, alpha_len: int = ALPHA_LEN,
                 skip_key_gen: bool = False, rng: Optional[random.Random] = None):
        if key_len > KEY_BITS:
            raise RuntimeError(
                f"PaillierGPULookupClient: key_len={key_len} exceeds the supported KEY_BITS={KEY_BITS} "
                "(n^2 is limited to 4096 bits)")
        self._embed_len = int(embed_len)
        self._key_len = int(key_len)
        self._alpha_len = int(alpha_len)
        self._chunk_len, self._chunk_num = chunk_geometry(self._embed_len, self._key_len)
        self._rng = rng
        self._ctx = Context.default()
        self._have_keys = False
        self._keys: dict = {}
        self._message_chunks = 0
        self._tables: Optional[tuple[np.ndarray, np.ndarray]] = None  # plain residues (g, noise)
        self._dev: Optional[LookupKey] = None
        if not skip_key_gen:
            k = kg.lookup_key_gen(self._key_len, self._alpha_len, rng)
            self._install(k["pk"]["g"], k["pk"]["n"], k["sk"]["phi"], k["sk"]["a"], k["sk"]["g_a_inv"])
            self._message_chunks = k["message_chunks"]

    # ---- key / table state ----
    def _install(self, g: int, n: int, phi: int, a: int, g_a_inv: int) -> None:
        n2 = n * n
        if n2.bit_length() > 32 * 128:
            raise RuntimeError(
                f"load_stringified_keys: n_squared has {n2.bit_length()} bits, exceeding the capacity of 4096 "
                f"bits (KEY_BITS={KEY_BITS})")
        self._keys = {"g": g, "n": n, "n_squared": n2, "phi": phi, "a": a, "g_a_inv": g_a_inv,
                      "inv": pow(phi, -1, n)}
        self._have_keys = True
        self._tables = None
        self._dev = None

    def _require_keys(self, what: str) -> None:
        if not self._have_keys:
            raise RuntimeError(f"{what}: keys not initialized")

    def _ensure_tables(self) -> tuple[np.ndarray, np.ndarray]:
        if self._tables is None:
            k = self._keys
            mc = self._message_chunks or (self._key_len * 2 + kg.MSG_BITS - 1) // kg.MSG_BITS
            self._message_chunks = mc
            g_tab = build_g_table_device(self._ctx, k["g"], k["n"], mc)
            noise = build_noise_table_device(self._ctx, k["g"], k["n"], self._rng)
            self._tables = (g_tab, noise)
        return self._tables

    def _device_key(self) -> LookupKey:
        if self._dev is None:
            g_tab, noise = self._ensure_tables()
            k = self._keys
            self._dev = _with_crt(LookupKey(self._ctx, k["n"], k["a"], k["g_a_inv"], g_tab, noise), k["phi"])
            # packed embeddings leave the digits above a chunk's bit string zero
            # (a chunk of s <= chunk_len bits is an integer below 2^s,
            # hamming/paillier.py:86-94): the row-group tables need cover no more
            rows = -(-min(self._chunk_len, 2 * self._embed_len) // kg.MSG_BITS)
            if rows < len(g_tab):
                self._dev.set_table_rows(rows)
        return self._dev

    @property
    def embed_len(self) -> int:
        return self._embed_len

    @property
    def key_len(self) -> int:
        return self._key_len

    @property
    def chunk_len(self) -> int:
        return self._chunk_len

    @property
    def chunk_num(self) -> int:
        return self._chunk_num

    def get_pk_hex(self) -> tuple[str, str]:
        self._require_keys("get_pk_hex")
        return format(self._keys["n"], "x"), format(self._keys["n_squared"], "x")

    def get_keys_hex(self) -> tuple[str, str, str, str]:
        self._require_keys("get_keys_hex")
        k = self._keys
        return (format(k["n"], "x"), format(k["n_squared"], "x"), format(k["phi"], "x"), format(k["inv"], "x"))

    def stringify_pk(self) -> str:
        self._require_keys("stringify_pk")
        k = self._keys
        g_n = limbs_to_ints(Modulus(self._ctx, k["n_squared"]).modexp([k["g"] % k["n_squared"]], k["n"]))[0]
        return json.dumps({"g": k["g"], "n": k["n"], "n_squared": k["n_squared"], "g_n": g_n})

    def stringify_sk(self) -> str:
        self._require_keys("stringify_sk")
        k = self._keys
        return json.dumps({"phi": k["phi"], "a": k["a"], "g_a_inv": k["g_a_inv"]})

    def stringify_config(self) -> str:
        return json.dumps({"message_chunks": self._message_chunks, "alpha_len": self._alpha_len,
                           "embed_len": self._embed_len, "key_len": self._key_len})

    def load_stringified_keys(self, pk_json: str, sk_json: str) -> None:
        pk, sk = json.loads(pk_json), json.loads(sk_json)
        self._install(_to_int(str(pk["g"])), _to_int(str(pk["n"])), _to_int(str(sk["phi"])),
                      _to_int(str(sk["a"])), _to_int(str(sk["g_a_inv"])))
        self._alpha_len = self._keys["a"].bit_length()

    def load_config(self, config: dict, tables: Optional[dict] = None) -> None:
        """lcu:1771-1836: geometry from `config`; tables optional (else built lazily)."""
        if not self._have_keys:
            raise RuntimeError("load_config: keys must be loaded first")
        key_len = int(config["key_len"])
        if key_len > KEY_BITS:
            raise RuntimeError(f"load_config: key_len={key_len} exceeds KEY_BITS={KEY_BITS}")
        self._embed_len = int(config["embed_len"])
        self._key_len = key_len
        self._alpha_len = int(config["alpha_len"])
        self._chunk_len, self._chunk_num = chunk_geometry(self._embed_len, self._key_len)
        self._message_chunks = int(config["message_chunks"])
        self._tables = None
        self._dev = None
        if tables is not None:
            if not isinstance(tables, dict):
                raise RuntimeError("load_config: tables must be a dict")
            W = width_for_bits(self._keys["n_squared"].bit_length())
            gt = tables.get("g_table")
            nt = tables.get("noise_table")
            if gt is None or nt is None:
                return
            rows = [gt[i] if i in gt else gt[str(i)] for i in range(self._message_chunks)] \
                if isinstance(gt, dict) else list(gt)
            if len(rows) != self._message_chunks or any(len(r) != 256 for r in rows):
                raise RuntimeError("load_config: g_table size mismatch")
            if len(nt) != kg.NOISE_TABLE_SIZE:
                raise RuntimeError("load_config: noise_table size mismatch")
            g_tab = ints_to_limbs([_to_int(v) for r in rows for v in r], W).reshape(self._message_chunks, 256, W)
            noise = ints_to_limbs([_to_int(v) for v in nt], W)
            self._tables = (g_tab, noise)

    def load_config_bytes(self, config: dict, tables: Optional[dict] = None) -> None:
        self.load_config(config, tables)

    def dump_tables(self) -> dict:
        """{"g_table": {i: [mpz]*256}, "noise_table": [mpz]*256} (lcu:1838-1862)."""
        self._require_keys("dump_tables")
        g_tab, noise = self._ensure_tables()
        flat = limbs_to_ints(g_tab.reshape(-1, g_tab.shape[-1]))
        return {"g_table": {i: [_out(v) for v in flat[i * 256:(i + 1) * 256]] for i in range(g_tab.shape[0])},
                "noise_table": [_out(v) for v in limbs_to_ints(noise)]}

    def dump_tables_bytes(self) -> dict:
        return self.dump_tables()

    # ---- encryption ----
    def _noise_idx(self, count: int) -> np.ndarray:
        if self._rng is not None:  # seeded: one randbytes call per batch (row-major count x 14 bytes)
            return np.frombuffer(self._rng.randbytes(count * kg.NOISE_MULTIPLES), dtype=np.uint8).reshape(
                count, kg.NOISE_MULTIPLES).copy()
        return np.frombuffer(os.urandom(count * kg.NOISE_MULTIPLES), dtype=np.uint8).reshape(
            count, kg.NOISE_MULTIPLES).copy()

    def encrypt_limbs(self, m: np.ndarray, noise_idx: Optional[np.ndarray] = None) -> np.ndarray:
        """Batched table encryption of plaintext limbs (count, W) -> ciphertext limbs."""
        dev = self._device_key()
        digits = plaintexts_to_digits(m, self._message_chunks)
        idx = self._noise_idx(len(m)) if noise_idx is None else np.ascontiguousarray(noise_idx, dtype=np.uint8)
        return dev.encrypt_digits(digits, idx)

    def _width(self) -> int:
        return width_for_bits(self._keys["n_squared"].bit_length())

    def _encrypt_embeddings(self, embeddings) -> tuple[int, np.ndarray]:
        self._require_keys("encrypt")
        bits = validate_embeddings(embeddings, self._embed_len)
        W = self._width()
        if bits.shape[0] == 0:
            return 0, np.zeros((0, W), dtype=np.uint32)
        m = embeddings_to_plaintexts(bits, self._key_len, W)
        if not plaintexts_below(self._embed_len, self._key_len, self._keys["n"]) and \
                not limbs_below(m, self._keys["n"]).all():
            raise ValueError("plaintext must be in range [0, n)")
        return bits.shape[0], self.encrypt_limbs(m)

    def encrypt(self, embeddings: Sequence[Sequence[int]]) -> list:
        """[batch][embed_len] bits -> [batch][chunk_num] ciphertexts (lcu:1873-2002).
        A (batch, embed_len) numpy array of 0/1 is taken as it is."""
        batch, c = self._encrypt_embeddings(embeddings)
        cts, cn = [_out(v) for v in limbs_to_ints(c)], self._chunk_num
        return [cts[i * cn:(i + 1) * cn] for i in range(batch)]

    def encrypt_bytes(self, embeddings) -> list[list[bytes]]:
        """``encrypt`` with minimal LE bytes per ciphertext."""
        batch, c = self._encrypt_embeddings(embeddings)
        by, cn = limbs_to_bytes(c), self._chunk_num
        return [by[i * cn:(i + 1) * cn] for i in range(batch)]

    def _plaintext_batch(self, values: Sequence) -> np.ndarray:
        try:
            m = _value_limbs(values, self._width())
        except DoesNotFit as e:
            raise RuntimeError("encrypt_cts: plaintext out of range [0, n)") from e
        if not limbs_below(m, self._keys["n"]).all():
            raise RuntimeError("encrypt_cts: plaintext out of range [0, n)")
        return m

    def encrypt_cts(self, plaintexts: Sequence, noise_idx: Optional[np.ndarray] = None) -> list:
        """Integers in [0, n) -> ciphertexts (lcu:2006-2131)."""
        self._require_keys("encrypt_cts")
        if not len(plaintexts):
            return []
        return [_out(c) for c in limbs_to_ints(self.encrypt_limbs(self._plaintext_batch(plaintexts), noise_idx))]

    def encrypt_cts_bytes(self, plaintexts: Sequence[bytes], noise_idx: Optional[np.ndarray] = None) -> list[bytes]:
        """``encrypt_cts`` with LE bytes in and minimal LE bytes out."""
        self._require_keys("encrypt_cts_bytes")
        if not len(plaintexts):
            return []
        return limbs_to_bytes(self.encrypt_limbs(self._plaintext_batch(plaintexts), noise_idx))

    # ---- decryption ----
    def _cipher_batch(self, values: Sequence) -> np.ndarray:
        try:
            C = _value_limbs(values, self._width())
        except DoesNotFit as e:
            raise ValueError("ciphertext must be in range [0, n^2)") from e
        if not limbs_below(C, self._keys["n_squared"]).all():
            raise ValueError("ciphertext must be in range [0, n^2)")
        return C

    def decrypt_cts(self, ciphers: Sequence) -> list[int]:
        """m = L(c^a mod n^2) * g_a_inv mod n (lcu:2440-2499; lookup.py:257-278)."""
        self._require_keys("decrypt_cts")
        if not len(ciphers):
            return []
        return limbs_to_ints(self._decrypt_key().decrypt(self._cipher_batch(ciphers)))

    def decrypt_cts_bytes(self, ciphers: Sequence[bytes]) -> list[bytes]:
        """``decrypt_cts`` with LE bytes in and minimal LE bytes out."""
        self._require_keys("decrypt_cts_bytes")
        if not len(ciphers):
            return []
        return limbs_to_bytes(self._decrypt_key().decrypt(self._cipher_batch(ciphers)))

    def _decrypt_key(self) -> LookupKey:
        if self._dev is not None:
            return self._dev
        k = self._keys
        if getattr(self, "_dec_only", None) is None or self._dec_only.n != k["n"]:
            self._dec_only = _with_crt(LookupKey(self._ctx, k["n"], k["a"], k["g_a_inv"], [], []), k["phi"])
        return self._dec_only

    def decode_hamming_client(self, ciphers_batch: Sequence[Sequence]) -> list[int]:
        """Decrypt chunks and count odd-position bits, fused on the GPU (lcu:2338-2437)."""
        self._require_keys("decode_hamming_client")
        batch = len(ciphers_batch)
        if batch == 0:
            return []
        cn = self._chunk_num
        if any(len(row) != cn for row in ciphers_batch):
            raise RuntimeError("decode_hamming_client: each inner list must have length == chunk_num")
        C = self._cipher_batch([c for row in ciphers_batch for c in row])
        counts = self._decrypt_key().decode_hamming(C, odd_mask(self._chunk_len))
        return counts.reshape(batch, cn).sum(axis=1, dtype=np.int64).tolist()

    def decrypt_then_reencrypt(self, ciphers: Sequence, key_a: dict, key_b: dict) -> list:
        """Decrypt under key A, re-encrypt under key B's tables (lcu:2134-2285)."""
        cs = [_to_int(c) for c in ciphers]
        if not cs:
            return []
        pa, sa, pb = key_a.get("pk", key_a), key_a.get("sk", key_a), key_b.get("pk", key_b)
        for d, names in ((pa, ("g", "n", "n_squared")), (sa, ("a", "g_a_inv")), (pb, ("g", "n", "n_squared"))):
            if not all(x in d for x in names):
                raise RuntimeError("key pk must contain g, n, n_squared; sk must contain a and g_a_inv")
        na = _to_int(pa["n"])
        dec = _with_crt(LookupKey(self._ctx, na, _to_int(sa["a"]), _to_int(sa["g_a_inv"]), [], []), sa.get("phi"))
        nb, gb = _to_int(pb["n"]), _to_int(pb["g"])
        if "message_chunks" in key_b:
            mc = int(key_b["message_chunks"])
        elif "key_len" in key_b:
            mc = (int(key_b["key_len"]) * 2 + kg.MSG_BITS - 1) // kg.MSG_BITS
        else:
            mc = (nb.bit_length() + kg.MSG_BITS - 1) // kg.MSG_BITS
        Wb = width_for_bits((nb * nb).bit_length())
        if "g_table" in key_b and "noise_table" in key_b:
            gt = key_b["g_table"]
            rows = [gt[i] if i in gt else gt[str(i)] for i in range(mc)] if isinstance(gt, dict) else list(gt)
            g_tab = ints_to_limbs([_to_int(v) for r in rows for v in r], Wb).reshape(mc, 256, Wb)
            noise = ints_to_limbs([_to_int(v) for v in key_b["noise_table"]], Wb)
        else:
            g_tab = build_g_table_device(self._ctx, gb, nb, mc)
            noise = build_noise_table_device(self._ctx, gb, nb, self._rng)
        enc = LookupKey(self._ctx, nb, 1, 1, g_tab, noise)
        if 8 * mc >= na.bit_length():
            # every plaintext digit has a row: decrypt, digits and encryption on the device
            out = dec.reencrypt(cs, enc, self._noise_idx(len(cs)))
        else:
            # a plaintext may carry more digits than key B's rows: check each on the host
            ms = dec.decrypt(cs)
            digits = plaintexts_to_digits(ints_to_limbs(limbs_to_ints(ms), Wb), mc)
            out = enc.encrypt_digits(digits, self._noise_idx(len(cs)))
        return [_out(c) for c in limbs_to_ints(out)]

    # ---- evaluator ----
    @staticmethod
    def _server(ct1: Sequence, ct2: Sequence, n2: int, what: str) -> np.ndarray:
        if len(ct1) != len(ct2):
            raise RuntimeError(f"{what}: ct1 and ct2 must have the same length")
        W = width_for_bits(n2.bit_length())
        return pointwise_mulmod_limbs(_value_limbs(ct1, W), _value_limbs(ct2, W), n2)

    def encode_hamming_client(self, ct1: Sequence, ct2: Sequence) -> list:
        self._require_keys("encode_hamming_client")
        res = self._server(ct1, ct2, self._keys["n_squared"], "encode_hamming_client")
        return [_out(v) for v in limbs_to_ints(res)]

    @staticmethod
    def _server_modulus(pk: dict, what: str) -> int:
        if "n_squared" not in pk:
            raise RuntimeError(f"{what}: pk must contain n_squared")
        return _to_int(pk["n_squared"])

    @staticmethod
    def encode_hamming_server(ct1: Sequence, ct2: Sequence, pk: dict) -> list:
        """ct1[i] * ct2[i] mod pk["n_squared"] (int, decimal or 0x-hex), lcu:2312-2335."""
        C = PaillierGPULookupClient
        n2 = C._server_modulus(pk, "encode_hamming_server")
        return [_out(v) for v in limbs_to_ints(C._server(ct1, ct2, n2, "encode_hamming_server"))]

    @staticmethod
    def encode_hamming_server_bytes(ct1: Sequence[bytes], ct2: Sequence[bytes], pk: dict) -> list[bytes]:
        """``encode_hamming_server`` with LE bytes in and minimal LE bytes out."""
        C = PaillierGPULookupClient
        n2 = C._server_modulus(pk, "encode_hamming_server_bytes")
        return limbs_to_bytes(C._server(ct1, ct2, n2, "encode_hamming_server_bytes"))


# Backwards-compatible alias (lcu:2669-2670, 2720)
PaillierGPUClient = PaillierGPULookupClient
