"""``PaillierGPUClient`` -- drop-in for ``cuhepy.paillier._gpu_ext``.

Same constructor, methods, properties and return types as the pybind11 class
of ``src/cuhepy/paillier/_gpu_ext/paillier_GPU_client.cu:735-1165``; the work
runs in the gfx950 kernels of ``libchx.so`` through the C ABI (include/chx.h).

Differences from the reference extension, each taking the CPU scheme's
semantics (paillier/scheme.py):
  * a plaintext chunk >= n raises ValueError (the extension reduces it mod n,
    cu:518-520);
  * r is drawn on the host as scheme.py:47-60 does (uniform in [1, n-1],
    gcd(r, n) = 1), not as (urandom mod (n-1)) + 1 with a fallback to r = 1
    (cu:376-404); a ``random.Random`` can be injected for reproducibility;
  * hex inputs wider than the key width raise instead of being truncated
    (cu:235-254).
Added (no reference counterpart):
  * batched ``encrypt_cts``/``decrypt_cts``/``decrypt_cts_crt`` over integers,
    so ``Paillier.encrypt``/``decrypt`` (scheme.py:63-104) can run on the GPU;
  * the same path in the little-endian bytes convention of
    hamming/paillier.py:101, 162 and of the msgpack packets of
    benchmarks/bfv_client_matrix.py:49-69 (minimal LE bytes per ciphertext):
    ``encrypt_bytes``, ``encode_hamming_server_bytes``, ``encrypt_cts_bytes``,
    ``decrypt_cts_bytes``; ``decode_hamming_client`` takes bytes too. These
    never build Python integers: bytes <-> limb arrays in numpy.
Host conversions are batched numpy work (bit packing, range checks, hex and
bytes codecs); without an injected ``random.Random`` the r values are drawn
in one batch from the OS and their coprimality with n is checked with one
gcd of their product mod n (a tree of device mulmods), redrawing one at a
time only if that gcd is not 1.
"""

from __future__ import annotations

import json
import math
import random
from typing import Optional, Sequence

import numpy as np

from ..engine import Context, Modulus, PaillierKey
from ..limbs import (DoesNotFit, bytes_to_limbs, hex_to_limbs, ints_to_limbs, limbs_below, limbs_to_bytes,
                     limbs_to_hex, limbs_to_ints, width_for_bits)
from . import keygen as kg
from .packing import chunk_geometry, embeddings_to_plaintexts, odd_mask, plaintexts_below, validate_embeddings

KEY_BITS = 1024  # largest prime size: n^2 of 4096 bits = 128 limbs

_MOD_CACHE: dict[int, Modulus] = {}


def _modulus(value: int) -> Modulus:
    m = _MOD_CACHE.get(value)
    if m is None:
        if len(_MOD_CACHE) > 16:
            _MOD_CACHE.clear()
        m = _MOD_CACHE[value] = Modulus(Context.default(), value)
    return m


def _parse_hex(s) -> int:
    if isinstance(s, (bytes, bytearray)):
        return int.from_bytes(s, "little")
    if isinstance(s, int):
        return s
    t = str(s).strip()
    if t[:2] in ("0x", "0X"):
        t = t[2:]
    try:
        return int(t, 16) if t else 0
    except ValueError as e:
        raise ValueError(f"Invalid hex digit in {s!r}") from e


def cipher_limbs(values: Sequence, limbs: int) -> np.ndarray:
    """Ciphertexts as hex strings (the reference's format), LE bytes or ints
    -> (count, limbs) uint32; hex and bytes batches are converted in bulk."""
    if values and all(type(v) is str for v in values):
        return hex_to_limbs(values, limbs)
    if values and all(isinstance(v, (bytes, bytearray, memoryview)) for v in values):
        return bytes_to_limbs(values, limbs)
    return ints_to_limbs([_parse_hex(v) for v in values], limbs)


def pointwise_mulmod_limbs(a: np.ndarray, b: np.ndarray, n_squared: int) -> np.ndarray:
    """Row-wise a * b mod n^2 on the GPU (gpu_pointwise_mulmod_hex, cu:672-732)."""
    if len(a) != len(b):
        raise RuntimeError("encode_hamming_*: ct1 and ct2 must have the same length")
    mod = _modulus(n_squared)
    if len(a) == 0:
        return np.zeros((0, mod.limbs), dtype=np.uint32)
    return mod.mulmod(a, b)


def pointwise_mulmod(ct1: Sequence[int], ct2: Sequence[int], n_squared: int) -> list[int]:
    """c1[i] * c2[i] mod n^2 on the GPU (gpu_pointwise_mulmod_hex, cu:672-732)."""
    if len(ct1) != len(ct2):
        raise RuntimeError("encode_hamming_*: ct1 and ct2 must have the same length")
    if not ct1:
        return []
    W = _modulus(n_squared).limbs
    return limbs_to_ints(pointwise_mulmod_limbs(ints_to_limbs(ct1, W), ints_to_limbs(ct2, W), n_squared))


def _server_operands(ct1: Sequence, ct2: Sequence, n2: int) -> tuple[np.ndarray, np.ndarray]:
    if len(ct1) != len(ct2):
        raise RuntimeError("encode_hamming_*: ct1 and ct2 must have the same length")
    W = width_for_bits(n2.bit_length())
    return cipher_limbs(ct1, W), cipher_limbs(ct2, W)


class PaillierGPUClient:
    """Paillier (g = n + 1) client whose batched arithmetic runs on the GPU."""

    @staticmethod
    def max_key_len() -> int:
        return KEY_BITS

    def __init__(self, embed_len: int = 512, key_len: int = KEY_BITS, skip_key_gen: bool = False,
                 rng: Optional[random.Random] = None):
        if key_len > KEY_BITS:
            raise RuntimeError(
                f"PaillierGPUClient: key_len={key_len} exceeds the supported KEY_BITS={KEY_BITS} "
                "(n^2 is limited to 4096 bits)")
        self._embed_len = int(embed_len)
        self._key_len = int(key_len)
        self._chunk_len, self._chunk_num = chunk_geometry(self._embed_len, self._key_len)
        self._rng = rng
        self._ctx = Context.default()
        self._key: Optional[PaillierKey] = None
        self._n = self._phi = self._inv = 0
        self._have_keys = False
        if not skip_key_gen:
            k = kg.paillier_key_gen(self._key_len, rng)
            self._set_keys(k["pk"]["n"], k["sk"]["phi"], k["sk"]["inv"], k["p"], k["q"])

    # ---- key state ----
    def _set_keys(self, n: int, phi: int, inv: int, p: Optional[int] = None, q: Optional[int] = None) -> None:
        if p is None or q is None:
            f = kg.factor_from_phi(n, phi)
            if f is not None:
                p, q = f
        if (n * n).bit_length() > 32 * 128:
            raise RuntimeError(
                f"load_stringified_keys: n_squared has {(n * n).bit_length()} bits, exceeding the capacity "
                f"of 4096 bits (KEY_BITS={KEY_BITS})")
        self._n, self._phi, self._inv = n, phi, inv
        self._key = PaillierKey(self._ctx, n, phi, inv, p, q)
        self._have_keys = True

    def _require_keys(self, what: str) -> PaillierKey:
        if not self._have_keys or self._key is None:
            raise RuntimeError(f"{what}: keys not initialized")
        return self._key

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
        if not self._have_keys:
            raise RuntimeError("get_pk_hex: keys not initialized")
        return format(self._n, "x"), format(self._n * self._n, "x")

    def get_keys_hex(self) -> tuple[str, str, str, str]:
        if not self._have_keys:
            raise RuntimeError("get_keys_hex: keys not initialized")
        return (format(self._n, "x"), format(self._n * self._n, "x"), format(self._phi, "x"),
                format(self._inv, "x"))

    def stringify_pk(self) -> str:
        if not self._have_keys:
            raise RuntimeError("stringify_pk: keys not initialized")
        n = self._n
        return json.dumps({"g": str(n + 1), "n": str(n), "n_squared": str(n * n)})

    def stringify_sk(self) -> str:
        if not self._have_keys:
            raise RuntimeError("stringify_sk: keys not initialized")
        return json.dumps({"phi": str(self._phi), "inv": str(self._inv)})

    def stringify_config(self) -> str:
        return json.dumps({"embed_len": self._embed_len, "key_len": self._key_len})

    def load_stringified_keys(self, pk_json: str, sk_json: str) -> None:
        pk, sk = json.loads(pk_json), json.loads(sk_json)
        self._set_keys(int(str(pk["n"])), int(str(sk["phi"])), int(str(sk["inv"])))

    def load_config(self, config: dict) -> None:
        key_len = int(config["key_len"])
        if key_len > KEY_BITS:
            raise RuntimeError(f"load_config: key_len={key_len} exceeds KEY_BITS={KEY_BITS}")
        self._embed_len = int(config["embed_len"])
        self._key_len = key_len
        self._chunk_len, self._chunk_num = chunk_geometry(self._embed_len, self._key_len)

    # ---- batched integer API (added) ----
    def _n_modulus(self) -> Modulus:
        return _modulus(self._n)

    def _draw_r(self, count: int, width: int) -> np.ndarray:
        """(count, width) r values uniform in [1, n-1] with gcd(r, n) = 1
        (scheme.py:47-60)."""
        if self._rng is not None:  # reproducible: one value at a time from the injected generator
            return ints_to_limbs([kg.random_r(self._n, self._rng) for _ in range(count)], width)
        R = kg.random_below_limbs(self._n, count, width)
        if count and math.gcd(self._n_modulus().product(R[:, :self._n_modulus().limbs]), self._n) != 1:
            # some r shares a factor with n: redraw those, one at a time
            for i, r in enumerate(limbs_to_ints(R)):
                if math.gcd(r, self._n) != 1:
                    R[i] = ints_to_limbs([kg.random_r(self._n)], width)[0]
        return R

    def _plaintext_limbs(self, m: np.ndarray) -> np.ndarray:
        if not limbs_below(m, self._n).all():
            raise ValueError("plaintext must be in range [0, n)")
        return m

    def _cipher_batch(self, values: Sequence, width: int) -> np.ndarray:
        try:
            C = cipher_limbs(values, width)
        except DoesNotFit as e:
            raise ValueError("ciphertext must be in range [0, n^2)") from e
        if not limbs_below(C, self._n * self._n).all():
            raise ValueError("ciphertext must be in range [0, n^2)")
        return C

    def _encrypt_limbs(self, key: PaillierKey, m: np.ndarray, rs: Optional[Sequence[int]] = None) -> np.ndarray:
        self._plaintext_limbs(m)
        if rs is None:
            r = self._draw_r(len(m), key.width)
        else:
            if len(rs) != len(m):
                raise ValueError("rs must match plaintexts")
            r = ints_to_limbs([int(v) for v in rs], key.width)
        return key.encrypt(m, r)

    def encrypt_cts(self, plaintexts: Sequence[int], rs: Optional[Sequence[int]] = None) -> list[int]:
        """c_i = g^m_i r_i^n mod n^2 for m_i in [0, n) (scheme.py:63-83)."""
        key = self._require_keys("encrypt_cts")
        ms = [int(m) for m in plaintexts]
        if any(m < 0 for m in ms):
            raise ValueError("plaintext must be in range [0, n)")
        if not ms:
            if rs:
                raise ValueError("rs must match plaintexts")
            return []
        try:
            m = ints_to_limbs(ms, key.width)
        except DoesNotFit as e:
            raise ValueError("plaintext must be in range [0, n)") from e
        return limbs_to_ints(self._encrypt_limbs(key, m, rs))

    def encrypt_cts_bytes(self, plaintexts: Sequence[bytes]) -> list[bytes]:
        """``encrypt_cts`` with LE bytes in and minimal LE bytes out."""
        key = self._require_keys("encrypt_cts_bytes")
        if not plaintexts:
            return []
        try:
            m = bytes_to_limbs(plaintexts, key.width)
        except DoesNotFit as e:
            raise ValueError("plaintext must be in range [0, n)") from e
        return limbs_to_bytes(self._encrypt_limbs(key, m))

    def decrypt_cts(self, ciphers: Sequence[int]) -> list[int]:
        """m_i = L(c_i^phi mod n^2) * inv mod n (scheme.py:86-104)."""
        key = self._require_keys("decrypt_cts")
        cs = [int(c) for c in ciphers]
        if not cs:
            return []
        if any(c < 0 for c in cs):
            raise ValueError("ciphertext must be in range [0, n^2)")
        return limbs_to_ints(key.decrypt(self._cipher_batch(cs, key.width)))

    def decrypt_cts_bytes(self, ciphers: Sequence[bytes]) -> list[bytes]:
        """``decrypt_cts`` with LE bytes in and minimal LE bytes out."""
        key = self._require_keys("decrypt_cts_bytes")
        if not ciphers:
            return []
        return limbs_to_bytes(key.decrypt(self._cipher_batch(ciphers, key.width)))

    def decrypt_cts_crt(self, ciphers: Sequence[int]) -> list[int]:
        """CRT decryption (two half-width modexps per ciphertext)."""
        key = self._require_keys("decrypt_cts_crt")
        cs = [int(c) for c in ciphers]
        if not cs:
            return []
        if any(c < 0 for c in cs):
            raise ValueError("ciphertext must be in range [0, n^2)")
        return limbs_to_ints(key.decrypt_crt(self._cipher_batch(cs, key.width)))

    # ---- reference surface ----
    def _encrypt_embeddings(self, embeddings) -> tuple[int, np.ndarray]:
        key = self._require_keys("encrypt")
        bits = validate_embeddings(embeddings, self._embed_len)
        if bits.shape[0] == 0:
            return 0, np.zeros((0, key.width), dtype=np.uint32)
        m = embeddings_to_plaintexts(bits, self._key_len, key.width)
        if not plaintexts_below(self._embed_len, self._key_len, self._n):
            self._plaintext_limbs(m)
        r = self._draw_r(len(m), key.width)
        return bits.shape[0], key.encrypt(m, r)

    def encrypt(self, embeddings: Sequence[Sequence[int]]) -> list[list[str]]:
        """[batch][embed_len] bits -> [batch][chunk_num] hex ciphertexts (cu:872-974).
        A (batch, embed_len) numpy array of 0/1 is taken as it is."""
        batch, c = self._encrypt_embeddings(embeddings)
        hx, cn = limbs_to_hex(c), self._chunk_num
        return [hx[i * cn:(i + 1) * cn] for i in range(batch)]

    def encrypt_bytes(self, embeddings) -> list[list[bytes]]:
        """``encrypt`` with minimal LE bytes per ciphertext."""
        batch, c = self._encrypt_embeddings(embeddings)
        by, cn = limbs_to_bytes(c), self._chunk_num
        return [by[i * cn:(i + 1) * cn] for i in range(batch)]

    def encode_hamming_client(self, ct1: Sequence[str], ct2: Sequence[str]) -> list[str]:
        """ct1[i] * ct2[i] mod n^2 with the client's own key (cu:976-983)."""
        self._require_keys("encode_hamming_client")
        n2 = self._n * self._n
        return limbs_to_hex(pointwise_mulmod_limbs(*_server_operands(ct1, ct2, n2), n2))

    @staticmethod
    def encode_hamming_server(ct1: Sequence[str], ct2: Sequence[str], pk: dict) -> list[str]:
        """Evaluator: ct1[i] * ct2[i] mod pk["n_squared"] (hex), cu:986-992."""
        n2 = _parse_hex(pk["n_squared"])
        return limbs_to_hex(pointwise_mulmod_limbs(*_server_operands(ct1, ct2, n2), n2))

    @staticmethod
    def encode_hamming_server_bytes(ct1: Sequence[bytes], ct2: Sequence[bytes], pk: dict) -> list[bytes]:
        """``encode_hamming_server`` with LE bytes in and out."""
        n2 = _parse_hex(pk["n_squared"])
        return limbs_to_bytes(pointwise_mulmod_limbs(*_server_operands(ct1, ct2, n2), n2))

    def decode_hamming_client(self, ciphers_hex_batch: Sequence[Sequence[str]]) -> list[int]:
        """Decrypt each chunk and count odd-position bits (cu:995-1038), fused on
        the GPU. Ciphertexts are hex strings (the reference's format), LE bytes
        or ints."""
        key = self._require_keys("decode_hamming_client")
        batch = len(ciphers_hex_batch)
        if batch == 0:
            return []
        cn = self._chunk_num
        if any(len(row) != cn for row in ciphers_hex_batch):
            raise RuntimeError("decode_hamming_client: each inner list must have length == chunk_num")
        flat = [c for row in ciphers_hex_batch for c in row]
        counts = key.decode_hamming(self._cipher_batch(flat, key.width), odd_mask(self._chunk_len))
        return counts.reshape(batch, cn).sum(axis=1, dtype=np.int64).tolist()


# the reference module also exports the compile-time width
PaillierGPUClient.KEY_BITS = KEY_BITS
