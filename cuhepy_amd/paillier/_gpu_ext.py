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
Added (no reference counterpart): batched ``encrypt_cts``/``decrypt_cts``/
``decrypt_cts_crt`` over integers, so ``Paillier.encrypt``/``decrypt`` can run
on the GPU.
"""

from __future__ import annotations

import json
import random
from typing import Optional, Sequence

import numpy as np

from ..engine import Context, Modulus, PaillierKey
from ..limbs import ints_to_limbs, limbs_to_ints, width_for_bits
from . import keygen as kg
from .packing import chunk_geometry, embeddings_to_plaintexts, odd_mask, validate_embeddings

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


def pointwise_mulmod(ct1: Sequence[int], ct2: Sequence[int], n_squared: int) -> list[int]:
    """c1[i] * c2[i] mod n^2 on the GPU (gpu_pointwise_mulmod_hex, cu:672-732)."""
    if len(ct1) != len(ct2):
        raise RuntimeError("encode_hamming_*: ct1 and ct2 must have the same length")
    if not ct1:
        return []
    mod = _modulus(n_squared)
    out = mod.mulmod(ints_to_limbs(ct1, mod.limbs), ints_to_limbs(ct2, mod.limbs))
    return limbs_to_ints(out)


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
    def _draw_r(self, count: int) -> list[int]:
        return [kg.random_r(self._n, self._rng) for _ in range(count)]

    def encrypt_cts(self, plaintexts: Sequence[int], rs: Optional[Sequence[int]] = None) -> list[int]:
        """c_i = g^m_i r_i^n mod n^2 for m_i in [0, n) (scheme.py:63-83)."""
        key = self._require_keys("encrypt_cts")
        ms = [int(m) for m in plaintexts]
        for m in ms:
            if not 0 <= m < self._n:
                raise ValueError("plaintext must be in range [0, n)")
        rr = self._draw_r(len(ms)) if rs is None else [int(r) for r in rs]
        if len(rr) != len(ms):
            raise ValueError("rs must match plaintexts")
        if not ms:
            return []
        out = key.encrypt(ints_to_limbs(ms, key.width), ints_to_limbs(rr, key.width))
        return limbs_to_ints(out)

    def decrypt_cts(self, ciphers: Sequence[int]) -> list[int]:
        """m_i = L(c_i^phi mod n^2) * inv mod n (scheme.py:86-104)."""
        key = self._require_keys("decrypt_cts")
        cs = [int(c) for c in ciphers]
        n2 = self._n * self._n
        for c in cs:
            if not 0 <= c < n2:
                raise ValueError("ciphertext must be in range [0, n^2)")
        if not cs:
            return []
        return limbs_to_ints(key.decrypt(ints_to_limbs(cs, key.width)))

    def decrypt_cts_crt(self, ciphers: Sequence[int]) -> list[int]:
        """CRT decryption (two half-width modexps per ciphertext)."""
        key = self._require_keys("decrypt_cts_crt")
        cs = [int(c) for c in ciphers]
        n2 = self._n * self._n
        for c in cs:
            if not 0 <= c < n2:
                raise ValueError("ciphertext must be in range [0, n^2)")
        if not cs:
            return []
        return limbs_to_ints(key.decrypt_crt(ints_to_limbs(cs, key.width)))

    # ---- reference surface ----
    def encrypt(self, embeddings: Sequence[Sequence[int]]) -> list[list[str]]:
        """[batch][embed_len] bits -> [batch][chunk_num] hex ciphertexts (cu:872-974)."""
        key = self._require_keys("encrypt")
        bits = validate_embeddings(embeddings, self._embed_len)
        batch = bits.shape[0]
        if batch == 0:
            return []
        m = embeddings_to_plaintexts(bits, self._key_len, key.width)
        ms = limbs_to_ints(m)
        for v in ms:
            if v >= self._n:
                raise ValueError("plaintext must be in range [0, n)")
        r = ints_to_limbs(self._draw_r(len(ms)), key.width)
        cts = limbs_to_ints(key.encrypt(m, r))
        cn = self._chunk_num
        return [[format(c, "x") for c in cts[i * cn:(i + 1) * cn]] for i in range(batch)]

    def encode_hamming_client(self, ct1: Sequence[str], ct2: Sequence[str]) -> list[str]:
        """ct1[i] * ct2[i] mod n^2 with the client's own key (cu:976-983)."""
        self._require_keys("encode_hamming_client")
        out = pointwise_mulmod([_parse_hex(c) for c in ct1], [_parse_hex(c) for c in ct2], self._n * self._n)
        return [format(v, "x") for v in out]

    @staticmethod
    def encode_hamming_server(ct1: Sequence[str], ct2: Sequence[str], pk: dict) -> list[str]:
        """Evaluator: ct1[i] * ct2[i] mod pk["n_squared"] (hex), cu:986-992."""
        n2 = _parse_hex(pk["n_squared"])
        out = pointwise_mulmod([_parse_hex(c) for c in ct1], [_parse_hex(c) for c in ct2], n2)
        return [format(v, "x") for v in out]

    def decode_hamming_client(self, ciphers_hex_batch: Sequence[Sequence[str]]) -> list[int]:
        """Decrypt each chunk and count odd-position bits (cu:995-1038), fused on the GPU."""
        key = self._require_keys("decode_hamming_client")
        batch = len(ciphers_hex_batch)
        if batch == 0:
            return []
        cn = self._chunk_num
        flat: list[int] = []
        for row in ciphers_hex_batch:
            if len(row) != cn:
                raise RuntimeError("decode_hamming_client: each inner list must have length == chunk_num")
            flat.extend(_parse_hex(c) for c in row)
        n2 = self._n * self._n
        for c in flat:
            if not 0 <= c < n2:
                raise ValueError("ciphertext must be in range [0, n^2)")
        counts = key.decode_hamming(ints_to_limbs(flat, key.width), odd_mask(self._chunk_len))
        return [int(counts[i * cn:(i + 1) * cn].sum()) for i in range(batch)]


# the reference module also exports the compile-time width
PaillierGPUClient.KEY_BITS = KEY_BITS
