"""Bit packing between binary embeddings and Paillier plaintexts (host, numpy).

Encoding (hamming/paillier.py:81-95, encrypt_kernel cu:489-514): every bit b_j
becomes the pair "0 b_j"; the padded stream of 2*embed_len bits is cut into
chunks of chunk_len = 2*key_len bits (the last one may be shorter) and each
chunk is read MSB-first as an integer.

Decoding (hamming/paillier.py:97-110, cu:1015-1036): each plaintext is
left-padded to chunk_len bits and the '1's at odd string positions are
counted; position t (MSB-first) is LSB bit chunk_len-1-t, so the count is
popcount(m & odd_mask(chunk_len)) (make_odd_mask_host, cu:410-424), which the
GPU computes in the decrypt kernel.
"""

from __future__ import annotations

import numpy as np


def chunk_geometry(embed_len: int, key_len: int) -> tuple[int, int]:
    """(chunk_len, chunk_num) as hamming/paillier.py:26-32."""
    chunk_len = key_len * 2
    if embed_len > key_len:
        chunk_num = 2 * embed_len // chunk_len + int((2 * embed_len) % chunk_len != 0)
    else:
        chunk_num = 1
    return chunk_len, chunk_num


def _validate_generic(embeddings, embed_len: int) -> np.ndarray:
    try:
        arr = np.asarray(embeddings)
    except ValueError as e:  # ragged nested lists
        raise RuntimeError("encrypt: each embedding must have length == embed_len") from e
    if arr.size == 0:
        return np.zeros((0, embed_len), dtype=np.uint8)
    if arr.ndim != 2 or arr.shape[1] != embed_len:
        raise RuntimeError("encrypt: each embedding must have length == embed_len")
    if not np.isin(arr, (0, 1)).all():
        raise RuntimeError("encrypt: embeddings must be binary 0/1")
    return arr.astype(np.uint8)


def validate_embeddings(embeddings, embed_len: int) -> np.ndarray:
    """(batch, embed_len) uint8 of 0/1, raising RuntimeError like the reference
    GPU client (cu:885-896) on a wrong length or a non-binary value.

    A numpy array is checked in place. Nested lists go through one
    ``bytes(row)`` per row (a C loop over the row's ints, about 10 ns per bit,
    a third of what ``np.asarray`` of the nested lists costs); rows that are
    not lists of small ints take the generic path, which reports the error."""
    if isinstance(embeddings, np.ndarray):
        arr = embeddings
        if arr.size == 0:
            return np.zeros((0, embed_len), dtype=np.uint8)
        if arr.ndim != 2 or arr.shape[1] != embed_len:
            raise RuntimeError("encrypt: each embedding must have length == embed_len")
        if arr.dtype == np.bool_:
            return np.ascontiguousarray(arr).view(np.uint8)
        if arr.dtype.kind in "iu":
            if (arr.min() < 0) or (arr.max() > 1):
                raise RuntimeError("encrypt: embeddings must be binary 0/1")
            return np.ascontiguousarray(arr, dtype=np.uint8)
        return _validate_generic(arr, embed_len)
    try:
        batch = len(embeddings)
    except TypeError:
        return _validate_generic(embeddings, embed_len)
    if batch == 0:
        return np.zeros((0, embed_len), dtype=np.uint8)
    try:
        if any(len(row) != embed_len for row in embeddings):
            raise RuntimeError("encrypt: each embedding must have length == embed_len")
        raw = b"".join(map(bytes, embeddings))
    except (TypeError, ValueError):  # not sequences of ints in [0, 256)
        return _validate_generic(embeddings, embed_len)
    if len(raw) != batch * embed_len:  # a row with wider items (an int64 array, say)
        return _validate_generic(embeddings, embed_len)
    arr = np.frombuffer(raw, dtype=np.uint8).reshape(batch, embed_len)
    if arr.max() > 1:
        raise RuntimeError("encrypt: embeddings must be binary 0/1")
    return arr


# byte v -> the 16-bit value with bit k of v at bit 2k (the "0 b" pairs of 8
# bits); _SPREAD32 does the same for a 16-bit group into one 32-bit limb
_SPREAD = np.array([sum(((v >> k) & 1) << (2 * k) for k in range(8)) for v in range(256)], dtype="<u2")
_V16 = np.arange(1 << 16)
_SPREAD32 = (_SPREAD[_V16 & 255].astype(np.uint32) | (_SPREAD[_V16 >> 8].astype(np.uint32) << 16))
del _V16


def embeddings_to_plaintexts(bits: np.ndarray, key_len: int, limbs: int) -> np.ndarray:
    """(batch, embed_len) 0/1 -> (batch*chunk_num, limbs) uint32 plaintexts.

    Chunk c holds embedding bits [c key_len, c key_len + e) (e = key_len, or
    fewer for the last chunk) and, read MSB-first as "0 b" pairs, equals
    sum_i b_(c key_len + i) 4^(e - 1 - i). Left-padded with zeros to a
    multiple of 16 bits and packed MSB-first into bytes, the chunk's bytes in
    reverse order are its 16-bit groups little-endian; spreading each group's
    bits to the even bits of 32 gives the limbs."""
    batch, embed_len = bits.shape
    chunk_len, chunk_num = chunk_geometry(embed_len, key_len)
    out = np.zeros((batch, chunk_num, limbs), dtype=np.uint32)
    for c in range(chunk_num):
        j0 = c * key_len
        j1 = min(j0 + key_len, embed_len)
        if j1 <= j0:
            continue
        if 2 * (j1 - j0) > limbs * 32:
            raise ValueError("chunk wider than the plaintext width")
        part = bits[:, j0:j1]
        pad = -(j1 - j0) % 16
        if pad:
            part = np.concatenate([np.zeros((batch, pad), dtype=np.uint8), part], axis=1)
        groups = np.ascontiguousarray(np.packbits(part, axis=1)[:, ::-1]).view("<u2")
        out[:, c, :groups.shape[1]] = _SPREAD32[groups]
    return out.reshape(batch * chunk_num, limbs)


def plaintexts_below(embed_len: int, key_len: int, n: int) -> bool:
    """Every packed chunk is below n: a chunk has at most 2 min(embed_len,
    key_len) bits and n >= 2^(bits(n) - 1)."""
    return 2 * min(embed_len, key_len) <= n.bit_length() - 1


def odd_mask(chunk_len: int) -> int:
    """Bits at MSB-first odd positions t -> LSB index chunk_len-1-t."""
    mask = 0
    for t in range(1, chunk_len, 2):
        mask |= 1 << (chunk_len - 1 - t)
    return mask


def plaintexts_to_digits(m: np.ndarray, chunks: int) -> np.ndarray:
    """Base-2^8 digits of each plaintext, little-endian (lookup.py:239, lcu:889-911)."""
    a = np.ascontiguousarray(m, dtype=np.uint32)
    by = a.view(np.uint8).reshape(a.shape[0], a.shape[1] * 4)
    if by.shape[1] >= chunks:
        if by.shape[1] > chunks and by[:, chunks:].any():
            raise ValueError("plaintext has more base-256 digits than message_chunks")
        return np.ascontiguousarray(by[:, :chunks])
    out = np.zeros((a.shape[0], chunks), dtype=np.uint8)
    out[:, : by.shape[1]] = by
    return out
