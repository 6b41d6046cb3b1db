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


def validate_embeddings(embeddings, embed_len: int) -> np.ndarray:
    """(batch, embed_len) uint8 of 0/1, raising RuntimeError like the reference
    GPU client (cu:885-896) on a wrong length or a non-binary value."""
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


def embeddings_to_plaintexts(bits: np.ndarray, key_len: int, limbs: int) -> np.ndarray:
    """(batch, embed_len) 0/1 -> (batch*chunk_num, limbs) uint32 plaintexts."""
    batch, embed_len = bits.shape
    chunk_len, chunk_num = chunk_geometry(embed_len, key_len)
    padded = np.zeros((batch, 2 * embed_len), dtype=np.uint8)
    padded[:, 1::2] = bits
    out = np.zeros((batch, chunk_num, limbs * 32), dtype=np.uint8)
    for c in range(chunk_num):
        lo = c * chunk_len
        hi = min(lo + chunk_len, 2 * embed_len)
        if hi <= lo:
            continue
        avail = hi - lo
        if avail > limbs * 32:
            raise ValueError("chunk wider than the plaintext width")
        out[:, c, :avail] = padded[:, lo:hi][:, ::-1]  # LSB-first
    packed = np.packbits(out.reshape(batch * chunk_num, limbs * 32), axis=1, bitorder="little")
    return np.ascontiguousarray(packed.view("<u4").astype(np.uint32))


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
