"""Conversions between Python integers and little-endian 32-bit limb arrays.

A batch is a numpy ``uint32`` array of shape ``(count, limbs)`` -- the "AoS"
layout of the reference's ``cgbn_mem_t`` arrays (paillier_GPU_client.cu:310-313)
and of the C ABI's host-buffer entry points. Values that do not fit raise
ValueError instead of being truncated (the reference's ``hex_to_le_words``
silently truncates, cu:235-254).
"""

from __future__ import annotations

from typing import Iterable, Sequence

import numpy as np

_ZERO = b"\0"


class DoesNotFit(ValueError):
    """A value is negative or wider than the limb width."""


def width_for_bits(bits: int) -> int:
    """Smallest built width (32, 64 or 128 limbs) that holds `bits` bits."""
    for w in (32, 64, 128):
        if bits <= 32 * w:
            return w
    raise ValueError(f"{bits}-bit values exceed the widest built width (4096 bits)")


def ints_to_limbs(values: Iterable[int], limbs: int) -> np.ndarray:
    vals = list(values)
    nbytes = 4 * limbs
    try:
        raw = b"".join(int(v).to_bytes(nbytes, "little") for v in vals)
    except OverflowError as e:
        raise DoesNotFit(f"value does not fit in {limbs} limbs (or is negative)") from e
    return np.frombuffer(raw, dtype="<u4").reshape(len(vals), limbs).astype(np.uint32)  # a writable copy


def int_to_limbs(value: int, limbs: int) -> np.ndarray:
    return ints_to_limbs([value], limbs)[0].copy()


def limbs_to_ints(arr: np.ndarray) -> list[int]:
    a = np.ascontiguousarray(arr, dtype=np.uint32)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    raw = a.astype("<u4", copy=False).tobytes()
    step = 4 * a.shape[1]
    return [int.from_bytes(raw[i:i + step], "little") for i in range(0, len(raw), step)]


def limbs_to_int(arr: Sequence[int]) -> int:
    return limbs_to_ints(np.asarray(arr, dtype=np.uint32))[0]


def limb_count(value: int) -> int:
    return max(1, (value.bit_length() + 31) // 32)


def limbs_below(arr: np.ndarray, bound: int) -> np.ndarray:
    """Row-wise arr < bound as a bool array (no conversion to Python ints):
    the most significant limb where a row differs from the bound decides."""
    a = np.ascontiguousarray(arr, dtype=np.uint32)
    count, W = a.shape
    if bound.bit_length() > 32 * W:
        return np.ones(count, dtype=bool)
    b = int_to_limbs(bound, W)
    ne = a != b
    top = W - 1 - np.argmax(ne[:, ::-1], axis=1)
    return ne.any(axis=1) & (a[np.arange(count), top] < b[top])


def _as_bytes(v) -> bytes:
    if isinstance(v, bytes):
        return v
    if isinstance(v, (bytearray, memoryview)):
        return bytes(v)
    raise ValueError(f"expected a bytes-like value, got {type(v).__name__}")


def bytes_to_limbs(values: Sequence, limbs: int) -> np.ndarray:
    """Little-endian byte strings (any length, trailing zero bytes allowed)
    -> (count, limbs) uint32, each as ``int.from_bytes(v, "little")`` (the
    wire convention of hamming/paillier.py:101, 162); a value wider than
    the width raises ValueError."""
    nbytes = 4 * limbs
    vals = [_as_bytes(v) for v in values]
    if any(len(v) > nbytes for v in vals):
        vals = [v.rstrip(_ZERO) for v in vals]
        if any(len(v) > nbytes for v in vals):
            raise DoesNotFit(f"value does not fit in {limbs} limbs")
    raw = b"".join(v.ljust(nbytes, _ZERO) for v in vals)
    return np.frombuffer(raw, dtype="<u4").reshape(len(vals), limbs).astype(np.uint32)


def limbs_to_bytes(arr: np.ndarray) -> list[bytes]:
    """(count, limbs) -> minimal little-endian byte strings, b"\\x00" for zero
    (``int(v).to_bytes(max(1, (v.bit_length() + 7) // 8), "little")``)."""
    a = np.ascontiguousarray(arr, dtype=np.uint32)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    raw = a.astype("<u4", copy=False).tobytes()
    step = 4 * a.shape[1]
    return [raw[i:i + step].rstrip(_ZERO) or _ZERO for i in range(0, len(raw), step)]


def limbs_to_hex(arr: np.ndarray) -> list[str]:
    """(count, limbs) -> lower-case hex without leading zeros ("0" for zero),
    as ``format(v, "x")``."""
    a = np.ascontiguousarray(arr, dtype=np.uint32)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    raw = np.ascontiguousarray(a[:, ::-1]).astype(">u4").tobytes()  # big-endian rows
    step = 4 * a.shape[1]
    return [raw[i:i + step].hex().lstrip("0") or "0" for i in range(0, len(raw), step)]


def hex_to_limbs(values: Sequence[str], limbs: int) -> np.ndarray:
    """Hex strings (optional 0x, surrounding blanks) -> (count, limbs) uint32;
    a value wider than the width raises ValueError."""
    nhex = 8 * limbs
    parts = []
    for v in values:
        t = v.strip()
        if t[:2] in ("0x", "0X"):
            t = t[2:]
        t = t.lstrip("0")
        if len(t) > nhex:
            raise DoesNotFit(f"value does not fit in {limbs} limbs")
        parts.append(t.rjust(nhex, "0"))
    try:
        raw = bytes.fromhex("".join(parts))
    except ValueError as e:
        raise ValueError("invalid hex digit") from e
    be = np.frombuffer(raw, dtype=">u4").reshape(len(parts), limbs)
    return np.ascontiguousarray(be[:, ::-1], dtype=np.uint32)
