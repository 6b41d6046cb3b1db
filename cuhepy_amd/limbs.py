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
        raise ValueError(f"value does not fit in {limbs} limbs (or is negative)") from e
    arr = np.frombuffer(raw, dtype="<u4").reshape(len(vals), limbs)
    return np.ascontiguousarray(arr, dtype=np.uint32)


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
