"""FNV-1a 64 over the little-endian bytes of numeric arrays (reference: engine/include/sinksim/hash.hpp:15-56).

Doubles are hashed as their 8 bytes with negative zero normalised to zero, 32-bit integers as 4 bytes, bytes as
themselves; one hasher can be fed several arrays in turn (the chained "all" hash of the compiled files).
"""
from __future__ import annotations

import numpy as np

_OFFSET = 14695981039346656037
_PRIME = 1099511628211
_MASK = (1 << 64) - 1


def _canonical_bytes(a: np.ndarray) -> bytes:
    a = np.asarray(a)
    if a.dtype == np.float64:
        a = np.where(a == 0.0, 0.0, a).astype("<f8")
    elif a.dtype == np.int32:
        a = a.astype("<i4")
    elif a.dtype == np.uint8:
        pass
    else:
        raise TypeError(f"FNV hashing supports f64, i32 and u8 arrays, not {a.dtype}")
    return np.ascontiguousarray(a).tobytes()


class Fnv1a64:
    __slots__ = ("h",)

    def __init__(self) -> None:
        self.h = _OFFSET

    def update(self, data: bytes) -> None:
        h = self.h
        for b in data:
            h = ((h ^ b) * _PRIME) & _MASK
        self.h = h

    def update_array(self, a: np.ndarray) -> None:
        self.update(_canonical_bytes(a))

    def value(self) -> int:
        return self.h

    def hex(self) -> str:
        return f"{self.h:016x}"
