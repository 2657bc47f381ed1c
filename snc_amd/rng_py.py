"""SplitMix64 draws in Python for small host-side choices (sample lengths)."""
from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1


def splitmix64_stream(seed: int, count: int):
    s = seed & _MASK
    out = []
    for _ in range(count):
        s = (s + 0x9E3779B97F4A7C15) & _MASK
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        out.append(z ^ (z >> 31))
    return out


def splitmix_uniform_ints(n: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """n integers uniform in [lo, hi] (inclusive)."""
    span = hi - lo + 1
    return np.array([lo + (v >> 11) * span // (1 << 53) for v in splitmix64_stream(seed, n)],
                    dtype=np.int32)
