"""Seeded synthetic inputs shared by the tests and the benchmark.

splitmix64 (Steele, Lea, Flood 2014): state += 0x9E3779B97F4A7C15, then the
output mix (z ^ z >> 30) * 0xBF58476D1CE4E5B9, (z ^ z >> 27) * 0x94D049BB133111EB,
z ^ z >> 31. Uniform M31 values are next() mod p.
"""

from __future__ import annotations

import numpy as np

P = (1 << 31) - 1
GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

# batched wide commitment: 4 columns of 2^14, 10 of 2^18, 20 of 2^20 (ascending)
WIDE_LOG_SIZES = [14] * 4 + [18] * 10 + [20] * 20
WIDE_SEED = 3


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Outputs start .. start + n - 1 of the stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(start + 1, start + n + 1, dtype=np.uint64) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def m31_columns(seed: int, log_sizes) -> list:
    """Columns drawn consecutively from one stream, values next() mod p."""
    out, pos = [], 0
    for l in log_sizes:
        n = 1 << l
        out.append((splitmix64(seed, n, pos) % np.uint64(P)).astype(np.uint32))
        pos += n
    return out
