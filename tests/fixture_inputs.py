"""Deterministic synthetic inputs for fixtures and tests.

The generator is splitmix64 over (seed, index) in numpy uint64 arithmetic, so
inputs are reproducible on any machine without storing them; only integer
operations and exact float scalings are used.
"""
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n):
    with np.errstate(over="ignore"):
        z = (np.arange(n, dtype=np.uint64) + np.uint64(seed) * np.uint64(1000003)) * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def u8_image(rows, cols, seed, lo=0, hi=256):
    """Integer-valued filterbank in [lo, hi) (uint8 when it fits)."""
    z = splitmix64(seed, rows * cols)
    v = (z % np.uint64(hi - lo)).astype(np.int64) + lo
    return v.reshape(rows, cols)


def noise_image(rows, cols, seed):
    """Approximately N(0,1) fp32 noise: sum of 4 uniform 16-bit variates,
    centred and scaled by exact powers of two (no transcendental calls)."""
    z = splitmix64(seed, rows * cols)
    s = np.zeros(rows * cols, np.int64)
    for k in range(4):
        s += ((z >> np.uint64(16 * k)) & np.uint64(0xFFFF)).astype(np.int64)
    # sum of 4 U{0..65535}: mean 131070, sd ~37837; scale by 2^-15 ~ sd 1.15
    v = (s - 131070).astype(np.float64) / 32768.0
    return v.astype(np.float32).reshape(rows, cols)
