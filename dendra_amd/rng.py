"""Counter-based random numbers shared by the host generators, the HIP
kernels and the oracle.

``uniform(seed, stream, i)`` is a pure function of its integer arguments
(splitmix64 finaliser applied twice), so any consumer -- numpy here, the
HIP Poisson sources, the C oracle -- reproduces the same stream without
sharing generator state.  The reference draws from torch Generators
(dendra/models/rng.py:205-206), whose CPU and GPU streams differ; a
counter RNG makes both sides of every parity test see identical inputs.
"""

from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z + _GOLD) & _M64
        z = ((z ^ (z >> np.uint64(30))) * _C1) & _M64
        z = ((z ^ (z >> np.uint64(27))) * _C2) & _M64
        return z ^ (z >> np.uint64(31))


def bits(seed: int, stream: int, idx) -> np.ndarray:
    """64 random bits for counters ``idx`` of (seed, stream)."""
    key = _mix(np.asarray([(seed & 0xFFFFFFFF) | ((stream & 0xFFFFFFFF) << 32)],
                          dtype=np.uint64))[0]
    i = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix((key + i) & _M64)


def uniform(seed: int, stream: int, idx) -> np.ndarray:
    """Uniform doubles in [0, 1) (53-bit)."""
    return (bits(seed, stream, idx) >> np.uint64(11)).astype(np.float64) * 2.0**-53


def uniform_n(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    return uniform(seed, stream, np.arange(start, start + n, dtype=np.uint64))


def normal_n(seed: int, stream: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller over counters 2k, 2k+1."""
    u1 = uniform(seed, stream, np.arange(0, 2 * n, 2, dtype=np.uint64))
    u2 = uniform(seed, stream, np.arange(1, 2 * n, 2, dtype=np.uint64))
    u1 = np.maximum(u1, 2.0**-53)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def integers_n(seed: int, stream: int, n: int, low: int, high: int) -> np.ndarray:
    """Integers uniform on [low, high] inclusive."""
    u = uniform_n(seed, stream, n)
    return (low + np.floor(u * (high - low + 1))).astype(np.int64)
