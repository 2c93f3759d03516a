"""Counter-based synthetic inputs shared by the tests, the oracle checks and bench.py.

Element i of stream `seed` is splitmix64(seed, i): z = seed + (i + 1) * 0x9E3779B97F4A7C15, then
the splitmix64 finaliser (Steele, Lea & Flood 2014). Being counter-based, any slice of a stream
can be generated independently and identically on any machine. Base seed 0x4E554D5253 ("NUMRS")
plus a per-workload index.
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 0x4E554D5253
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_CHUNK = 1 << 22


def _mix(z: np.ndarray) -> np.ndarray:
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def u64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n raw 64-bit outputs of stream `seed`, starting at counter `offset`."""
    out = np.empty(n, np.uint64)
    s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        for lo in range(0, n, _CHUNK):
            hi = min(n, lo + _CHUNK)
            idx = np.arange(offset + lo + 1, offset + hi + 1, dtype=np.uint64)
            z = s + idx * _GOLDEN
            out[lo:hi] = _mix(z)
    return out


def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0, offset: int = 0) -> np.ndarray:
    """f64 in [lo, hi) with 53 random mantissa bits."""
    u = (u64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return lo + (hi - lo) * u


def normal(seed: int, n: int, dtype=np.float64, offset: int = 0) -> np.ndarray:
    """N(0,1) by Box-Muller on pairs of uniforms (cos branch only)."""
    out = np.empty(n, dtype)
    for lo in range(0, n, _CHUNK):
        hi = min(n, lo + _CHUNK)
        r = u64(seed, 2 * (hi - lo), 2 * (offset + lo))
        u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))  # (0, 1]
        u2 = (r[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        out[lo:hi] = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).astype(dtype)
    return out


def int_uniform(seed: int, n: int, lo: int, hi: int, offset: int = 0) -> np.ndarray:
    """int64 uniform in [lo, hi] inclusive (multiply-shift range reduction, bias < 2^-40)."""
    span = np.uint64(hi - lo + 1)
    out = np.empty(n, np.int64)
    for a in range(0, n, _CHUNK):
        b = min(n, a + _CHUNK)
        r = u64(seed, b - a, offset + a) >> np.uint64(32)
        out[a:b] = (r * span >> np.uint64(32)).astype(np.int64) + lo
    return out


def scaled_pow2(seed: int, n: int, emin: int = -10, emax: int = 10, dtype=np.float64) -> np.ndarray:
    """sign * 2^e * u with e uniform-integer in [emin, emax] and u uniform in [1, 2)."""
    r = u64(seed, n)
    sign = np.where((r & np.uint64(1)) == 0, 1.0, -1.0)
    e = ((r >> np.uint64(1)) % np.uint64(emax - emin + 1)).astype(np.int64) + emin
    u = 1.0 + (u64(seed ^ 0x5A5A5A5A, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return (sign * np.ldexp(u, e)).astype(dtype)
