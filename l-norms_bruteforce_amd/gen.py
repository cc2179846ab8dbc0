"""Seeded synthetic integer matrices in the reference's input format.

Replaces gen_mtx.m (gen_mtx.m:4-5: n x n uniform integers in [-r, r],
space-separated, one row per line) with a reproducible generator:
SplitMix64 as the bit source and Lemire's unbiased bounded-integer method.
"""
from __future__ import annotations

import numpy as np

_GAMMA = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Outputs offset+1 .. offset+count of SplitMix64 started at `seed`."""
    idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        state = np.uint64(seed & _M64) + idx * np.uint64(_GAMMA)
        return _mix(state)


def _mulhi_lo(x: np.ndarray, s: int):
    """128-bit product x*s split into (hi, lo) 64-bit words, s < 2^32."""
    s64 = np.uint64(s)
    xl = x & np.uint64(0xFFFFFFFF)
    xh = x >> np.uint64(32)
    with np.errstate(over="ignore"):
        pl = xl * s64
        ph = xh * s64
        lo = pl + (ph << np.uint64(32))
        carry = (lo < pl).astype(np.uint64)
        hi = (ph >> np.uint64(32)) + carry
    return hi, lo


def uniform_ints(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    """`count` integers uniform on [lo, hi] (Lemire rejection, exact)."""
    s = hi - lo + 1
    if not 1 <= s < (1 << 32):
        raise ValueError("range must hold 1 .. 2^32-1 values")
    x = splitmix64(seed, count)
    h, l = _mulhi_lo(x, s)
    thresh = ((1 << 64) - s) % s
    bad = np.nonzero(l < np.uint64(thresh))[0]
    out = h.astype(np.int64)
    nxt = count
    for i in bad:                       # probability s/2^64 per draw
        while True:
            y = int(splitmix64(seed, 1, nxt)[0])
            nxt += 1
            m = y * s
            if (m & _M64) >= thresh:
                out[i] = m >> 64
                break
    return out + lo


def gen_matrix(rows: int, cols: int, r: int, seed: int) -> np.ndarray:
    """rows x cols int32 matrix, entries uniform on [-r, r]."""
    return uniform_ints(seed, rows * cols, -r, r).reshape(rows, cols).astype(np.int32)


def matrix_text(M) -> str:
    """The reference input format: space-separated, one row per line."""
    return "".join(" ".join(str(int(v)) for v in row) + "\n" for row in np.asarray(M))


def write_matrix(path, M) -> None:
    with open(path, "w") as f:
        f.write(matrix_text(M))
