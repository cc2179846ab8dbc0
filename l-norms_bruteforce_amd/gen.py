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


def _draw(seed: int, k: int, lo: int, hi: int) -> int:
    """The k-th parameter of an item: one draw from sub-stream (seed, k)."""
    return int(uniform_ints(((seed << 8) | k) & _M64, 1, lo, hi)[0])


def batch_item(i: int, seed0: int = 1_000_000, d3_max_rows: int = 20, r: int = 100) -> dict:
    """Item i of the mixed batch workload: an m x n file (m rows uniform in
    [12, 30], n columns uniform in [8, 64]), d = 1, 2, 3 with probability
    60/25/15 %, entries uniform in [-r, r].  A quarter of the items get 1-3
    duplicated rows and 1-3 duplicated columns and the preprocessing flag 'y';
    a tenth of the d = 1 items use marginals ('m').  d = 3 items keep at most
    d3_max_rows rows (3^19 = 1.2e9 labelings at the default 20), which bounds
    the batch at about 1e12 labelings in total.

    Returns {"seed", "text" (file bytes), "d", "stat", "marginal"}: the
    arguments the reference CLI takes after the file name."""
    seed = seed0 + i
    cols = _draw(seed, 0, 8, 64)
    rows = _draw(seed, 1, 12, 30)
    u = _draw(seed, 2, 0, 99)
    d = 1 if u < 60 else (2 if u < 85 else 3)
    if d == 3:
        rows = min(rows, d3_max_rows)
    dup = _draw(seed, 3, 0, 3) == 0
    marginal = 1 if d == 1 and _draw(seed, 4, 0, 9) == 0 else 0
    M = gen_matrix(rows, cols, r, seed)
    if dup:
        nr, nc = _draw(seed, 5, 1, 3), _draw(seed, 6, 1, 3)
        src_r = uniform_ints(((seed << 8) | 7) & _M64, nr, 1, rows - 1)
        src_c = uniform_ints(((seed << 8) | 8) & _M64, nc, 1, cols - 1)
        M = np.vstack([M, M[src_r]])
        M = np.hstack([M, M[:, src_c]])
    return {"seed": seed, "text": matrix_text(M).encode(), "d": d, "stat": "y" if dup else "n",
            "marginal": marginal}


def batch_workload(count: int = 4096, seed0: int = 1_000_000, d3_max_rows: int = 20, r: int = 100) -> list:
    """Items 0..count-1 of the mixed batch workload (see batch_item)."""
    return [batch_item(i, seed0, d3_max_rows, r) for i in range(count)]
