"""Rank ranges (opts rank_begin / rank_end) and the reference's labeling
count.

The reference enumerates total_num_to_calc ranks (L_OpenMP.c:352-367):
(unsigned long long) pow(d, k-1) for d >= 3, which differs from d^(k-1)
from 2^53 on.  The engine enumerates [0, min(count, d^(k-1))); a count
above d^(k-1) only revisits labelings (their digits below k-1 repeat), which
never win the reference's strict comparison.

CPU tests: the engine's default end of the rank range, found through the
range checks, which run before any device is touched.  GPU tests: the
maximum and first maximising rank of ranges on every kernel family against
the oracle's segment walk (lo_walk, the reference's per-worker loop), which
covers the split of a range into whole units and the shorter aligned units
at its ends."""
import ctypes

import numpy as np
import pytest

import oracle

# (d, rows): shapes where (unsigned long long) pow(d, rows - 1) differs from
# d^(rows - 1) (both signs), and two where it is exact
POW_SHAPES = [(3, 35), (3, 36), (3, 37), (3, 41), (5, 24), (5, 28), (7, 23), (9, 20), (3, 20), (10, 20)]


def _count(d, R):
    return min(oracle.total(R, d), d ** (R - 1))


def _small(R, C, seed):
    return np.random.default_rng(seed).integers(-9, 10, size=(R, C)).astype(np.int32)


@pytest.mark.parametrize("d,R", POW_SHAPES)
def test_default_end_is_the_reference_count(pkg, d, R):
    count = _count(d, R)
    if (d, R) in ((3, 35), (5, 28)):
        assert oracle.total(R, d) > d ** (R - 1)        # pow rounds up: every labeling
    if (d, R) in ((3, 36), (3, 41), (9, 20)):
        assert oracle.total(R, d) < d ** (R - 1)        # pow rounds down: the top ranks are left out
    A = _small(R, 5, d * 100 + R)
    # a range starting at the default end is empty ...
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(A, d, 0, rank_begin=count)
    assert e.value.code == -1 and "[%d, %d)" % (count, count) in str(e.value)
    # ... and one rank before it is not (no GPU here: it fails at the device,
    # after the range checks; with a GPU it enumerates that one rank)
    try:
        pkg.calc_lnorm(A, d, 0, rank_begin=count - 1)
    except pkg.LnormError as err:
        assert err.code == -4, str(err)


def test_range_errors(pkg):
    A = _small(6, 4, 1)
    for kw in ({"rank_begin": 5, "rank_end": 5}, {"rank_begin": 7, "rank_end": 3}, {"rank_end": 2 ** 5 + 1},
               {"rank_begin": 2 ** 5}):
        with pytest.raises(pkg.LnormError) as e:
            pkg.calc_lnorm(A, 1, 0, **kw)
        assert e.value.code == -1, kw
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(_small(6, 4, 2), 3, 0, rank_end=3 ** 5 + 1)
    assert e.value.code == -1
    # 4^32 = 2^64 labelings: the reference's count wraps to 0
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(_small(33, 3, 3), 4, 0)
    assert e.value.code == -3
    # batches enumerate whole problems
    b = pkg._Batch([(A, 1, 0)], True)
    o = pkg.default_opts(rank_begin=1)
    assert pkg.hip_lib().lnorm_hip_calc_batch_ex(b.ptr, b.count, ctypes.byref(o), None) == -1
    assert "rank_begin" in pkg.last_error()


def _agree(pkg, A, n, m, rb, re, **kw):
    norm, strat, res = pkg.calc_lnorm(A, n, m, rank_begin=rb, rank_end=re, **kw)
    best, rank, ostrat = oracle.walk(A, n, m, rb, re - 1)
    assert (norm, res["best_rank"], strat) == (best, rank, ostrat), (A.shape, n, m, rb, re, kw, res["kernel_name"])
    assert res["total"] == re - rb
    return res


# (rows, columns, d, marginal, entry range, kernel)
KERNEL_CASES = [
    (22, 40, 1, 0, 100, "blocked"),
    (22, 30, 1, 1, 100, "blocked"),
    (22, 100, 1, 0, 3000, "blocked"),      # column sums beyond 2^15: 32-bit words over W waves
    (22, 40, 1, 0, 100, "table16"),
    (21, 33, 1, 1, 100, "table16"),
    (18, 600, 1, 0, 20, "table16"),        # the W-wave tabulated walk
    (22, 40, 1, 0, 3000, "table32"),
    (18, 600, 1, 1, 3000, "table32"),      # 32-bit words over W waves, LM
    (11, 200, 3, 0, 5000, "table32"),      # d-ary, 32-bit words over W waves
    (10, 300, 3, 0, 5000, "columns"),
    (22, 40, 1, 0, 100, "plain"),
    (20, 200, 1, 0, 50, "wide"),
    (16, 30, 2, 0, 50, "table16"),
    (16, 30, 2, 0, 50, "plain"),
    (14, 32, 3, 0, 50, "table16"),
    (11, 200, 3, 0, 20, "table16"),        # W waves
    (14, 32, 3, 0, 5000, "table32"),
    (14, 32, 3, 0, 50, "plain16"),
    (14, 32, 3, 0, 50, "plain"),
    (11, 300, 3, 0, 30, "wide"),
    (9, 20, 5, 0, 50, "table16"),
    (8, 12, 7, 0, 50, "plain16"),
    (6, 20, 10, 0, 50, "general"),
    (7, 9, 10, 0, 50, "auto"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("R,C,n,m,r,kernel", KERNEL_CASES)
def test_ranges_against_segment_walk(pkg, R, C, n, m, r, kernel):
    rng = np.random.default_rng(R * 1000 + C * 10 + n)
    A = rng.integers(-r, r + 1, size=(R, C)).astype(np.int32)
    full = (2 if n <= 2 else n) ** (R - 1)
    print("ranges", R, C, n, m, kernel, flush=True)
    whole = pkg.calc_lnorm(A, n, m, kernel=kernel)
    ranges = [(0, min(full, 1 << 18)), (full - min(full, 200000), full), (full // 3, full // 3 + 1),
              (full // 2 - 12345, full // 2 + 54321)]
    for _ in range(4):
        size = int(rng.integers(1, min(full, 300000) + 1))
        lo = int(rng.integers(0, full - size + 1))
        ranges.append((lo, lo + size))
    for rb, re in ranges:
        rb, re = max(0, rb), min(full, re)
        res = _agree(pkg, A, n, m, rb, re, kernel=kernel)
        assert res["kernel_name"].split("<")[0] == whole[2]["kernel_name"].split("<")[0]
    # the whole range as two halves: the larger half's maximum is the norm
    cut = int(rng.integers(1, full))
    lo = pkg.calc_lnorm(A, n, m, kernel=kernel, rank_end=cut)
    hi = pkg.calc_lnorm(A, n, m, kernel=kernel, rank_begin=cut)
    first = lo if lo[0] >= hi[0] else hi
    assert (max(lo[0], hi[0]), first[1], first[2]["best_rank"]) == (whole[0], whole[1], whole[2]["best_rank"])


@pytest.mark.gpu
@pytest.mark.parametrize("d,R", [(3, 35), (3, 36), (3, 41), (5, 28), (9, 20)])
def test_reference_count_beyond_2_53(pkg, d, R):
    """The last ranks of the reference's count for d^(k-1) >= 2^53, where
    pow() rounds: the range up to the default end is the reference's, and
    its maximum and first rank are the segment walk's (also at 2^60 and
    beyond, in the 64-bit rank arithmetic of every walk)."""
    count = _count(d, R)
    A = np.random.default_rng(d * 7 + R).integers(-40, 41, size=(R, 6)).astype(np.int32)
    print("count", d, R, count, flush=True)
    for rb in (count - 5000, count - 1):
        norm, strat, res = pkg.calc_lnorm(A, d, 0, rank_begin=rb)
        assert res["total"] == count - rb
        assert (norm, res["best_rank"], strat) == oracle.walk(A, d, 0, rb, count - 1)
    mid = d ** (R - 1) // 2 + 98765
    _agree(pkg, A, d, 0, mid, mid + 150000)


@pytest.mark.gpu
def test_ranges_int64(pkg):
    """The long long variant over rank ranges, with entries beyond int32."""
    rng = np.random.default_rng(64)
    for R, C, n in ((20, 12, 1), (16, 9, 2), (12, 7, 3), (8, 5, 6)):
        A = rng.integers(-100, 101, size=(R, C)).astype(np.int64) * (2 ** 33 + 5)
        full = (2 if n <= 2 else n) ** (R - 1)
        for rb, re in ((0, full), (full // 5 + 3, full // 5 + 70001), (full - 4321, full)):
            re = min(re, full)
            norm, strat, res = pkg.calc_lnorm64(A, n, 0, rank_begin=rb, rank_end=re)
            assert (norm, res["best_rank"], strat) == oracle.walk64(A, n, 0, rb, re - 1), (R, C, n, rb, re)
            assert res["total"] == re - rb
