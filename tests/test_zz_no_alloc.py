"""lnorm_hip_calc_noalloc (every device buffer carved from the library's
static per-device scratch array) against lnorm_hip_calc_ex on the same
problems: bit-identical norm, strategy, maximising rank and kernel.  Each
case runs two calls in a row, the second on a smaller input, so contents
the first call left in the shared array cannot leak into the second.
GPU only: every call goes through the C ABI."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _rand(seed, R, C, r):
    return np.random.default_rng(seed).integers(-r, r + 1, size=(R, C)).astype(np.int32)


def _same(pkg, A, n, marg, **kw):
    want = pkg.calc_lnorm(A, n, marg, **kw)
    got = pkg.calc_lnorm(A, n, marg, noalloc=True, **kw)
    assert got[0] == want[0] and got[1] == want[1], (A.shape, n, marg, kw)
    for k in ("best_rank", "total", "kernel", "kernel_name"):
        assert got[2][k] == want[2][k], (A.shape, n, marg, kw, k, got[2][k], want[2][k])
    return got


# (first problem, smaller second problem): (R, C, range, n, marginal, kernel options)
PAIRS = [
    # blocked16 (AUTO from 2^33 sign vectors), then the tabulated L1 walk
    ((36, 40, 100, 1, 0, {}), (20, 20, 10, 1, 0, {})),
    # LM through the blocked kernel, then a plain LM walk
    ((35, 30, 100, 1, 1, {"kernel": "blocked"}), (16, 18, 50, 1, 1, {"kernel": "plain"})),
    # 32-bit blocked kernel (int16-range entries), then a small one
    ((30, 100, 30000, 1, 0, {"kernel": "blocked"}), (16, 40, 30000, 1, 0, {"kernel": "blocked"})),
    # L3 tabulated walk, then L2
    ((18, 32, 50, 3, 0, {}), (12, 20, 50, 2, 0, {})),
    # more units than one launch of the array holds (2^25 segments of 2^5)
    ((31, 24, 100, 1, 0, {"kernel": "plain", "seg_log": 5}), (25, 16, 100, 1, 0, {"kernel": "plain", "seg_log": 4})),
    # the general kernel at the reference's widest matrices: its waves are
    # limited by the state slices the array holds
    ((8, 4096, 20, 10, 0, {}), (6, 3000, 20, 4, 0, {})),
    ((12, 2100, 100, 1, 0, {}), (10, 2049, 100, 1, 1, {})),
    # W-wave tabulated walks, then a narrower one
    ((16, 300, 30, 3, 0, {}), (14, 150, 30, 4, 0, {})),
    ((22, 1000, 10, 1, 0, {}), (20, 500, 10, 1, 0, {})),
]


@pytest.mark.parametrize("first,second", PAIRS, ids=[str(i) for i in range(len(PAIRS))])
def test_noalloc_matches_calc_ex(pkg, first, second):
    for seed, (R, C, r, n, marg, kw) in enumerate((first, second)):
        print("noalloc", R, C, r, n, marg, kw, flush=True)
        A = _rand(1000 + 7 * R + C + seed, R, C, r)
        _same(pkg, A, n, marg, **kw)


def test_noalloc_small_against_oracle(pkg):
    # two calls in a row, the second smaller, checked against the oracle too
    for R, C, n in ((18, 24, 3), (10, 12, 1)):
        print("noalloc oracle", R, C, n, flush=True)
        A = _rand(77 + R, R, C, 40)
        norm, strat, res = pkg.calc_lnorm(A, n, 0, noalloc=True)
        assert (norm, strat, res["best_rank"]) == oracle.calc_lnorm(A, n, 0, 8), (R, C, n)


def test_noalloc_errors(pkg):
    # argument checks are those of lnorm_hip_calc_ex
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(np.ones((5, 5), np.int32), 11, 0, noalloc=True)
    assert e.value.code == -3
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(np.ones((5, 5), np.int32), 2, 0, noalloc=True, kernel="blocked")
    assert e.value.code == -5
