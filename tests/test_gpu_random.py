"""AUTO's routing over random shapes: several hundred seeded problems with
random rows, columns, d (1..10), marginals and entry ranges (small, around
the 16-bit column-sum bound, int16-range, and close to the int32 bound on
Sum |A|), each small enough for the oracle, through the engine's own kernel
choice and through a random rank range.  Norm, strategy and rank must equal
the oracle's, whatever kernel AUTO picks.  GPU only."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

# the most ranks a random problem may have (the oracle walks them on the CPU)
MAX_RANKS = 1 << 16


def _shape(rng):
    d = int(rng.choice([1, 1, 1, 2, 2, 3, 3, 4, 5, 6, 7, 8, 9, 10]))
    base = 2 if d <= 2 else d
    rmax = 1
    while base ** rmax <= MAX_RANKS and rmax < 63:
        rmax += 1
    R = int(rng.integers(1, rmax + 1))
    C = int(rng.choice([int(rng.integers(1, 70)), int(rng.integers(60, 300)), int(rng.integers(250, 1100))]))
    return d, R, C


def _entries(rng, R, C, kind):
    if kind == 0:
        r = 50
    elif kind == 1:    # column sums around 2^15
        r = max(1, int(2 * 32767 // max(R, 1)))
    elif kind == 2:
        r = 30000
    else:              # Sum |A| close to INT32_MAX
        r = max(1, int(2 * (2 ** 31 - 1) // (R * C) * 0.95))
    return rng.integers(-r, r + 1, size=(R, C)).astype(np.int64).clip(-(2 ** 31 - 1), 2 ** 31 - 1).astype(np.int32)


@pytest.mark.parametrize("block", range(8))
def test_auto_random_shapes_vs_oracle(pkg, block):
    rng = np.random.default_rng(9000 + block)
    kernels = {}
    for i in range(50):
        d, R, C = _shape(rng)
        marg = int(d == 1 and C > 1 and rng.integers(0, 3) == 0)
        A = _entries(rng, R, C, int(rng.integers(0, 4)))
        if np.abs(A.astype(np.int64)).sum() > 2 ** 31 - 1:
            continue
        print("random", block, i, A.shape, d, marg, flush=True)
        norm, strat, res = pkg.calc_lnorm(A, d, marg)
        want = oracle.calc_lnorm(A, d, marg, 4)
        assert (norm, strat, res["best_rank"]) == want, (block, i, A.shape, d, marg, res["kernel_name"])
        kernels[res["kernel_name"].split("<")[0]] = kernels.get(res["kernel_name"].split("<")[0], 0) + 1
        # a random shard of the same problem
        full = (2 if d <= 2 else d) ** (R - 1)
        if full > 1:
            rb = int(rng.integers(0, full))
            re = int(rng.integers(rb + 1, full + 1))
            n2, s2, r2 = pkg.calc_lnorm(A, d, marg, rank_begin=rb, rank_end=re)
            assert (n2, r2["best_rank"], s2) == oracle.walk(A, d, marg, rb, re - 1), (block, i, rb, re)
    print("kernels", kernels, flush=True)
    assert len(kernels) >= 3
