"""The int_type = long long variant (lnorm_hip_calc_ex64): 64-bit matrices
whose sums exceed int32, against the oracle's 64-bit walks (norm, strategy
and maximising rank) on both kinds of plain walk (doubles below sum |A| =
2^52, 64-bit integers above it or on request), the bound between them, the
scaling identity on the GPU, delegation to the int32 kernels inside the
int32 bound, and error codes.  GPU only."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _rand64(rng, R, C, bits):
    return rng.integers(-(2 ** bits), 2 ** bits, size=(R, C)).astype(np.int64)


# (n, marginal, rows, column counts): the column counts cover the
# register-resident 64-bit widths and, beyond them, the general kernel
CASES = [
    (1, 0, 14, [1, 7, 33, 64, 70]),
    (1, 1, 13, [2, 9, 48, 65]),
    (2, 0, 12, [3, 17, 32, 40]),
    (3, 0, 8, [5, 24, 30]),
    (4, 0, 7, [4, 16, 20]),
    (5, 0, 6, [3, 8, 12]),
    (7, 0, 5, [6, 9]),
    (10, 0, 4, [8, 10]),
]


@pytest.mark.parametrize("n,marg,R,cols", CASES)
def test_int64_vs_oracle(pkg, n, marg, R, cols):
    rng = np.random.default_rng(500 + 10 * n + marg)
    for C in cols:
        for bits in (33, 40):
            A = _rand64(rng, R, C, bits)
            print("int64", n, marg, A.shape, bits, flush=True)
            want = oracle.calc_lnorm64(A, n, marg, 8)
            norm, strat, res = pkg.calc_lnorm64(A, n, marg)
            assert (norm, strat, res["best_rank"]) == want, (n, marg, A.shape, bits, res["kernel_name"])
            assert "64<" in res["kernel_name"]
            if res["kernel_name"].startswith("plainf64"):
                # the same problem on the 64-bit integer walk
                norm, strat, res = pkg.calc_lnorm64(A, n, marg, word_bits=64)
                assert (norm, strat, res["best_rank"]) == want, (n, marg, A.shape, bits, res["kernel_name"])
                assert res["kernel_name"].startswith("plain64<"), res["kernel_name"]


def test_int64_general_equals_plain(pkg):
    rng = np.random.default_rng(9)
    for n, R, C in ((1, 16, 20), (2, 12, 10), (3, 9, 8)):
        A = _rand64(rng, R, C, 36)
        p = pkg.calc_lnorm64(A, n, 0, kernel="plain")
        g = pkg.calc_lnorm64(A, n, 0, kernel="general")
        assert p[:2] == g[:2] and p[2]["best_rank"] == g[2]["best_rank"]
        assert p[2]["kernel_name"].startswith("plainf64"), p[2]["kernel_name"]
        assert g[2]["kernel_name"].startswith("general64")


def _sum_exactly(rng, R, C, total):
    """An R x C int64 matrix of random signs whose sum |A| is exactly total."""
    A = np.full((R, C), total // (R * C), np.int64)
    A[-1, -1] += total - int(A.sum())
    A *= rng.choice(np.array([-1, 1], np.int64), size=(R, C))
    assert int(np.abs(A).sum()) == total
    return A


@pytest.mark.parametrize("n,marg", [(1, 0), (1, 1), (2, 0), (3, 0), (5, 0)])
def test_int64_double_walk_bound(pkg, n, marg):
    """The plain walks run in doubles up to sum |A| = 2^52 - 1 and on 64-bit
    integers from 2^52 on; both agree with the oracle on either side, and
    with each other, including random matrices whose entries differ in
    magnitude."""
    rng = np.random.default_rng(620 + 10 * n + marg)
    for total, name in ((2 ** 52 - 1, "plainf64<"), (2 ** 52, "plain64<"), (2 ** 52 + 12345, "plain64<")):
        R, C = (11, 9) if n <= 2 else (8, 8)
        A = _sum_exactly(rng, R, C, total)
        A[rng.integers(0, R), :] //= 3   # rows of different magnitude
        A[-1, -1] += (1 if A[-1, -1] >= 0 else -1) * (total - int(np.abs(A).sum()))
        assert int(np.abs(A).sum()) == total
        print("int64 bound", n, marg, total, flush=True)
        want = oracle.calc_lnorm64(A, n, marg, 8)
        norm, strat, res = pkg.calc_lnorm64(A, n, marg)
        assert (norm, strat, res["best_rank"]) == want, (total, res["kernel_name"])
        assert res["kernel_name"].startswith(name), (total, res["kernel_name"])
    shape = {1: (26, 30), 2: (26, 30), 3: (14, 24), 5: (9, 8)}[n]
    A = rng.integers(-(2 ** 40), 2 ** 40, size=shape).astype(np.int64)   # sum |A| below 2^50
    print("int64 double vs integer", n, marg, A.shape, flush=True)
    f = pkg.calc_lnorm64(A, n, marg)
    i = pkg.calc_lnorm64(A, n, marg, word_bits=64)
    assert f[2]["kernel_name"].startswith("plainf64<") and i[2]["kernel_name"].startswith("plain64<")
    assert f[:2] == i[:2] and f[2]["best_rank"] == i[2]["best_rank"]
    assert oracle.objective64(A, n, marg, f[1]) == f[0]


def test_int64_scaling_on_gpu(pkg):
    """L(c*A) = c * L(A) with the same maximising rank: a 2^29-vector L1
    and a 3^13-labeling L3, scaled far beyond int32."""
    from lnorm_amd import gen
    for (R, C, n, c) in ((30, 20, 1, 2 ** 26 + 1), (14, 16, 3, 2 ** 31 + 5)):
        A = gen.gen_matrix(R, C, 100, 77 + n)
        base = pkg.calc_lnorm(A, n, 0)
        big = pkg.calc_lnorm64(A.astype(np.int64) * c, n, 0)
        assert big[0] == c * base[0] and big[0] > 2 ** 39
        assert big[1] == base[1] and big[2]["best_rank"] == base[2]["best_rank"]


def test_int64_inside_int32_bound_uses_int32_kernels(pkg):
    from lnorm_amd import gen
    A = gen.gen_matrix(24, 40, 100, 5)
    a = pkg.calc_lnorm(A, 1, 0)
    b = pkg.calc_lnorm64(A.astype(np.int64), 1, 0)
    assert a[:2] == b[:2] and b[2]["kernel_name"] == a[2]["kernel_name"]
    assert b[2]["kernel_name"].startswith(("blocked", "table16")) and "64<" not in b[2]["kernel_name"]


@pytest.mark.parametrize("n,marg", [(1, 0), (1, 1), (2, 0)])
def test_int64_wide_double_walk(pkg, n, marg):
    """Beyond the one-wave widths (L1 / LM 64 columns, L2 32) the walk in
    doubles runs over W waves (128, 256 or 512 columns; L2 128 or 256):
    against the oracle at every width, at segment lengths below and at one
    8-step group (q = 0, 1, 2, 3) and beyond, over a rank range, and against
    the general kernel on a larger problem."""
    rng = np.random.default_rng(660 + 10 * n + marg)
    widths = (65, 100, 128, 129, 256, 257, 400, 512) if n == 1 else (33, 100, 128, 129, 256)
    for C in widths:
        A = _rand64(rng, 9, C, 36)
        print("int64 wide", n, marg, A.shape, flush=True)
        want = oracle.calc_lnorm64(A, n, marg, 8)
        norm, strat, res = pkg.calc_lnorm64(A, n, marg)
        assert (norm, strat, res["best_rank"]) == want, (n, marg, A.shape, res["kernel_name"])
        assert res["kernel_name"].startswith("widef64<"), res["kernel_name"]
    A = _rand64(rng, 12, 200, 36)
    want = oracle.calc_lnorm64(A, n, marg, 8)
    for q in (0, 1, 2, 3, 5):
        print("int64 wide seg_log", n, marg, q, flush=True)
        norm, strat, res = pkg.calc_lnorm64(A, n, marg, seg_log=q)
        assert (norm, strat, res["best_rank"]) == want, (q, res["kernel_name"])
        assert f",q={q}>" in res["kernel_name"], res["kernel_name"]
    rb, re = 37, 2 ** 11 - 5
    print("int64 wide range", n, marg, rb, re, flush=True)
    norm, strat, res = pkg.calc_lnorm64(A, n, marg, rank_begin=rb, rank_end=re)
    assert (norm, res["best_rank"], strat) == oracle.walk64(A, n, marg, rb, re - 1)
    B = _rand64(rng, 20, 300 if n == 1 else 200, 30)
    print("int64 wide vs general", n, marg, B.shape, flush=True)
    w = pkg.calc_lnorm64(B, n, marg)
    g = pkg.calc_lnorm64(B, n, marg, kernel="general")
    assert w[2]["kernel_name"].startswith("widef64<") and g[2]["kernel_name"].startswith("general64<")
    assert w[:2] == g[:2] and w[2]["best_rank"] == g[2]["best_rank"]


# (d, rows, columns): beyond the one-wave widths, every wave count (2, 4, 8)
WIDE_DARY = [(3, 7, (25, 48, 49, 96, 97, 192)), (4, 6, (17, 32, 64, 65, 128)), (5, 5, (9, 16, 32, 64)),
             (7, 5, (9, 16, 32, 33, 64)), (10, 4, (9, 16, 32, 64))]


@pytest.mark.parametrize("n,R,cols", WIDE_DARY)
def test_int64_wide_double_dary(pkg, n, R, cols):
    """The d-ary walk in doubles over 2, 4 or 8 waves beyond one wave's
    d * NY <= 80 doubles: against the oracle at every wave count, at forced
    segment lengths (a segment shorter than one 8-step combine included),
    over a rank range, and against the general kernel."""
    rng = np.random.default_rng(690 + n)
    nw = {3: 24, 4: 16}.get(n, 8)
    seen = set()
    for C in cols:
        A = _rand64(rng, R, C, 36)
        print("int64 wide dary", n, A.shape, flush=True)
        want = oracle.calc_lnorm64(A, n, 0, 8)
        norm, strat, res = pkg.calc_lnorm64(A, n, 0)
        assert (norm, strat, res["best_rank"]) == want, (n, A.shape, res["kernel_name"])
        assert res["kernel_name"].startswith("widef64<kind=3"), res["kernel_name"]
        seen.add(int(res["kernel_name"].split("ny=")[1].split(",")[0]) // nw)
    assert seen == {2, 4, 8}, seen
    A = _rand64(rng, R + 1, cols[1], 36)
    want = oracle.calc_lnorm64(A, n, 0, 8)
    for q in (1, 2, 3):
        print("int64 wide dary seg_log", n, q, flush=True)
        norm, strat, res = pkg.calc_lnorm64(A, n, 0, seg_log=q)
        assert (norm, strat, res["best_rank"]) == want, (q, res["kernel_name"])
        assert f",q={q}>" in res["kernel_name"], res["kernel_name"]
    total = n ** R
    rb, re = 5, total - 3
    norm, strat, res = pkg.calc_lnorm64(A, n, 0, rank_begin=rb, rank_end=re)
    assert (norm, res["best_rank"], strat) == oracle.walk64(A, n, 0, rb, re - 1)
    B = _rand64(rng, R + 3, cols[2], 30)
    print("int64 wide dary vs general", n, B.shape, flush=True)
    w = pkg.calc_lnorm64(B, n, 0)
    g = pkg.calc_lnorm64(B, n, 0, kernel="general")
    assert w[2]["kernel_name"].startswith("widef64<") and g[2]["kernel_name"].startswith("general64<")
    assert w[:2] == g[:2] and w[2]["best_rank"] == g[2]["best_rank"]


def test_int64_double_walks_few_rows(pkg):
    """1, 2 and 3 rows (a single labeling, segments of 1 and 2 steps) on
    every walk in doubles, against the oracle."""
    rng = np.random.default_rng(710)
    for n, marg, C, name in ((1, 0, 40, "plainf64<"), (1, 1, 40, "plainf64<"), (1, 0, 200, "widef64<"),
                             (1, 1, 300, "widef64<"), (2, 0, 100, "widef64<"), (3, 0, 20, "plainf64<"),
                             (3, 0, 100, "widef64<"), (6, 0, 40, "widef64<")):
        for R in (1, 2, 3):
            A = _rand64(rng, R, C, 36)
            print("int64 few rows", n, marg, A.shape, flush=True)
            want = oracle.calc_lnorm64(A, n, marg, 8)
            norm, strat, res = pkg.calc_lnorm64(A, n, marg)
            assert (norm, strat, res["best_rank"]) == want, (n, marg, A.shape, res["kernel_name"])
            assert res["kernel_name"].startswith(name), (n, marg, A.shape, res["kernel_name"])
            # without the strategy (the call must still wait for its timing events)
            norm2, _, res2 = pkg.calc_lnorm64(A, n, marg, want_strategy=False)
            assert norm2 == want[0], (n, marg, A.shape)


def test_int64_errors(pkg):
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm64(np.full((3, 2), 2 ** 62, np.int64), 1, 0)
    assert e.value.code == -2
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm64(np.full((12, 3), 2 ** 40, np.int64), 1, 0, kernel="blocked")
    assert e.value.code == -5
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm64(np.full((12, 3), 2 ** 40, np.int64), 2, 1)
    assert e.value.code == -1
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm64(np.full((12, 3), 2 ** 40, np.int64), 1, 0, word_bits=16)
    assert e.value.code == -1
    # the W-wave walk runs in doubles only: refused with the integer walks
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm64(np.full((12, 100), 2 ** 33, np.int64), 1, 0, kernel="wide", word_bits=64)
    assert e.value.code == -5
