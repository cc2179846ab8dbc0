"""The tabulated walks in 32-bit words (LNORM_KERNEL_TABLE32: one column per
word, biased by 2^31, v_sad_u32) against the oracle: bit-exact norm,
strategy and maximising Gray rank at every instantiated width and wave
count, at forced segment lengths and at the row counts where the tables
start to apply; AUTO's choice of them for int32 inputs whose column sums
exceed 16 bits; entries next to the int32 contract's bound (sum |A| <=
2^31 - 1); and equality with the 32-bit blocked and plain walks on larger
problems.  GPU only: every call goes through the C ABI."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _rand(rng, R, C, r):
    return rng.integers(-r, r + 1, size=(R, C)).astype(np.int32)


def _agree(eng, A, n, marg, **kw):
    want = oracle.calc_lnorm(A, n, marg, 8)
    norm, strat, res = eng.calc_lnorm(A, n, marg, **kw)
    assert (norm, strat, res["best_rank"]) == want, (A.shape, n, marg, kw, res["kernel_name"])
    return res


def _m_of(res):
    return int(res["kernel_name"].split("M=")[1].split(",")[0])


def _big_sums(A, first_col=0):
    return int(np.abs(A[:, first_col:].astype(np.int64)).sum(axis=0).max()) > 32767


# (d, rows, single-wave column counts, W-wave column counts).  Single wave:
# every instantiated width (8, 16, 24, 32, 48, 64, 96 words and the widest,
# 40 at d = 5) and counts between them.  W waves of 32 words (d = 8, 9: 24;
# d = 10: 16): one count per wave count (2, 3, 4, 5, 6, 8 waves for d <= 6,
# not 5 at d = 6; up to 4 waves beyond).
TABLE32 = [
    (2, 14, [1, 5, 8, 9, 17, 24, 30, 33, 48, 50, 64, 80, 96], [100, 150, 180, 250]),
    (3, 10, [1, 5, 8, 9, 17, 24, 30, 33, 48, 50, 64], [65, 100, 150, 170, 256]),
    (4, 8, [3, 8, 16, 24, 31, 40, 48], [49, 80, 128, 150, 190, 200]),
    (5, 7, [7, 16, 24, 32, 33, 40], [41, 90, 110, 140, 180, 256]),
    (6, 7, [5, 8, 9, 16, 24, 32], [33, 70, 120, 170, 220]),
    (7, 6, [5, 16, 24, 32], [33, 80, 128]),
    (8, 6, [5, 16, 24], [25, 60, 96]),
    (9, 5, [5, 9, 24], [25, 72, 90]),
    (10, 5, [5, 16], [17, 40, 64]),
]


@pytest.mark.parametrize("d,R,cols,wcols", TABLE32)
def test_table32_walk_vs_oracle(pkg, d, R, cols, wcols):
    rng = np.random.default_rng(1300 + d)
    for C in cols + wcols:
        # int16-range entries: column sums beyond 32767, where AUTO takes it
        A = _rand(rng, R, C, 20000)
        print("table32", d, A.shape, flush=True)
        res = _agree(pkg, A, d, 0)
        want = "table32w<" if C in wcols else "table32<"
        if _big_sums(A):
            assert res["kernel_name"].startswith(want), res["kernel_name"]
        # small entries: the same walk when asked for
        B = _rand(rng, R, C, 60)
        res = _agree(pkg, B, d, 0, kernel="table32")
        assert res["kernel_name"].startswith(want), res["kernel_name"]
    for C in (cols[len(cols) // 2], wcols[0], wcols[-1]):
        A = _rand(rng, R, C, 20000)
        M = _m_of(pkg.calc_lnorm(A, d, 0, kernel="table32")[2])
        for q in (M, M + 1, M + 2):
            if q > R - 1:
                continue
            print("table32 seg_log", d, A.shape, q, flush=True)
            res = _agree(pkg, A, d, 0, seg_log=q, kernel="table32")
            assert f",q={q}>" in res["kernel_name"], res["kernel_name"]
        # exactly M enumerated rows: one block per segment; fewer: refused
        A = _rand(rng, M + 1, C, 20000)
        assert _agree(pkg, A, d, 0, kernel="table32")["kernel_name"].startswith("table32")
        with pytest.raises(pkg.LnormError) as e:
            pkg.calc_lnorm(_rand(rng, M, C, 20000), d, 0, kernel="table32")
        assert e.value.code == -5


# (rows, |.| column counts) of the 32-bit L1 / LM walk: one wave at every
# instantiated width (8 .. 48 in steps of 8, 64, 80, 96, 128, 160, 192
# words) and between them; W waves at 256, 320, 384, 512 and 1024 words
TABLE32_BIN = [1, 7, 8, 9, 16, 17, 24, 33, 40, 41, 48, 60, 64, 65, 80, 96, 100, 128, 129, 160, 192]
TABLE32_BIN_W = [193, 256, 300, 320, 384, 500, 512, 513, 1000, 1024]


@pytest.mark.parametrize("marg", [0, 1])
def test_table32_binary_vs_oracle(pkg, marg):
    rng = np.random.default_rng(1400 + marg)
    for R in (7, 12):
        for nabs in TABLE32_BIN + TABLE32_BIN_W:
            C = nabs + marg
            A = _rand(rng, R, C, 20000)
            print("table32 L1", marg, A.shape, flush=True)
            res = _agree(pkg, A, 1, marg)
            want = "table32w<L" if nabs > 192 else "table32<L"
            if _big_sums(A, marg):
                assert res["kernel_name"].startswith(want), res["kernel_name"]
            res = _agree(pkg, _rand(rng, R, C, 60), 1, marg, kernel="table32")
            assert res["kernel_name"].startswith(want), res["kernel_name"]
    for nabs in (40, 128, 160, 300, 1000):
        A = _rand(rng, 15, nabs + marg, 20000)
        M = _m_of(pkg.calc_lnorm(A, 1, marg, kernel="table32")[2])
        for q in (M, M + 1, M + 3, 14):
            print("table32 L1 seg_log", marg, A.shape, q, flush=True)
            res = _agree(pkg, A, 1, marg, kernel="table32", seg_log=q)
            assert f"M={M},q={q}>" in res["kernel_name"], res["kernel_name"]
        assert _agree(pkg, _rand(rng, M + 1, nabs + marg, 20000), 1, marg)["kernel_name"].startswith("table32")
        with pytest.raises(pkg.LnormError) as e:
            pkg.calc_lnorm(_rand(rng, M, nabs + marg, 20000), 1, marg, kernel="table32")
        assert e.value.code == -5


def _near_bound(rng, R, C):
    """Entries whose absolute sum is INT32_MAX - 7 (the int32 contract's
    edge): every partial sum and the norm stay within int32."""
    A = rng.integers(-(2 ** 31 // (R * C)), 2 ** 31 // (R * C), size=(R, C)).astype(np.int64)
    A[A == 0] = 1
    rest = (2 ** 31 - 1 - 7) - int(np.abs(A).sum()) + abs(int(A[-1, -1]))
    assert 0 < rest < 2 ** 31
    A[-1, -1] = rest if A[-1, -1] > 0 else -rest
    assert int(np.abs(A).sum()) == 2 ** 31 - 1 - 7
    return A.astype(np.int32)


@pytest.mark.parametrize("n,R,C", [(1, 14, 100), (1, 10, 600), (3, 9, 40), (3, 8, 150), (5, 6, 30), (10, 5, 12)])
def test_table32_near_int32_bound(pkg, n, R, C):
    rng = np.random.default_rng(1500 + n + C)
    A = _near_bound(rng, R, C)
    print("table32 bound", n, A.shape, flush=True)
    res = _agree(pkg, A, n, 0)
    assert res["kernel_name"].startswith("table32"), res["kernel_name"]
    if n == 1:
        _agree(pkg, A, n, 1, kernel="table32")


def test_table32_binary_equals_blocked32_large(pkg):
    """2^29 sign vectors of int16 entries: the 32-bit tabulated walk equals
    the 32-bit blocked kernel (norm, strategy, rank) at 128 columns and the
    binary W-wave walk at 1000, and the strategy's objective is the norm."""
    from lnorm_amd import gen
    for (R, C, marg, other) in ((30, 128, 0, "blocked"), (30, 101, 1, "blocked"), (26, 1000, 0, "wide")):
        A = gen.gen_matrix(C, R, 30000, 105 + C).T.copy()
        print("table32 vs", other, R, C, marg, flush=True)
        t = pkg.calc_lnorm(A, 1, marg)
        o = pkg.calc_lnorm(A, 1, marg, kernel=other)
        assert t[2]["kernel_name"].startswith("table32"), t[2]["kernel_name"]
        assert t[:2] == o[:2] and t[2]["best_rank"] == o[2]["best_rank"], (R, C, marg)
        assert oracle.objective(A, 1, marg, t[1]) == t[0]


def test_table32_equals_plain_larger(pkg):
    """d-ary problems of 3^15 .. 8^7 labelings with int16-range entries: the
    32-bit tabulated walk (one wave and W waves) equals the 32-bit plain or
    general walk, and the strategy's objective is the norm."""
    from lnorm_amd import gen
    for (R, C, d, other) in ((16, 40, 3, "plain"), (14, 200, 3, "general"), (11, 30, 4, "plain"),
                             (8, 16, 8, "plain"), (9, 100, 6, "general")):
        A = gen.gen_matrix(R, C, 20000, 70 + d + C)
        print("table32 vs", other, R, C, d, flush=True)
        t = pkg.calc_lnorm(A, d, 0)
        o = pkg.calc_lnorm(A, d, 0, kernel=other)
        assert t[2]["kernel_name"].startswith("table32"), t[2]["kernel_name"]
        assert t[:2] == o[:2] and t[2]["best_rank"] == o[2]["best_rank"], (R, C, d)
        assert oracle.objective(A, d, 0, t[1]) == t[0]
