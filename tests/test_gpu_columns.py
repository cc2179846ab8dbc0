"""The column walk (LNORM_KERNEL_COLUMNS: one wave per segment, the columns
over the lanes, label sums in registers, the rows in LDS) against the oracle:
every kind (L1, LM, L2, d = 3..10) at the instantiated widths (64 x columns
per lane, exactly and with padding), forced segment lengths (q = 0 .. 4),
rows whose LDS copy exceeds 64 KiB, AUTO's choice where only the general
kernel applied before, and equality with the general kernel on larger
problems.  GPU only."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CPL = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64)


def _max_cpl(d):
    return max(c for c in CPL if (c if d <= 1 else d * c) <= 128)


def _agree(pkg, A, n, m, **kw):
    norm, strat, res = pkg.calc_lnorm(A, n, m, **kw)
    want = oracle.calc_lnorm(A, n, m, 8)
    assert (norm, strat, res["best_rank"]) == want, (A.shape, n, m, kw, res["kernel_name"], norm, want[0])
    return res


# (d, rows): small enough for the oracle at every width (d = 1: L1 and LM),
# and for LDS at the widest (9 x 4096 words: 144 KiB)
ROWS = {1: 9, 2: 9, 3: 7, 4: 6, 5: 5, 6: 5, 7: 4, 8: 4, 9: 4, 10: 4}


@pytest.mark.parametrize("d", sorted(ROWS))
def test_columns_every_width_vs_oracle(pkg, d):
    rng = np.random.default_rng(700 + d)
    R = ROWS[d]
    seen = set()
    for c in CPL:
        if c > _max_cpl(d):
            break
        for C in (64 * c, 64 * c - 37 if c > 1 else 41):
            for m in ((0, 1) if d == 1 else (0,)):
                # LM: column 0 is the marginal, the other C - 1 take the lanes
                A = rng.integers(-3000, 3001, size=(R, C)).astype(np.int32)
                print("columns", d, m, A.shape, flush=True)
                res = _agree(pkg, A, d, m, kernel="columns")
                assert res["kernel_name"].startswith("cols<"), res["kernel_name"]
                seen.add(int(res["kernel_name"].split("cpl=")[1].split(",")[0]))
    assert seen == {c for c in CPL if c <= _max_cpl(d)}


@pytest.mark.parametrize("d", [1, 2, 3, 6, 10])
def test_columns_forced_segments(pkg, d):
    rng = np.random.default_rng(720 + d)
    R = ROWS[d] + (3 if d == 1 else 1)
    A = rng.integers(-500, 501, size=(R, 150)).astype(np.int32)
    for q in (1, 2, 3, 4):
        print("columns seg_log", d, q, flush=True)
        res = _agree(pkg, A, d, 0, kernel="columns", seg_log=q)
        assert f",q={q}>" in res["kernel_name"], res["kernel_name"]
    if d == 1:
        _agree(pkg, A, 1, 1, kernel="columns", seg_log=2)


def test_columns_lds_beyond_64k(pkg):
    """Rows whose LDS copy exceeds 64 KiB (17 x 1024 words, 68 KiB; 20 x 1536
    words, 120 KiB) still run, and agree with the oracle."""
    rng = np.random.default_rng(731)
    A = rng.integers(-100, 101, size=(17, 1000)).astype(np.int32)
    print("columns LDS 68 KiB", flush=True)
    res = _agree(pkg, A, 1, 0, kernel="columns")
    assert "cpl=16" in res["kernel_name"]
    B = rng.integers(-100, 101, size=(20, 1500)).astype(np.int32)
    print("columns LDS 120 KiB, rank range", flush=True)
    rb, re = 3 ** 19 // 2, 3 ** 19 // 2 + 20000
    norm, strat, res = pkg.calc_lnorm(B, 3, 0, kernel="columns", rank_begin=rb, rank_end=re)
    assert "cpl=24" in res["kernel_name"]
    assert (norm, res["best_rank"], strat) == oracle.walk(B, 3, 0, rb, re - 1)


def test_columns_auto_and_general(pkg):
    """AUTO takes the column walk where only the general kernel applied
    (int32 entries beyond the 16-bit walks, wider than the 32-bit tabulated
    walks), and it equals the general kernel there."""
    rng = np.random.default_rng(740)
    for R, C, d, r in ((12, 300, 3, 20000), (9, 400, 4, 20000), (8, 200, 7, 20000), (6, 700, 10, 3000),
                       (16, 1500, 1, 3000), (12, 700, 2, 30000)):
        A = rng.integers(-r, r + 1, size=(R, C)).astype(np.int32)
        print("columns auto", R, C, d, flush=True)
        a = pkg.calc_lnorm(A, d, 0)
        assert a[2]["kernel_name"].startswith("cols<"), (R, C, d, a[2]["kernel_name"])
        g = pkg.calc_lnorm(A, d, 0, kernel="general")
        assert a[:2] == g[:2] and a[2]["best_rank"] == g[2]["best_rank"], (R, C, d)
        assert oracle.objective(A, d, 0, a[1]) == a[0]


def test_columns_long_segments(pkg):
    """Segments at the walks' limits: 2^20 ranks for L1 (binary counter) and,
    for L2, whose column walk is the d-ary form with 15 counter digits, a
    requested 2^20 capped at 2^15; both against the general kernel."""
    rng = np.random.default_rng(750)
    for n, R, C in ((1, 22, 300), (2, 22, 700)):
        A = rng.integers(-20000, 20001, size=(R, C)).astype(np.int32)
        print("columns long segments", n, A.shape, flush=True)
        c = pkg.calc_lnorm(A, n, 0, kernel="columns", seg_log=20)
        assert c[2]["kernel_name"].endswith(",q=%d>" % (20 if n == 1 else 15)), c[2]["kernel_name"]
        g = pkg.calc_lnorm(A, n, 0, kernel="general")
        assert c[:2] == g[:2] and c[2]["best_rank"] == g[2]["best_rank"], (n, A.shape)

