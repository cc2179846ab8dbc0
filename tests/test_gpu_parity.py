"""Parity of the HIP engine with the oracle and the reference's golden
vectors (bit-exact norms, strategies and maximising Gray ranks).  GPU only:
every call goes through the C ABI of lib/liblnorm_hip.so."""
import ctypes

import numpy as np
import pytest

import oracle
from oracle import host_oracle as ho

pytestmark = pytest.mark.gpu

# rows the tabulated walk keeps in its subset tables, per d, at up to 8 words
# (16 columns); beyond, d = 3 keeps 6 rows up to 16 words, 5 up to 32 and 4
# beyond, d = 5 4 rows up to 16 words and 3 up to 32, d = 6 3 rows up to 24
# words, d = 7, 8, 9 3 rows up to 16 words, d = 10 2 rows
TABLE_ROWS = {2: 5, 3: 6, 4: 4, 5: 4, 6: 4, 7: 3, 8: 3, 9: 3, 10: 3}


def _laid_out(case):
    argv = ["L", str(case["threads"]), "m.txt"] + case["args"]
    threads, _, _, n0, stat, marg = ho.parse_args(argv, "omp")
    p = ho.load(case["matrix"].encode(), n0, stat, marg, "omp", threads)
    A = np.array(p.A, dtype=np.int32).reshape(p.R, p.C) if p.R else None
    return p, A


def _rand(rng, R, C, r):
    return rng.integers(-r, r + 1, size=(R, C)).astype(np.int32)


def _agree(eng, A, n, marg, **kw):
    want = oracle.calc_lnorm(A, n, marg, 8)
    norm, strat, res = eng.calc_lnorm(A, n, marg, **kw)
    assert (norm, strat, res["best_rank"]) == want, (A.shape, n, marg, kw, res["kernel_name"])
    return res


def test_selftest(pkg):
    print("selftest", flush=True)
    assert pkg.selftest() == 0, pkg.last_error()


def test_valu_rate_probe(pkg):
    """The measured rate of the walks' |.|-accumulates (v_sad_u16, v_sad_u32)
    lies below the published int32 VALU peak (256 CUs x 64 lanes x 2.4 GHz)
    and above half of it, the two at the same rate; v_add_u32 runs at least
    at that rate (it measured faster: VOP2 adds are not bound by the 64 lanes
    per clock of the published figure); a bad op is an error."""
    peak = 256 * 64 * 2.4e9
    rates = {}
    for op in ("sad_u16", "sad_u32", "add_u32"):
        print("valu rate", op, flush=True)
        rates[op], ms = pkg.valu_rate(op)
        print("  %.4e lane-ops/s, %.2f ms" % (rates[op], ms), flush=True)
    for op in ("sad_u16", "sad_u32"):
        assert 0.5 * peak < rates[op] < 1.01 * peak, rates
    assert abs(rates["sad_u16"] / rates["sad_u32"] - 1) < 0.03, rates
    assert rates["add_u32"] > 0.97 * rates["sad_u32"], rates
    with pytest.raises(KeyError):
        pkg.valu_rate("mul")
    rc = pkg.hip_lib().lnorm_hip_valu_rate(3, ctypes.byref(ctypes.c_double()), None)
    assert rc == -1, rc


def test_golden_vectors_all_kernels(pkg, golden):
    n_checked = 0
    for case in golden:
        if case["returncode"] != 0:
            continue
        p, A = _laid_out(case)
        if p.R == 0:
            continue
        want_line = case["stdout"].strip().splitlines()[-1]
        for kernel in ("auto", "plain", "general", "blocked", "plain16", "table16", "table32"):
            try:
                norm, strat, res = pkg.calc_lnorm(A, p.n, p.marginal, kernel=kernel)
            except pkg.LnormError as e:
                assert e.code == -5 and kernel in ("plain", "blocked", "plain16", "table16", "table32"), (
                    case["name"], kernel, str(e))
                continue
            line, fname, lines = ho.result_lines(p, norm, strat)
            assert line == want_line, (case["name"], kernel, res["kernel_name"])
            assert {fname: "".join(s + "\n" for s in lines)} == case["strategy"], (case["name"], kernel)
            n_checked += 1
    assert n_checked > 180


@pytest.mark.parametrize("n,marg", [(1, 0), (1, 1), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (7, 0),
                                    (8, 0), (9, 0), (10, 0)])
def test_random_vs_oracle_plain_and_general(pkg, n, marg):
    rng = np.random.default_rng(200 + 10 * n + marg)
    rmax = {1: 17, 2: 15, 3: 10, 4: 8, 5: 7, 6: 6, 7: 6, 8: 5, 9: 5, 10: 5}[n]
    for trial in range(12):
        R = int(rng.integers(1, rmax + 1))
        C = int(rng.integers(1, 41))
        A = _rand(rng, R, C, int(rng.choice([1, 2, 9, 100, 5000])))
        print("random n=%d marg=%d trial=%d shape=%s" % (n, marg, trial, A.shape), flush=True)
        _agree(pkg, A, n, marg, kernel="general")
        if C <= {1: 128, 2: 64, 3: 40, 4: 32, 5: 24, 6: 24}.get(n, 16):
            _agree(pkg, A, n, marg, kernel="plain")
        words = {2: 64, 3: 48, 4: 32, 5: 32, 6: 24}.get(n, 16)
        if n >= 2 and C <= 2 * words and int(np.abs(A).sum(axis=0).max()) <= 32767:
            res = _agree(pkg, A, n, marg, kernel="plain16")
            assert res["kernel_name"].startswith("plain16")
            if R - 1 >= TABLE_ROWS[n]:
                res = _agree(pkg, A, n, marg, kernel="table16")
                assert res["kernel_name"].startswith("table16")


@pytest.mark.parametrize("marg", [0, 1])
@pytest.mark.parametrize("P", [1, 2, 4, 8])
def test_blocked_vs_oracle(pkg, marg, P):
    rng = np.random.default_rng(300 + P + 17 * marg)
    kl = 6 + {1: 0, 2: 1, 4: 2, 8: 3}[P]
    # instantiated widths: at most 32 packed words, P * (words + marg) <= 192
    cmax = min(64 + marg, 2 * (192 // P - marg) + marg)
    for trial in range(6):
        R = int(rng.integers(kl + 2, kl + 9))
        C = int(rng.integers(2 + marg, cmax + 1))
        r = int(rng.choice([1, 3, 127, 300]))
        A = _rand(rng, R, C, r)
        print("blocked P=%d marg=%d trial=%d shape=%s r=%d" % (P, marg, trial, A.shape, r), flush=True)
        res = _agree(pkg, A, 1, marg, kernel="blocked", patterns_per_lane=P)
        assert res["kernel"] == pkg.KERNEL_BLOCKED


@pytest.mark.parametrize("marg", [0, 1])
@pytest.mark.parametrize("P", [1, 2, 4])
def test_blocked32_vs_oracle(pkg, marg, P):
    # 32-bit words: int16 entries (two int8 MFMA planes), up to 128 |.| columns
    rng = np.random.default_rng(500 + P + 31 * marg)
    kl = 6 + {1: 0, 2: 1, 4: 2}[P]
    for trial in range(5):
        R = int(rng.integers(kl + 4, kl + 9))
        C = int(rng.integers(2 + marg, 129 + marg))
        r = int(rng.choice([2, 300, 30000]))
        A = _rand(rng, R, C, r)
        print("blocked32 P=%d marg=%d trial=%d shape=%s r=%d" % (P, marg, trial, A.shape, r), flush=True)
        res = _agree(pkg, A, 1, marg, kernel="blocked", patterns_per_lane=P, word_bits=32)
        assert res["kernel_name"].startswith("blocked32"), res["kernel_name"]


def test_int16_entries_auto_and_blocked32(pkg):
    from lnorm_amd import gen
    # int16 entries: column sums beyond 2^15; AUTO takes the 32-bit tabulated
    # walk, and the blocked kernel runs them in 32-bit words
    A = gen.gen_matrix(128, 20, 30000, 105).T.copy()       # 20 x 128, 2^19 sign vectors
    res = _agree(pkg, A, 1, 0)
    assert res["kernel_name"].startswith("table32<L1,words=128"), res["kernel_name"]
    res = _agree(pkg, A, 1, 0, kernel="blocked")
    assert res["kernel_name"].startswith("blocked32<nc=32,P=2,W=4"), res["kernel_name"]
    res = _agree(pkg, A[:, :50].copy(), 1, 1)
    assert res["kernel_name"].startswith("table32<LM"), res["kernel_name"]
    res = _agree(pkg, A[:, :50].copy(), 1, 1, kernel="blocked")
    assert res["kernel_name"].startswith("blocked32"), res["kernel_name"]


def test_blocked32_table_matches_numpy(pkg):
    rng = np.random.default_rng(19)
    for marg, P, r, C in [(0, 4, 30000, 100), (1, 2, 20000, 41), (0, 1, 100, 7)]:
        kl = 6 + {1: 0, 2: 1, 4: 2}[P]
        A = _rand(rng, kl + 4, C, r)
        tab = pkg.debug_table(A, marg, P, word_bits=32)
        cols = list(range(marg, C))
        W = 1 if len(cols) <= 32 else (2 if len(cols) <= 64 else 4)
        NC = ((len(cols) + W - 1) // W + 7) // 8 * 8
        NW = W * NC + marg
        tab = tab.reshape(P, NW, 64)
        for p in range(64 * P):
            signs = np.array([1 if (p >> i) & 1 else -1 for i in range(kl)])
            Y = signs @ A[:kl].astype(np.int64)
            rr, L = divmod(p, 64)
            for c in range(W * NC):
                y = int(Y[cols[c]]) if c < len(cols) else 0
                assert int(tab[rr, c, L]) == (y + 2 ** 31) & 0xFFFFFFFF, (marg, P, p, c)
            if marg:
                assert int(tab[rr, W * NC, L]) == int(Y[0]) & 0xFFFFFFFF


def test_blocked_table_matches_numpy(pkg):
    rng = np.random.default_rng(9)
    for marg, P, r in [(0, 4, 127), (1, 2, 127), (0, 2, 3000), (1, 8, 20000), (0, 1, 5)]:
        kl = 6 + {1: 0, 2: 1, 4: 2, 8: 3}[P]
        C = 37
        A = _rand(rng, kl + 3, C, r)
        A[:, marg:] //= max(1, (kl + 3) * r // 30000)    # keep |.| column sums < 2^15
        tab = pkg.debug_table(A, marg, P)
        npk = ((C - marg) + 1) // 2
        npk = max(4, npk + (npk & 1))
        NW = npk + marg
        tab = tab.reshape(P, NW, 64)
        cols = list(range(marg, C))
        for p in range(64 * P):
            signs = np.array([1 if (p >> i) & 1 else -1 for i in range(kl)])
            Y = signs @ A[:kl].astype(np.int64)
            rr, L = divmod(p, 64)
            for c in range(npk):
                lo = cols[2 * c] if 2 * c < len(cols) else None
                hi = cols[2 * c + 1] if 2 * c + 1 < len(cols) else None
                ylo = Y[lo] if lo is not None else 0
                yhi = Y[hi] if hi is not None else 0
                assert int(tab[rr, c, L]) == ((int(yhi) + 32768) << 16) | (int(ylo) + 32768), (marg, P, p, c)
            if marg:
                assert int(tab[rr, npk, L]) == int(Y[0]) & 0xFFFFFFFF


def test_segment_size_independence(pkg):
    rng = np.random.default_rng(4)
    A = _rand(rng, 16, 12, 2)          # tiny entries: many ties
    want = oracle.calc_lnorm(A, 1, 0, 8)
    for s in (1, 3, 7, 12, 15):
        norm, strat, res = pkg.calc_lnorm(A, 1, 0, kernel="plain", seg_log=s)
        assert (norm, strat, res["best_rank"]) == want
    for wl in (1, 2, 5):
        norm, strat, res = pkg.calc_lnorm(A, 1, 0, kernel="blocked", wave_log2=wl, patterns_per_lane=2)
        assert (norm, strat, res["best_rank"]) == want
    B = _rand(rng, 9, 7, 3)
    want3 = oracle.calc_lnorm(B, 3, 0, 8)
    for s in (1, 2, 5, 8):
        for k in ("plain", "plain16", "table16"):
            norm, strat, res = pkg.calc_lnorm(B, 3, 0, kernel=k, seg_log=s)
            assert (norm, strat, res["best_rank"]) == want3
    want2 = oracle.calc_lnorm(A, 2, 0, 8)
    for s in (1, 4, 9, 15):
        norm, strat, res = pkg.calc_lnorm(A, 2, 0, kernel="plain16", seg_log=s)
        assert (norm, strat, res["best_rank"]) == want2


def test_errors(pkg):
    big = np.full((3, 2), 2 ** 30, dtype=np.int32)
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(big, 1)
    assert e.value.code == -2
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(np.ones((42, 3), np.int32), 3)        # maxRows(3) = 41
    assert e.value.code == -3
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(np.ones((12, 3), np.int32), 2, kernel="blocked")
    assert e.value.code == -5
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(np.ones((3, 3), np.int32), 11)
    assert e.value.code == -3


def test_simple_abi_and_batch(pkg):
    rng = np.random.default_rng(12)
    probs = []
    for i in range(20):
        n = [1, 1, 2, 3, 4][i % 5]
        marg = 1 if (n == 1 and i % 2) else 0
        R = int(rng.integers(2, 12 if n <= 2 else 7))
        probs.append((_rand(rng, R, int(rng.integers(1, 30)), 50), n, marg))
    out = pkg.calc_batch(probs)
    for (A, n, marg), (norm, strat, status) in zip(probs, out):
        assert status == 0
        assert (norm, strat) == oracle.calc_lnorm(A, n, marg, 4)[:2]
        assert pkg.calc_lnorm_simple(A, n, marg) == (norm, strat)


def test_workloads_vs_oracle(pkg):
    from lnorm_amd import gen
    # 20x20 L_1 from seed 101 (2^19 sign vectors) on every kernel
    A = gen.gen_matrix(20, 20, 10, 101)
    for k in ("auto", "plain", "blocked", "general"):
        print("cfg1 kernel", k, flush=True)
        _agree(pkg, A, 1, 0, kernel=k)
    # L_2 and L_3 of the 18 x 32 file (3^17 labelings at d = 3)
    B = gen.gen_matrix(18, 32, 50, 103)
    for k in ("auto", "plain", "plain16"):
        print("cfg3 L2/L3", k, flush=True)
        _agree(pkg, B, 2, 0, kernel=k)
        _agree(pkg, B, 3, 0, kernel=k)
    assert _agree(pkg, B, 3, 0)["kernel_name"].startswith("table16<")


def test_blocked_equals_plain_large(pkg):
    from lnorm_amd import gen
    # 2^35 marginal sign vectors (36 x 36) and 2^33 for a 34 x 40 matrix
    for (R, C, r, seed, marg) in [(36, 36, 100, 102, 1), (34, 40, 100, 7, 0), (33, 64, 127, 104, 0)]:
        A = gen.gen_matrix(R, C, r, seed)
        print("large blocked vs plain", R, C, marg, flush=True)
        nb, sb, rb = pkg.calc_lnorm(A, 1, marg, kernel="blocked")
        npl, sp, rp = pkg.calc_lnorm(A, 1, marg, kernel="plain")
        assert (nb, sb, rb["best_rank"]) == (npl, sp, rp["best_rank"])
        assert oracle.objective(A, 1, marg, sb) == nb


@pytest.mark.parametrize("n,marg,R,cols", [
    (1, 0, 14, [129, 200, 256, 300, 384, 500, 700, 1000]),
    (1, 1, 13, [130, 400, 1024]),
    (2, 0, 12, [200, 300, 512]),
    (1, 0, 4, [200]),
])
def test_wide_walk_vs_oracle(pkg, n, marg, R, cols):
    """The W-wave walk (columns split over the waves of a workgroup, LDS
    combine every 8 steps): bit-exact norm, strategy and rank for widths the
    register-resident walks do not hold."""
    rng = np.random.default_rng(700 + 10 * n + marg + R)
    for C in cols:
        A = _rand(rng, R, C, 40)
        print("wide", n, marg, A.shape, flush=True)
        # forced: AUTO takes the tabulated walks where the sums fit 16 bits
        res = _agree(pkg, A, n, marg, kernel="wide")
        assert res["kernel_name"].startswith("wide<"), res["kernel_name"]
    # narrower problems on the forced W-wave walk (AUTO keeps them elsewhere)
    A = _rand(rng, R, 65, 40)
    assert _agree(pkg, A, n, marg, kernel="wide")["kernel_name"].startswith("wide<")


# (d, rows, column counts): the first count is the narrowest that one wave's
# packed walk does not hold; the others reach the 2-, 4- and 8-wave shapes
# (d * words-per-wave state registers: 40 words at d = 3 ... 12 at d >= 9)
WIDE16 = [
    (3, 12, [97, 160, 161, 301, 640]),
    (4, 10, [65, 112, 224, 447]),
    (5, 8, [65, 96, 180, 384]),
    (6, 8, [49, 80, 160, 320]),
    (7, 7, [33, 64, 128, 255]),
    (8, 7, [33, 56, 112, 224]),
    (9, 6, [33, 48, 96, 192]),
    (10, 6, [33, 47, 96, 191]),
]


@pytest.mark.parametrize("d,R,cols", WIDE16)
def test_wide16_walk_vs_oracle(pkg, d, R, cols):
    """The packed d-ary walk over W waves: bit-exact norm, strategy and rank
    beyond the widths one wave holds, at the default segment length and at
    forced ones (d^q steps per segment, combined in groups of 8 with a
    shorter last group)."""
    rng = np.random.default_rng(800 + d)
    for C in cols:
        A = _rand(rng, R, C, 60)
        print("wide16", d, A.shape, flush=True)
        res = _agree(pkg, A, d, 0, kernel="wide")
        assert res["kernel_name"].startswith("wide16<"), res["kernel_name"]
    # at these widths AUTO takes a W-wave walk (tabulated where its widths
    # reach)
    A = _rand(rng, R, cols[-1], 60)
    assert _agree(pkg, A, d, 0)["kernel_name"].startswith(("table16w<", "wide16<"))
    A = _rand(rng, R, cols[1], 60)
    for q in (2, 3, 5):   # q = 5 at d >= 9 (5 enumerated rows): one segment
        print("wide16 seg_log", d, A.shape, q, flush=True)
        res = _agree(pkg, A, d, 0, seg_log=q, kernel="wide")
        assert res["kernel_name"].startswith("wide16<") and f"q={q}>" in res["kernel_name"], res["kernel_name"]


# (d, rows, column counts) for the tabulated walk: every instantiated width
# (8..48 words) and odd column counts; rows leave several high digits
TABLE16 = [
    (2, 15, [1, 5, 16, 24, 33, 40, 56, 64, 65, 100, 128, 129, 160, 192]),
    (3, 12, [1, 5, 16, 17, 24, 33, 40, 48, 56, 64, 65, 80, 96, 97, 112, 128]),
    (4, 10, [3, 16, 24, 31, 40, 48, 56, 64, 65, 80, 96]),
    (5, 8, [7, 16, 24, 33, 40, 48, 56, 64, 65, 80]),
    (6, 7, [9, 16, 24, 32, 40, 48, 49, 56, 64]),
    (7, 6, [5, 16, 24, 32, 33, 40, 48, 56, 64]),
    (8, 6, [5, 16, 24, 32, 33, 40, 48]),
    (9, 5, [5, 16, 24, 32, 33, 40, 48]),
    (10, 5, [5, 16, 24, 32]),
]


@pytest.mark.parametrize("d,R,cols", TABLE16)
def test_table16_walk_vs_oracle(pkg, d, R, cols):
    """The tabulated packed walk (the M fastest rows as per-label subset
    tables): bit-exact norm, strategy and rank at every width, at forced
    segment lengths q = M .. M+2 (one to d^2 high steps per segment), and
    with exactly M enumerated rows (a single block)."""
    rng = np.random.default_rng(900 + d)
    M = TABLE_ROWS[d]
    for C in cols:
        A = _rand(rng, R, C, 60)
        print("table16", d, A.shape, flush=True)
        res = _agree(pkg, A, d, 0)
        assert res["kernel_name"].startswith("table16<"), res["kernel_name"]
    A = _rand(rng, R, cols[1], 60)
    for q in (M, M + 1, M + 2):
        if q > R - 1:
            continue
        print("table16 seg_log", d, A.shape, q, flush=True)
        res = _agree(pkg, A, d, 0, seg_log=q, kernel="table16")
        assert f",q={q}>" in res["kernel_name"], res["kernel_name"]
    A = _rand(rng, M + 1, 9, 60)
    res = _agree(pkg, A, d, 0, kernel="table16")
    assert res["kernel_name"].startswith("table16<"), res["kernel_name"]
    # fewer enumerated rows than table rows: AUTO takes the packed walk
    A = _rand(rng, M, 9, 60)
    assert not _agree(pkg, A, d, 0)["kernel_name"].startswith("table16<")
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(A, d, 0, kernel="table16")
    assert e.value.code == -5


@pytest.mark.parametrize("marg", [0, 1])
def test_table16_binary_vs_oracle(pkg, marg):
    """The tabulated L1 / LM walk (the 5 fastest rows as a sign-pattern
    table): bit-exact norm, strategy and rank at every width (4..32 words in
    steps of 2, then 40..64 words of
    |.| columns, odd counts, LM's column 0 outside the words), at forced
    segment lengths, with exactly 5 enumerated rows, and AUTO's choice of it
    beyond the 16-bit blocked kernel's 64 |.| columns."""
    rng = np.random.default_rng(950 + marg)
    for R in (6, 9, 14):
        for C in (1 + marg, 2, 7, 8, 9, 11, 16, 17, 19, 27, 31, 33, 35, 43, 51, 59, 64, 65, 100, 128 + marg):
            A = _rand(rng, R, C, 60)
            print("table16 L1", marg, A.shape, flush=True)
            res = _agree(pkg, A, 1, marg, kernel="table16")
            assert res["kernel_name"].startswith("table16<L"), res["kernel_name"]
    A = _rand(rng, 14, 40, 60)
    for q in (5, 6, 9, 13):
        print("table16 L1 seg_log", marg, q, flush=True)
        res = _agree(pkg, A, 1, marg, kernel="table16", seg_log=q)
        assert f",q={q}>" in res["kernel_name"], res["kernel_name"]
    A = _rand(rng, 14, 100, 60)
    assert _agree(pkg, A, 1, marg)["kernel_name"].startswith("table16<L"), "AUTO beyond 64 |.| columns"
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(_rand(rng, 5, 9, 60), 1, marg, kernel="table16")
    assert e.value.code == -5


@pytest.mark.parametrize("marg", [0, 1])
def test_table16_binary_wide_vs_oracle(pkg, marg):
    """The tabulated L1 / LM walk beyond 128 |.| columns (80..192 words, 4
    table rows): bit-exact against the oracle, AUTO's choice there, and equal
    to the W-wave walk on 2^29 sign vectors."""
    rng = np.random.default_rng(970 + marg)
    for R in (6, 12):
        for C in (129, 130, 161, 200, 256, 257, 320, 384 + marg):
            A = _rand(rng, R, C, 40)
            print("table16 wide", marg, A.shape, flush=True)
            res = _agree(pkg, A, 1, marg)
            assert res["kernel_name"].startswith("table16<L"), res["kernel_name"]
    from lnorm_amd import gen
    A = gen.gen_matrix(30, 300, 60, 71 + marg)
    t = pkg.calc_lnorm(A, 1, marg)
    w = pkg.calc_lnorm(A, 1, marg, kernel="wide")
    assert t[2]["kernel_name"].startswith("table16<L") and w[2]["kernel_name"].startswith("wide<")
    assert t[:2] == w[:2] and t[2]["best_rank"] == w[2]["best_rank"]


@pytest.mark.parametrize("marg", [0, 1])
def test_table16w_binary_vs_oracle(pkg, marg):
    """The tabulated L1 / LM walk over W waves beyond 384 |.| columns (8, 10,
    12 or 16 waves of 32 words, 16 of 64: partial pattern sums through LDS):
    bit-exact norm, strategy and rank at every shape, at forced segment
    lengths, with the fewest enumerated rows it accepts, and an error below
    that."""
    rng = np.random.default_rng(1010 + marg)
    for R in (6, 10):
        for C in (385 + marg, 400, 512 + marg, 513 + marg, 700, 1024 + marg, 1025 + marg, 1500, 2048 + marg):
            A = _rand(rng, R, C, 60)
            print("table16w L1", marg, A.shape, flush=True)
            res = _agree(pkg, A, 1, marg)
            assert res["kernel_name"].startswith("table16w<L"), res["kernel_name"]
    for C, M in ((450, 5), (620, 4), (740, 4), (900, 4), (1800, 4)):
        A = _rand(rng, 14, C, 60)
        for q in (M, M + 1, M + 3, 13):
            print("table16w L1 seg_log", marg, A.shape, q, flush=True)
            res = _agree(pkg, A, 1, marg, kernel="table16", seg_log=q)
            assert f"M={M},q={q}>" in res["kernel_name"], res["kernel_name"]
        A = _rand(rng, M + 1, C, 60)
        assert _agree(pkg, A, 1, marg)["kernel_name"].startswith("table16w<L")
        with pytest.raises(pkg.LnormError) as e:
            pkg.calc_lnorm(_rand(rng, M, C, 60), 1, marg, kernel="table16")
        assert e.value.code == -5


def test_table16w_binary_equals_wide_larger(pkg):
    """2^23 sign vectors of 1000 columns: the tabulated and binary W-wave
    walks agree on norm, strategy and rank, and the strategy's objective is
    the norm."""
    from lnorm_amd import gen
    for (R, C, marg) in ((24, 1000, 0), (22, 900, 1), (20, 500, 0)):
        A = gen.gen_matrix(R, C, 60, 80 + marg)
        print("table16w vs wide", R, C, marg, flush=True)
        t = pkg.calc_lnorm(A, 1, marg)
        w = pkg.calc_lnorm(A, 1, marg, kernel="wide")
        assert t[2]["kernel_name"].startswith("table16w<L") and w[2]["kernel_name"].startswith("wide<")
        assert t[:2] == w[:2] and t[2]["best_rank"] == w[2]["best_rank"], (R, C, marg)
        assert oracle.objective(A, 1, marg, t[1]) == t[0]


def test_table16_binary_equals_blocked_large(pkg):
    """2^33 sign vectors: the tabulated L1 / LM walk equals the blocked
    kernel (norm, strategy, rank), and the objective of the strategy is the
    norm."""
    from lnorm_amd import gen
    for (R, C, marg) in ((34, 40, 0), (34, 36, 1)):
        A = gen.gen_matrix(R, C, 100, 60 + marg)
        print("table16 vs blocked", R, C, marg, flush=True)
        t = pkg.calc_lnorm(A, 1, marg, kernel="table16")
        b = pkg.calc_lnorm(A, 1, marg, kernel="blocked")
        assert t[:2] == b[:2] and t[2]["best_rank"] == b[2]["best_rank"], (R, C, marg)
        assert oracle.objective(A, 1, marg, t[1]) == t[0]


# (d, rows, column counts) for the W-wave tabulated walk: beyond one wave's
# tables, at every wave count it reaches (2, 3, 4, 5, 6 and 8 waves for
# d <= 6, not 5 at d = 6; up to 4 beyond)
WIDE_T16 = [
    (2, 12, [193, 200, 256, 257, 350, 400, 512]),
    (3, 10, [129, 130, 256, 300, 350, 512]),
    (4, 9, [97, 128, 150, 200, 256, 300, 350, 512]),
    (5, 8, [81, 128, 129, 256, 300, 350, 512]),
    (6, 7, [65, 128, 150, 200, 350, 512]),
    (7, 7, [65, 128, 129, 256]),
    (8, 6, [49, 96, 97, 192]),
    (9, 6, [49, 96, 120, 192]),
    (10, 6, [33, 64, 65, 128]),
]
# words per wave of the W-wave tabulated walk, its wave counts, and the
# single-wave walk's widest word count (beyond which the W waves take over)
WT16_NW = {2: 32, 3: 32, 4: 32, 5: 32, 6: 32, 7: 32, 8: 24, 9: 24, 10: 16}
WT16_WAVES = {d: [w for w in (2, 3, 4, 5, 6, 8) if w <= (8 if d <= 6 else 4) and not (d == 6 and w == 5)]
              for d in range(2, 11)}
BLK16_MAX_WORDS = {2: 96, 3: 64, 4: 48, 5: 40, 6: 32, 7: 32, 8: 24, 9: 24, 10: 16}


@pytest.mark.parametrize("d,R,cols", WIDE_T16)
def test_table16w_walk_vs_oracle(pkg, d, R, cols):
    """The tabulated walk over W waves (partial subset tables summed through
    LDS, the block's labelings split over the waves): bit-exact norm,
    strategy and rank at every wave count, at forced segment lengths and
    with the fewest enumerated rows it accepts."""
    rng = np.random.default_rng(990 + d)
    seen = set()
    for C in cols:
        A = _rand(rng, R, C, 50)
        print("table16w", d, A.shape, flush=True)
        res = _agree(pkg, A, d, 0)
        assert res["kernel_name"].startswith("table16w<"), res["kernel_name"]
        seen.add(int(res["kernel_name"].split("words=")[1].split(",")[0]))
    # every wave count the walk reaches is covered
    reach = {WT16_NW[d] * w for w in WT16_WAVES[d] if WT16_NW[d] * w > BLK16_MAX_WORDS[d]}
    assert seen == reach, (d, sorted(seen), sorted(reach))
    A = _rand(rng, R, cols[1], 50)
    M = int(pkg.calc_lnorm(A, d, 0)[2]["kernel_name"].split("M=")[1].split(",")[0])
    for q in (M, M + 1, M + 2):
        if q > R - 1:
            continue
        print("table16w seg_log", d, A.shape, q, flush=True)
        res = _agree(pkg, A, d, 0, seg_log=q, kernel="table16")
        assert res["kernel_name"].startswith("table16w<") and f",q={q}>" in res["kernel_name"], res["kernel_name"]
    A = _rand(rng, M + 1, cols[1], 50)
    assert _agree(pkg, A, d, 0)["kernel_name"].startswith("table16w<")


def test_table16w_equals_wide16_larger(pkg):
    """Larger problems: the tabulated and packed W-wave walks agree on norm,
    strategy and rank, and the strategy's objective is the norm."""
    from lnorm_amd import gen
    for (R, C, d) in ((16, 300, 3), (12, 400, 4), (11, 150, 5), (9, 100, 9)):
        A = gen.gen_matrix(R, C, 50, 50 + d)
        print("table16w vs wide16", R, C, d, flush=True)
        t = pkg.calc_lnorm(A, d, 0)
        w = pkg.calc_lnorm(A, d, 0, kernel="wide")
        assert t[2]["kernel_name"].startswith("table16w<") and w[2]["kernel_name"].startswith("wide16<")
        assert t[:2] == w[:2] and t[2]["best_rank"] == w[2]["best_rank"], (R, C, d)
        assert oracle.objective(A, d, 0, t[1]) == t[0]


def test_table16_equals_plain16_larger(pkg):
    """At sizes the oracle does not reach: the tabulated walk and the packed
    walk agree on norm, strategy and rank (3^23 labelings for the 24 x 32 L3
    workload), and the objective of the strategy is the norm."""
    from lnorm_amd import gen
    for (R, C, d) in ((24, 32, 3), (16, 64, 3), (13, 40, 4), (11, 24, 6), (9, 16, 10)):
        A = gen.gen_matrix(R, C, 50, 40 + d)
        print("table16 vs plain16", R, C, d, flush=True)
        t = pkg.calc_lnorm(A, d, 0)
        p = pkg.calc_lnorm(A, d, 0, kernel="plain16")
        assert t[2]["kernel_name"].startswith("table16<"), t[2]["kernel_name"]
        assert t[:2] == p[:2] and t[2]["best_rank"] == p[2]["best_rank"], (R, C, d)
        assert oracle.objective(A, d, 0, t[1]) == t[0]


def test_wide16_equals_general_larger(pkg):
    from lnorm_amd import gen
    for (R, C, d) in ((16, 300, 3), (11, 150, 5), (8, 100, 9)):
        A = gen.gen_matrix(R, C, 50, 30 + d)
        print("wide16 vs general", R, C, d, flush=True)
        w = pkg.calc_lnorm(A, d, 0, kernel="wide")
        g = pkg.calc_lnorm(A, d, 0, kernel="general")
        assert w[2]["kernel_name"].startswith("wide16<"), w[2]["kernel_name"]
        assert w[:2] == g[:2] and w[2]["best_rank"] == g[2]["best_rank"]
        assert oracle.objective(A, d, 0, w[1]) == w[0]


def test_wide16_needs_16bit_sums(pkg):
    """Column sums above 32767 do not fit the packed halves: AUTO takes the
    32-bit tabulated W-wave walk (not the general kernel), and a forced
    packed W-wave walk is refused."""
    rng = np.random.default_rng(5)
    A = _rand(rng, 6, 120, 30000)
    res = _agree(pkg, A, 3, 0)
    assert res["kernel_name"].startswith("table32w<"), res["kernel_name"]
    assert _agree(pkg, A, 3, 0, kernel="general")["kernel_name"].startswith("general<")
    with pytest.raises(pkg.LnormError) as e:
        pkg.calc_lnorm(A, 3, 0, kernel="wide")
    assert e.value.code == -5


def test_wide_equals_general_larger(pkg):
    from lnorm_amd import gen
    for (R, C, n) in ((24, 600, 1), (20, 300, 2)):
        A = gen.gen_matrix(R, C, 30, 17 + n)
        print("wide vs general", R, C, n, flush=True)
        w = pkg.calc_lnorm(A, n, 0, kernel="wide")
        g = pkg.calc_lnorm(A, n, 0, kernel="general")
        assert w[:2] == g[:2] and w[2]["best_rank"] == g[2]["best_rank"]
        assert oracle.objective(A, n, 0, w[1]) == w[0]
