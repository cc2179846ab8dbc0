"""Pins the oracle (oracle/lnorm_oracle.c and oracle/host_oracle.py):
golden vectors of the reference L_OpenMP, an independent brute force,
closed forms and invariants.  CPU only."""
import itertools

import numpy as np
import pytest

import oracle
from oracle import host_oracle as ho


def _run_oracle_cli(case):
    """The oracle's restatement of `L_OpenMP T file d [p] [m]`: stdout lines,
    exit status, strategy files."""
    argv = ["L_OpenMP", str(case["threads"]), "matrix.txt"] + case["args"]
    threads, _, _, n0, stat, marg = ho.parse_args(argv, "omp")
    try:
        p = ho.load(case["matrix"].encode(), n0, stat, marg, "omp", threads)
    except ho.HostExit as e:
        return e.lines, 255, {}
    lines = list(p.lines)
    if p.R == 0:
        lines.append("This is a zero matrix.")
        norm, strat = 0, []
    else:
        A = np.array(p.A, dtype=np.int32).reshape(p.R, p.C)
        norm, strat, _ = oracle.calc_lnorm(A, p.n, p.marginal, threads)
    line, fname, slines = ho.result_lines(p, norm, strat)
    lines.append(line)
    return lines, 0, {fname: "".join(s + "\n" for s in slines)}


def test_golden_vectors(golden):
    assert len(golden) >= 60
    for case in golden:
        lines, rc, strat = _run_oracle_cli(case)
        assert rc == (case["returncode"] & 0xFF), case["name"]
        assert "".join(l + "\n" for l in lines) == case["stdout"], case["name"]
        assert strat == case["strategy"], case["name"]


def _rand(rng, R, C, r):
    return rng.integers(-r, r + 1, size=(R, C)).astype(np.int32)


@pytest.mark.parametrize("n,marg", [(1, 0), (1, 1), (2, 0), (3, 0), (4, 0), (5, 0)])
def test_walk_matches_bruteforce(n, marg):
    rng = np.random.default_rng(1000 + 10 * n + marg)
    for trial in range(25):
        R = int(rng.integers(1, 9 if n <= 2 else 6))
        C = int(rng.integers(1, 7))
        A = _rand(rng, R, C, int(rng.choice([1, 3, 50])))
        nb, rb = oracle.brute(A, n, marg)
        for threads in (1, 3, 8):
            nw, sw, rw = oracle.calc_lnorm(A, n, marg, threads)
            assert (nw, rw) == (nb, rb), (n, marg, A.tolist(), threads)
            assert sw == oracle.labels_of_rank(rb, R, n)
            assert oracle.objective(A, n, marg, sw) == nb


def test_digits_closed_form():
    # BRGC bit i of N is bit i of N ^ (N >> 1); the reflected ternary code
    # visits every labeling once and consecutive words differ in one digit by 1.
    for N in range(300):
        for i in range(9):
            assert oracle.digit(N, i, 2) == ((N ^ (N >> 1)) >> i) & 1
    for n in (3, 4, 5):
        seen = set()
        prev = None
        for N in range(n ** 4):
            w = tuple(oracle.digit(N, i, n) for i in range(4))
            seen.add(w)
            if prev is not None:
                diff = [abs(a - b) for a, b in zip(w, prev)]
                assert sorted(diff) == [0, 0, 0, 1]
            prev = w
        assert len(seen) == n ** 4


def test_closed_forms():
    assert oracle.calc_lnorm([[1, 1], [1, -1]], 1)[0] == 2                     # CHSH
    for k in (1, 2, 5, 9):
        assert oracle.calc_lnorm(np.eye(k, dtype=np.int32), 1)[0] == k        # identity
    assert oracle.calc_lnorm(np.ones((3, 5), np.int32), 1)[0] == 15           # all-ones
    u, v = np.array([1, -2, 3, 5]), np.array([2, 1, -1, 4, -3])
    assert oracle.calc_lnorm(np.outer(u, v).astype(np.int32), 1)[0] == 11 * 11  # rank one
    rng = np.random.default_rng(7)
    for R in (2, 3, 4):
        A = _rand(rng, R, 5, 20)
        # d >= k: every row gets its own label, L_k = sum |A|
        assert oracle.calc_lnorm(A, R)[0] == int(np.abs(A).sum())
    assert oracle.calc_lnorm(np.zeros((4, 3), np.int32), 1)[0] == 0
    assert oracle.calc_lnorm(np.zeros((4, 3), np.int32), 3)[0] == 0


def test_invariants():
    rng = np.random.default_rng(11)
    for _ in range(10):
        R, C = int(rng.integers(2, 8)), int(rng.integers(2, 8))
        A = _rand(rng, R, C, 30)
        l1 = oracle.calc_lnorm(A, 1)[0]
        assert oracle.calc_lnorm(A.T.copy(), 1)[0] == l1                       # transpose
        P = A[rng.permutation(R)][:, rng.permutation(C)]
        S = P * rng.choice([-1, 1], size=(R, 1)) * rng.choice([-1, 1], size=(1, C))
        assert oracle.calc_lnorm(S.astype(np.int32), 1)[0] == l1             # perms, sign flips
        assert oracle.calc_lnorm((3 * A).astype(np.int32), 1)[0] == 3 * l1   # scaling
        Z = np.vstack([A, np.zeros((1, C), np.int32)])
        assert oracle.calc_lnorm(Z, 1)[0] == l1                                # zero row
        D = np.vstack([A, A[:1]])
        prev = oracle.calc_lnorm(A, 1)[0]
        for d in (2, 3, 4):
            ld = oracle.calc_lnorm(A, d)[0]
            assert ld >= prev                                                   # monotone in d
            prev = ld
            Pc = A[:, rng.permutation(C)] * rng.choice([-1, 1], size=(1, C))
            assert oracle.calc_lnorm(Pc.astype(np.int32), d)[0] == ld          # column perms/signs
        assert oracle.calc_lnorm(D, 1)[0] >= l1


def test_worker_count_independence():
    rng = np.random.default_rng(5)
    A = _rand(rng, 10, 6, 2)        # small entries: many ties
    ref = oracle.calc_lnorm(A, 1, 0, 1)
    for t in (2, 3, 7, 64, 511):
        assert oracle.calc_lnorm(A, 1, 0, t) == ref
    ref3 = oracle.calc_lnorm(A[:7], 3, 0, 1)
    for t in (2, 5, 13):
        assert oracle.calc_lnorm(A[:7], 3, 0, t) == ref3


def test_walk_segments_compose():
    # the max over segments equals the whole range's max, first rank kept
    rng = np.random.default_rng(3)
    A = _rand(rng, 9, 5, 4)
    whole = oracle.walk(A, 2, 0, 0, 255)
    parts = [oracle.walk(A, 2, 0, a, a + 31) for a in range(0, 256, 32)]
    best = max(p[0] for p in parts)
    first = min(p[1] for p in parts if p[0] == best)
    assert (best, first) == whole[:2]


def test_planted_mistakes_are_caught(golden):
    # Each plausible slip in a restatement must break a golden vector:
    # dropping the fixed last row, flipping the label convention, or using
    # |y_0| for the marginal column.
    cases = {c["name"]: c for c in golden}

    def norm_of(case, mutate):
        argv = ["x", str(case["threads"]), "m.txt"] + case["args"]
        threads, _, _, n0, stat, marg = ho.parse_args(argv, "omp")
        p = ho.load(case["matrix"].encode(), n0, stat, marg, "omp", threads)
        A = np.array(p.A, dtype=np.int32).reshape(p.R, p.C)
        return mutate(A, p)

    def drop_last(A, p):
        return oracle.calc_lnorm(A[:-1], p.n, p.marginal)[0] if p.R > 1 else -1

    def abs_marginal(A, p):
        return oracle.calc_lnorm(A, p.n, 0)[0]

    def unswapped_marginal(A, p):
        B = A.copy()
        B[[0, -1]] = B[[-1, 0]]
        return oracle.calc_lnorm(B, p.n, p.marginal)[0]

    def want(name):
        return int(cases[name]["stdout"].strip().splitlines()[-1].split()[-1])

    assert norm_of(cases["cfg1_l1_20x20_seed101"], drop_last) != want("cfg1_l1_20x20_seed101")
    assert norm_of(cases["lm_16x40"], abs_marginal) != want("lm_16x40")
    assert norm_of(cases["prep_lm_n"], abs_marginal) != want("prep_lm_n")
    assert norm_of(cases["cfg2_lm_20x20"], unswapped_marginal) != want("cfg2_lm_20x20")
    # label convention: the strategy file of an asymmetric case pins 1 <-> +1
    c = cases["transpose_l1"]
    lines, _, strat = _run_oracle_cli(c)
    flipped = {k: "".join("%d\n" % -int(x) for x in v.split()) for k, v in strat.items()}
    assert strat == c["strategy"] and flipped != c["strategy"]


def test_host_preprocessing_preserves_norm():
    rng = np.random.default_rng(17)
    for trial in range(30):
        R, C = int(rng.integers(2, 7)), int(rng.integers(2, 7))
        M = _rand(rng, R, C, 3)
        # plant proportional rows/columns and zero lines
        if R > 2:
            M[2] = -2 * M[0] if trial % 2 else 3 * M[1]
        if C > 3:
            M[:, 3] = M[:, 1] * (-1 if trial % 3 else 2)
        if trial % 4 == 0:
            M[R - 1] = 0
        text = "".join(" ".join(str(v) for v in r) + "\n" for r in M).encode()
        for d in (1, 2, 3):
            for marg in ((0, 1) if d == 1 else (0,)):
                norms = set()
                for stat in ("n", "y", "r", "c"):
                    try:
                        p = ho.load(text, d, stat, marg, "omp", 4)
                    except ho.HostExit:
                        continue
                    if p.R == 0:
                        norms.add(0)
                        continue
                    A = np.array(p.A, dtype=np.int32).reshape(p.R, p.C)
                    norms.add(oracle.calc_lnorm(A, p.n, p.marginal, 4)[0])
                assert len(norms) == 1, (M.tolist(), d, marg, norms)


# ---- the int_type = long long variant -------------------------------------

@pytest.mark.parametrize("n,marg", [(1, 0), (1, 1), (2, 0), (3, 0), (4, 0)])
def test_int64_scaling_beyond_int32(n, marg):
    """L(c*A) = c*L(A) for c > 0 (same maximising rank) with c*L(A) far
    beyond int32: a narrowing anywhere in the 64-bit walk breaks it."""
    rng = np.random.default_rng(40 + n + marg)
    R = {1: 9, 2: 8, 3: 6, 4: 5}[n]
    A = rng.integers(-50, 51, size=(R, 7)).astype(np.int64)
    norm, strat, rank = oracle.calc_lnorm64(A, n, marg, 3)
    assert (norm, strat, rank) == oracle.calc_lnorm(A.astype(np.int32), n, marg, 3)
    for c in (2 ** 24 + 3, 2 ** 33 - 1):
        n2, s2, r2 = oracle.calc_lnorm64(A * c, n, marg, 3)
        assert n2 == c * norm and n2 > 2 ** 31
        assert (s2, r2) == (strat, rank)
    if not marg:
        # a negative factor: |c| scaling, same maximising rank
        n3, s3, r3 = oracle.calc_lnorm64(A * -(2 ** 30 + 7), n, marg, 3)
        assert (n3, s3, r3) == ((2 ** 30 + 7) * norm, strat, rank)


def test_int64_rank_one_closed_form():
    """L1 of u v^T is ||u||_1 ||v||_1, here about 2^44."""
    rng = np.random.default_rng(7)
    u = rng.integers(-(2 ** 20), 2 ** 20, size=8).astype(np.int64)
    v = rng.integers(-(2 ** 20), 2 ** 20, size=5).astype(np.int64)
    M = np.outer(u, v)
    want = int(np.abs(u).sum()) * int(np.abs(v).sum())
    assert want > 2 ** 40
    assert oracle.calc_lnorm64(M, 1, 0, 4)[0] == want
    assert oracle.brute64(M, 1, 0)[0] == want


@pytest.mark.parametrize("n,marg", [(1, 0), (1, 1), (2, 0), (3, 0), (5, 0)])
def test_int64_walk_matches_bruteforce(n, marg):
    """The incremental 64-bit walks against direct decoding and full
    recompute, entries up to 2^40."""
    rng = np.random.default_rng(100 + n + marg)
    R = {1: 10, 2: 9, 3: 6, 5: 4}[n]
    for _ in range(3):
        A = rng.integers(-(2 ** 40), 2 ** 40, size=(R, int(rng.integers(1, 6)))).astype(np.int64)
        norm, strat, rank = oracle.calc_lnorm64(A, n, marg, 4)
        assert (norm, rank) == oracle.brute64(A, n, marg)
        assert oracle.objective64(A, n, marg, strat) == norm


def _block_moves(D, q, par_mask):
    """Restatement of the GPU d-ary walks' step order (walk_p16 / walk_dary /
    k_plain64_dary): blocks of d steps, a counted move of digit j >= 1 at
    step b*d, then the d-1 moves of row 0 unrolled in the direction of digit
    0's parity.  Yields (step, row, from, to)."""
    def ctz_nib(cnt):
        inv = ~cnt & ((1 << 64) - 1)
        return ((inv & -inv).bit_length() - 1) >> 2
    biaspat = 0
    for i in range(1, q):
        biaspat |= (16 - D) << (4 * (i - 1))
    cnt, pm, out, tt = biaspat, par_mask, [], 0
    nblocks = D ** (q - 1) if q > 0 else 1

    def run0(up):
        nonlocal tt
        for k in range(1, D):
            f, t = (k - 1, k) if up else (D - k, D - k - 1)
            tt += 1
            out.append((tt, 0, f, t))
    if q > 0:
        run0(not (pm & 1))
        for _ in range(1, nblocks):
            jj = ctz_nib(cnt)
            cnt = ((cnt + 1) | (biaspat & ((1 << (4 * jj)) - 1))) & ((1 << 64) - 1)
            v = ((cnt >> (4 * jj)) & 15) - (16 - D)
            j = jj + 1
            p = (pm >> j) & 1
            f = D - v if p else v - 1
            pm ^= (1 << j) - 1
            tt += 1
            out.append((tt, j, f, f - 1 if p else f + 1))
            run0(not (pm & 1))
    return out


@pytest.mark.parametrize("D", range(2, 11))
def test_dary_block_order(D):
    """The kernels' block-unrolled step order equals the reflected d-ary Gray
    code of the reference (lo_digit, L_OpenMP.c:140,204): for aligned
    segments of d^q ranks starting at seg*d^q, step t moves exactly the row
    whose label changes between ranks N0+t-1 and N0+t, between those labels.
    The initial parity mask is the kernels' (odd d: every digit below q
    follows the segment's parity; even d: only digit q-1)."""
    for q in range(0, 5 if D <= 4 else 3):
        for seg in (0, 1, 2, 3, 6):
            ps = seg & 1
            pmask = (((1 << q) - 1) if ps else 0) if D % 2 else ((1 << (q - 1)) if (q > 0 and ps) else 0)
            moves = _block_moves(D, q, pmask)
            N0 = seg * D ** q
            assert len(moves) == D ** q - 1
            rows = q + 3
            prev = [oracle.digit(N0, i, D) for i in range(rows)]
            for (tt, j, f, t) in moves:
                cur = [oracle.digit(N0 + tt, i, D) for i in range(rows)]
                diff = [i for i in range(rows) if cur[i] != prev[i]]
                assert diff == [j] and (prev[j], cur[j]) == (f, t), (D, q, seg, tt)
                prev = cur


def _low_label(D, M, l, i, odd):
    """walk_blk16's low labels: rank l's digit i, complemented for odd h
    (every low digit for odd d, the top one for even d)."""
    x = l // D ** i
    r = x % (2 * D)
    c = r if r < D else 2 * D - 1 - r
    return D - 1 - c if (odd and (D % 2 or i == M - 1)) else c


def _tabulated_walk(D, M, q, seg, R):
    """Restatement of walk_blk16's enumeration order for one segment of
    d^q ranks (q >= M): yields (offset, labels of rows 0..R-2)."""
    qh = q - M
    ps = seg & 1
    pm = (((1 << qh) - 1) if ps else 0) if D % 2 else ((1 << (qh - 1)) if (qh > 0 and ps) else 0)
    hpar0 = ps if (D % 2 or qh == 0) else 0
    high = []
    x = seg
    for i in range(R - 1 - M):
        if i < qh:
            high.append(D - 1 if (pm >> i) & 1 else 0)
        else:
            r = x % (2 * D)
            high.append(r if r < D else 2 * D - 1 - r)
            x //= D
    biaspat = 0
    for i in range(qh):
        biaspat |= (16 - D) << (4 * i)
    cnt = biaspat
    for k in range(D ** qh):
        if k > 0:
            inv = ~cnt & ((1 << 64) - 1)
            j = ((inv & -inv).bit_length() - 1) >> 2
            cnt = ((cnt + 1) | (biaspat & ((1 << (4 * j)) - 1))) & ((1 << 64) - 1)
            v = ((cnt >> (4 * j)) & 15) - (16 - D)
            p = (pm >> j) & 1
            f = D - v if p else v - 1
            pm ^= (1 << j) - 1
            assert high[j] == f
            high[j] = f - 1 if p else f + 1
        odd = (hpar0 ^ k) & 1
        for l in range(D ** M):
            yield k * D ** M + l, [_low_label(D, M, l, i, odd) for i in range(M)] + list(high)


@pytest.mark.parametrize("D,M", [(2, 5), (2, 4), (3, 5), (3, 4), (4, 4), (5, 3), (6, 3), (7, 2), (8, 2), (9, 2),
                                 (10, 2)])
def test_tabulated_walk_order(D, M):
    """The tabulated walk's split of the Gray rank N = h * d^M + l: the low
    labels are those of rank l, complemented when h is odd (all of them for
    odd d, the top one for even d), and the high rows follow the reflected
    code of h.  Every rank of several aligned segments is compared with the
    oracle's digits (L_OpenMP.c:140, 204)."""
    R = M + 4
    for q in sorted({M, M + 1, M + 2} if D ** (M + 2) <= 20000 else {M, M + 1}):
        for seg in (0, 1, 2, 3, 5):
            if (seg + 1) * D ** q > D ** (R - 1):
                continue
            N0 = seg * D ** q
            for off, labels in _tabulated_walk(D, M, q, seg, R):
                want = [oracle.digit(N0 + off, i, D) for i in range(R - 1)]
                assert labels == want, (D, M, q, seg, off)


@pytest.mark.parametrize("D,M", [(2, 5), (3, 5), (4, 4), (5, 3), (7, 2), (10, 2)])
def test_tabulated_objective(D, M):
    """sum_c F_c(S_c) with F_c(S) = sum_l |u_c,l + sum_{i in S} A_i,l| (u_c
    over the rows outside the table, the last row in label 0) equals the
    objective; the d >= 4 form TotU + sum over non-empty S_c of
    F_c(S_c) - F_c({}) too.  The subset table holds -A_S + 2^15 per 16-bit
    half, so |u + A_S| is one unsigned 16-bit absolute difference."""
    rng = np.random.default_rng(40 + D)
    R, C = M + 4, 7
    A = rng.integers(-90, 91, size=(R, C)).astype(np.int64)
    for _ in range(30):
        labels = list(rng.integers(0, D, size=R - 1)) + [0]
        u = np.zeros((D, C), np.int64)
        for i in range(M, R):
            u[labels[i]] += A[i]
        masks = [sum(1 << i for i in range(M) if labels[i] == c) for c in range(D)]

        def F(c, S):
            AS = sum((A[i] for i in range(M) if (S >> i) & 1), np.zeros(C, np.int64))
            neg = (0x8000 - AS) & 0xFFFF
            return int(np.abs((u[c] + 0x8000) - neg).sum())
        want = oracle.objective(A.astype(np.int32), D, 0, [int(v) for v in labels[:-1]])
        assert sum(F(c, masks[c]) for c in range(D)) == want
        totU = sum(F(c, 0) for c in range(D))
        assert totU + sum(F(c, masks[c]) - F(c, 0) for c in range(D) if masks[c]) == want


def _tab_bin_walk(M, q, seg, R):
    """Restatement of walk_tab16_bin's order for one segment of 2^q ranks:
    yields (offset, BRGC bits of rows 0..R-2).  Low bits of pattern l:
    l ^ (l >> 1), the top one xor (h & 1); a high step k flips row M + ctz(k)
    to 1 ^ (bit ctz(k) + 1 of h), bit qh of h being the segment's parity."""
    qh = q - M
    ps = seg & 1
    h0 = seg << qh
    g0 = h0 ^ (h0 >> 1)
    high = [(g0 >> j) & 1 for j in range(R - 1 - M)]
    for k in range(1 << qh):
        if k > 0:
            j = (k & -k).bit_length() - 1
            hb = (k >> (j + 1)) & 1 if j + 1 < qh else ps
            high[j] = 1 ^ hb
        hp = (k & 1) if qh > 0 else ps
        for l in range(1 << M):
            S = (l ^ (l >> 1)) ^ (hp << (M - 1))
            yield (k << M) + l, [(S >> i) & 1 for i in range(M)] + list(high)


def test_tabulated_binary_walk_order():
    """The L1 / LM tabulated walk visits ranks in the reference's BRGC order
    (L_OpenMP.c:58-69): every rank of aligned segments of 2^q ranks, q = 5..8,
    against the oracle's digits."""
    M, R = 5, 11
    for q in (5, 6, 8):
        for seg in (0, 1, 2, 3, 6, 7):
            if (seg + 1) << q > 1 << (R - 1):
                continue
            N0 = seg << q
            for off, bits in _tab_bin_walk(M, q, seg, R):
                want = [oracle.digit(N0 + off, i, 2) for i in range(R - 1)]
                assert bits == want, (q, seg, off)


@pytest.mark.parametrize("marg", [0, 1])
def test_tabulated_binary_objective(marg):
    """sum_l |u_l + A_S,l| over the |.| columns (u = A_{R-1} + sum of the
    high rows' signed rows, A_S the low rows' signed sum), plus u_0 + A_S,0
    for LM, equals the objective; the table word -A_S + 2^15 makes each |.| one
    unsigned 16-bit absolute difference."""
    rng = np.random.default_rng(77 + marg)
    M, R, C = 5, 9, 7
    A = rng.integers(-90, 91, size=(R, C)).astype(np.int64)
    for _ in range(40):
        bits = list(rng.integers(0, 2, size=R - 1))
        s = [2 * b - 1 for b in bits] + [1]
        u = sum(s[i] * A[i] for i in range(M, R))
        S = sum(bits[i] << i for i in range(M))
        AS = sum((2 * ((S >> i) & 1) - 1) * A[i] for i in range(M))
        cols = range(marg, C)
        neg = [(0x8000 - int(AS[l])) & 0xFFFF for l in cols]
        got = sum(abs((int(u[l]) + 0x8000) - n) for l, n in zip(cols, neg))
        if marg:
            got += int(u[0] + AS[0])
        assert got == oracle.objective(A.astype(np.int32), 1, marg, [int(b) for b in bits])
