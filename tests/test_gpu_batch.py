"""Batched persistent launches (lnorm_hip_calc_batch_ex): every item's norm
and strategy equal the single-problem call's and the oracle's; statuses of
invalid items; HIP-graph capture and replay.  GPU only."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _rand(rng, R, C, r):
    return rng.integers(-r, r + 1, size=(R, C)).astype(np.int32)


def _workload_problems(items):
    from lnorm_amd import host
    out = []
    for w in items:
        p, _, exited = host.load(w["text"], w["d"], w["stat"], w["marginal"])
        assert not exited, w["seed"]
        out.append((p.matrix_laid_out(), p.n, p.marginal, p.total))
        host.free(p)
    return out


def test_batch_mixed_vs_oracle(pkg):
    print("batch mixed vs oracle", flush=True)
    rng = np.random.default_rng(31)
    probs = []
    for i in range(120):
        n = [1, 1, 1, 2, 3, 4, 5, 7, 10][i % 9]
        marg = 1 if (n == 1 and i % 3 == 0) else 0
        rmax = {1: 16, 2: 14, 3: 9, 4: 7, 5: 6, 7: 5, 10: 4}[n]
        R = int(rng.integers(1, rmax + 1))
        C = int(rng.integers(1, 70))
        probs.append((_rand(rng, R, C, 60), n, marg))
    # beyond the batched walks: 150 columns at n = 1, column sums > 32767 at n = 2
    probs.append((_rand(rng, 9, 150, 50), 1, 0))
    probs.append((_rand(rng, 8, 5, 9000), 2, 0))
    probs.append((_rand(rng, 6, 20, 20), 1, 1))
    out, st = pkg.calc_batch(probs, stats=True)
    assert st["items"] == len(probs) and st["fallback"] >= 2 and st["batched"] > 90, st
    for (A, n, marg), (norm, strat, status) in zip(probs, out):
        assert status == 0, (A.shape, n, marg)
        assert (norm, strat) == oracle.calc_lnorm(A, n, marg, 4)[:2], (A.shape, n, marg)


def test_batch_table_widths_vs_oracle(pkg):
    """The batched tabulated walks at every width (d = 2..4 and L1 / LM: 4..32
    words in steps of 2; d = 5..10: steps of 4, up to one wave's table
    width), at column counts on both sides of each width and with the fewest
    rows they tabulate."""
    print("batch table widths vs oracle", flush=True)
    rng = np.random.default_rng(47)
    rows = {1: (6, 14), 2: (6, 12), 3: (6, 8), 4: (5, 7), 5: (4, 6), 6: (4, 5), 7: (3, 5), 8: (3, 4), 9: (3, 4),
            10: (3, 4)}
    maxw = {1: 64, 2: 32, 3: 32, 4: 32, 5: 32, 6: 32, 7: 32, 8: 24, 9: 24, 10: 16}
    probs = []
    for d, (r0, r1) in rows.items():
        for C in (1, 7, 8, 9, 11, 16, 17, 21, 24, 25, 29, 33, 40, 41, 45, 48, 49, 53, 56, 57, 61, 64):
            if (C + 1) // 2 > maxw[d]:
                continue
            for marg in ((0, 1) if d == 1 else (0,)):
                if C <= marg:
                    continue
                R = int(rng.integers(r0, r1 + 1))
                probs.append((_rand(rng, R, C, 40), d, marg))
    out, st = pkg.calc_batch(probs, stats=True)
    assert st["fallback"] == 0, st
    for (A, n, marg), (norm, strat, status) in zip(probs, out):
        assert status == 0, (A.shape, n)
        assert (norm, strat) == oracle.calc_lnorm(A, n, marg, 8)[:2], (A.shape, n)


def test_batch_item_errors(pkg):
    print("batch item statuses", flush=True)
    good = np.arange(12, dtype=np.int32).reshape(4, 3) - 5
    probs = [(good, 1, 0), (good, 11, 0), (np.ones((64, 3), np.int32), 1, 0), (good, 2, 1),
             (np.full((3, 2), 2 ** 30, np.int32), 1, 0), (good, 3, 0)]
    out = pkg.calc_batch(probs)
    assert [s for _, _, s in out] == [0, -3, -3, -1, -2, 0]
    assert out[0][:2] == oracle.calc_lnorm(good, 1, 0, 2)[:2]
    assert out[5][:2] == oracle.calc_lnorm(good, 3, 0, 2)[:2]


def test_batch_equals_single_and_seg_log(pkg):
    from lnorm_amd import gen
    print("batch vs single, workload items 0..255", flush=True)
    probs = _workload_problems(gen.batch_workload(256))
    plain = [(A, n, m) for A, n, m, _ in probs]
    out, st = pkg.calc_batch(plain, stats=True)
    assert st["batched"] == len(plain), st
    forced = pkg.calc_batch(plain, seg_log=3)
    for (A, n, m, total), got, got3 in zip(probs, out, forced):
        norm, strat, res = pkg.calc_lnorm(A, n, m)
        assert got == (norm, strat, 0), (A.shape, n, m)
        assert got3 == got
        if total <= 1 << 18:
            assert (norm, strat) == oracle.calc_lnorm(A, n, m, 8)[:2]


def test_batch_graph_replay(pkg):
    print("batch graph capture and replay", flush=True)
    shapes = [(14, 30, 1, 0), (12, 20, 1, 1), (11, 17, 2, 0), (8, 25, 3, 0), (6, 9, 5, 0), (18, 64, 1, 0)]

    def batch(seed):
        rng = np.random.default_rng(seed)
        return [(_rand(rng, R, C, 100), n, m) for R, C, n, m in shapes]

    eager = {s: pkg.calc_batch(batch(s)) for s in (1, 2, 3)}
    replays = 0
    for rep, s in enumerate([1, 2, 3, 1, 2]):
        out, st = pkg.calc_batch(batch(s), stats=True, batch_graph=1)
        assert out == eager[s], (rep, s)
        replays += st["graph_replayed"]
        # eager work on the same stream between replays
        A, n, m = batch(s + 10)[0]
        assert pkg.calc_lnorm(A, n, m)[:2] == oracle.calc_lnorm(A, n, m, 4)[:2]
    assert replays == 4


@pytest.mark.timeout(400)
def test_batch_workload_4096(pkg):
    from lnorm_amd import gen
    print("batch workload: 4096 items", flush=True)
    probs = _workload_problems(gen.batch_workload(4096))
    plain = [(A, n, m) for A, n, m, _ in probs]
    out, st = pkg.calc_batch(plain, stats=True, batch_graph=1)
    assert st["batched"] == 4096 and st["fallback"] == 0, st
    print("batch stats", st, flush=True)
    checked_oracle = 0
    for k, ((A, n, m, total), got) in enumerate(zip(probs, out)):
        if k % 512 == 0:
            print("checking item", k, flush=True)
        assert got[2] == 0, k
        if total <= 1 << 16:
            assert got[:2] == oracle.calc_lnorm(A, n, m, 8)[:2], k
            checked_oracle += 1
        else:
            norm, strat, _ = pkg.calc_lnorm(A, n, m)
            assert got[:2] == (norm, strat), k
    assert checked_oracle > 100


@pytest.mark.timeout(300)
def test_batch_workload_uncapped_d3_items(pkg):
    """The mixed batch workload without the 20-row cap on d = 3 files: the
    d = 3 items with 22 or 23 rows after preprocessing (3^21 and 3^22
    labelings, beyond the persistent launches' 2^32, so the batch runs them
    on the single-problem kernels).  The batch's result equals the
    tabulated walk's and the packed walk's (two different kernels: norm,
    strategy and rank), and the strategy's objective is the norm."""
    from lnorm_amd import gen
    print("batch workload without the d = 3 row cap: items with 22-23 rows", flush=True)
    work = [w for w in gen.batch_workload(4096, d3_max_rows=30) if w["d"] == 3]
    probs = [(A, n, m) for A, n, m, total in _workload_problems(work) if A.shape[0] in (22, 23)]
    assert len(probs) >= 40
    out, st = pkg.calc_batch(probs, stats=True)
    assert st["fallback"] == len(probs) and st["batched"] == 0, st
    for k, ((A, n, m), got) in enumerate(zip(probs, out)):
        assert got[2] == 0, k
        t_norm, t_strat, t_res = pkg.calc_lnorm(A, n, m, kernel="table16")
        p_norm, p_strat, p_res = pkg.calc_lnorm(A, n, m, kernel="plain16")
        assert (t_norm, t_strat, t_res["best_rank"]) == (p_norm, p_strat, p_res["best_rank"]), (k, A.shape)
        assert got[:2] == (t_norm, t_strat), (k, A.shape)
        assert oracle.objective(A, n, m, t_strat) == t_norm, (k, A.shape)
