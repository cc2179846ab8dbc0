"""The launch-shape autotuner (lnorm_hip_autotune, lnorm_hip_autotune_batch):
every candidate is timed, the cached choice is the fastest, and results with
autotune on are bit-identical to the default shapes.  GPU only."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _rate(e):
    return e["labelings_per_s"]


@pytest.mark.parametrize("case", ["blocked16", "blocked16_marg", "blocked32", "plain16_d3", "plain_L1", "table32_L1"])
def test_autotune_single(pkg, case):
    from lnorm_amd import gen
    print("autotune", case, flush=True)
    kw = {}
    if case == "blocked16":
        A, n, m = gen.gen_matrix(36, 40, 100, 7), 1, 0
    elif case == "blocked16_marg":
        A, n, m = gen.gen_matrix(34, 30, 100, 8), 1, 1
    elif case == "blocked32":
        A, n, m, kw = gen.gen_matrix(33, 40, 30000, 9), 1, 0, {"kernel": "blocked"}
    elif case == "table32_L1":
        A, n, m = gen.gen_matrix(33, 40, 30000, 9), 1, 0      # int16 entries: AUTO takes table32
    elif case == "plain16_d3":
        A, n, m = gen.gen_matrix(18, 32, 50, 103), 3, 0
    else:
        A, n, m, kw = gen.gen_matrix(30, 24, 50, 10), 1, 0, {"kernel": "plain"}
    best, entries = pkg.autotune(A, n, m, **kw)
    assert entries, case
    assert all(e["ms"] > 0 and _rate(e) > 0 for e in entries)
    top = max(entries, key=_rate)
    if case.startswith("blocked"):
        # the fastest shape, or the default one when it is within 2% of it
        assert best["kernel"] == pkg.KERNEL_BLOCKED
        chosen = [e for e in entries if (e["patterns_per_lane"], e["wave_log2"]) ==
                  (best["patterns_per_lane"], best["wave_log2"])]
        assert len(chosen) == 1 and _rate(chosen[0]) * 1.02 >= _rate(top)
        assert len({(e["patterns_per_lane"], e["wave_log2"]) for e in entries}) == len(entries) > 4
    else:
        assert best["block"] == top["block"] and len(entries) == 3
    ref = pkg.calc_lnorm(A, n, m, **kw)
    tuned = pkg.calc_lnorm(A, n, m, autotune=1, **kw)
    assert tuned[:2] == ref[:2] and tuned[2]["best_rank"] == ref[2]["best_rank"], case
    if case.startswith("blocked"):
        assert tuned[2]["patterns_per_lane"] == best["patterns_per_lane"]
    if A.shape[0] <= 20:
        assert ref[:2] == oracle.calc_lnorm(A, n, m, 8)[:2]


def test_autotune_batch(pkg):
    from lnorm_amd import gen, host
    print("autotune batch", flush=True)
    probs = []
    for w in gen.batch_workload(160):
        p, _, exited = host.load(w["text"], w["d"], w["stat"], w["marginal"])
        assert not exited
        probs.append((p.matrix_laid_out(), p.n, p.marginal))
        host.free(p)
    best, entries = pkg.autotune_batch(probs)
    assert len(entries) == 12
    top = min(entries, key=lambda e: e["ms"])
    assert (best["batch_streams"], best["batch_cap_log"]) == (top["batch_streams"], top["batch_cap_log"])
    ref = pkg.calc_batch(probs)
    assert pkg.calc_batch(probs, autotune=1) == ref
    assert pkg.calc_batch(probs, autotune=1, batch_graph=1) == ref


def test_autotune_cache_entry_from_a_taller_problem(pkg):
    """The cache is keyed by kernel shape, not by row count: a shape tuned on
    a tall matrix (P up to 8 needs 6 + 3 low rows and more high rows) must not
    make a shorter problem of the same width fail; it runs with the shape it
    would have had without tuning."""
    from lnorm_amd import gen
    tall = gen.gen_matrix(36, 36, 100, 21)
    best, _ = pkg.autotune(tall, 1, 1)
    assert best["kernel"] == pkg.KERNEL_BLOCKED
    for rows in (8, 11, 14, 20):
        A = tall[-rows:].copy()
        # the blocked kernel forced where it applies (AUTO takes the tabulated
        # walk below 2^33 sign vectors)
        kw = {"kernel": "blocked"} if rows >= 11 else {}
        ref = pkg.calc_lnorm(A, 1, 1, **kw)
        got = pkg.calc_lnorm(A, 1, 1, autotune=1, **kw)
        assert got[:2] == ref[:2] and got[2]["best_rank"] == ref[2]["best_rank"], rows
        assert got[:2] == oracle.calc_lnorm(A, 1, 1, 8)[:2]
