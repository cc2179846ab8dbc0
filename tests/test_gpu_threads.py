"""Concurrent calls into the C ABI from several host threads (ctypes drops
the GIL during the call): single calls on every kernel family, calls that
allocate nothing (one static scratch array per device), 64-bit calls, a
batch and calls that fail, interleaved in shuffled order over a thread
pool.  Every call must equal the same call made alone (norm, strategy, rank
and kernel), and each failing call must get its own error text (the
library keeps it per thread).  Calls on one device serialise on its
context.  GPU only."""
import random
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (rows, columns, d, marginal, kernel, entry range)
SINGLE = [
    (20, 40, 1, 0, "auto", 127), (18, 30, 1, 1, "auto", 127), (16, 300, 3, 0, "auto", 50),
    (14, 20, 3, 0, "plain16", 100), (20, 60, 1, 0, "table16", 100), (12, 200, 2, 0, "auto", 2000),
    (10, 700, 4, 0, "columns", 20000), (22, 64, 1, 0, "blocked", 127), (9, 30, 10, 0, "auto", 100),
    (12, 150, 3, 0, "table32", 20000), (16, 500, 1, 1, "auto", 100), (8, 40, 6, 0, "general", 100),
    (18, 60, 2, 0, "plain", 100), (20, 700, 1, 0, "wide", 100),
]


def _tasks():
    rng = np.random.default_rng(4242)
    tasks = []
    for R, C, n, m, kernel, r in SINGLE:
        A = rng.integers(-r, r + 1, size=(R, C)).astype(np.int32)
        tasks.append(("single", (A, n, m, kernel)))
    for R, C, n in ((20, 40, 1), (12, 200, 2)):   # through lnorm_hip_calc_noalloc
        tasks.append(("noalloc", (rng.integers(-100, 101, size=(R, C)).astype(np.int32), n, 0)))
    A64 = rng.integers(-100, 101, size=(16, 40)).astype(np.int64) * (2 ** 30 + 7)
    tasks.append(("int64", (A64, 1, 0)))
    probs = [(rng.integers(-50, 51, size=(int(rng.integers(4, 14)), int(rng.integers(2, 40)))).astype(np.int32),
              int(d), 0) for d in rng.choice([1, 2, 3], size=24)]
    tasks.append(("batch", (probs,)))
    # failing calls: d = 11, and a column sum beyond the int32 contract
    tasks.append(("error", (np.ones((4, 4), np.int32), 11, 0, "auto")))
    tasks.append(("error", (np.full((4, 4), 2 ** 30, np.int32), 1, 0, "auto")))
    return tasks


def _run(pkg, task):
    kind, args = task
    try:
        if kind == "single":
            A, n, m, kernel = args
            norm, strat, res = pkg.calc_lnorm(A, n, m, kernel=kernel)
            return ("ok", norm, tuple(strat), res["best_rank"], res["kernel_name"])
        if kind == "noalloc":
            A, n, m = args
            norm, strat, res = pkg.calc_lnorm(A, n, m, noalloc=True)
            return ("ok", norm, tuple(strat), res["best_rank"], res["kernel_name"])
        if kind == "int64":
            A, n, m = args
            norm, strat, res = pkg.calc_lnorm64(A, n, m)
            return ("ok", norm, tuple(strat), res["best_rank"], res["kernel_name"])
        if kind == "batch":
            out = pkg.calc_batch(args[0])
            return ("ok",) + tuple((int(l), tuple(s), int(st)) for l, s, st in out)
        A, n, m, kernel = args
        pkg.calc_lnorm(A, n, m, kernel=kernel)
        return ("no error",)
    except pkg.LnormError as e:
        return ("error", e.code, str(e))


def test_concurrent_calls_equal_sequential(pkg):
    tasks = _tasks()
    print("threads: sequential", len(tasks), flush=True)
    want = [_run(pkg, t) for t in tasks]
    assert [w[0] for w in want].count("error") == 2, [(i, tasks[i][0], SINGLE[i] if i < len(SINGLE) else None, w)
                                                      for i, w in enumerate(want) if w[0] == "error"]
    assert want[-2][2] != want[-1][2]
    order = [i for i in range(len(tasks)) for _ in range(3)]
    random.Random(7).shuffle(order)
    print("threads: concurrent", len(order), flush=True)
    with ThreadPoolExecutor(max_workers=6) as ex:
        got = list(ex.map(lambda i: (i, _run(pkg, tasks[i])), order))
    for i, g in got:
        assert g == want[i], (i, tasks[i][0], g, want[i])
