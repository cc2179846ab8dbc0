"""End-to-end drop-in check: the L_HIP binary (C host + HIP engine through
the C ABI) reproduces the reference's stdout and strategy file byte for byte
on every golden case (L_OpenMP argv form), and L_CUDA's output format in its
own argv form."""
import os
import subprocess

import pytest

from oracle import host_oracle as ho

pytestmark = pytest.mark.gpu


def _run(pkg, args, cwd):
    return subprocess.run([pkg.CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def test_cli_matches_reference_byte_for_byte(pkg, golden, tmp_path):
    for case in golden:
        d = tmp_path / case["name"]
        d.mkdir()
        (d / "matrix.txt").write_text(case["matrix"])
        print("cli", case["name"], flush=True)
        r = _run(pkg, ["--omp", str(case["threads"]), "matrix.txt"] + case["args"], d)
        assert r.returncode == (case["returncode"] & 0xFF), (case["name"], r.stderr)
        assert r.stdout == case["stdout"], case["name"]
        files = {f: (d / f).read_text() for f in os.listdir(d) if f.startswith("strategy_")}
        assert files == case["strategy"], case["name"]


def test_cli_argv_golden_runs(pkg, golden_argv, tmp_path):
    # golden argv lines that the reference accepts (numeric prefixes such as
    # "2abc" and "4x"): the whole stdout and strategy file, byte for byte
    ran = 0
    for case in golden_argv:
        if case["returncode"] != 0:
            continue
        d = tmp_path / case["name"]
        d.mkdir()
        (d / "matrix.txt").write_text(case["matrix"])
        print("cli argv", case["name"], flush=True)
        r = _run(pkg, ["--omp"] + case["argv"], d)
        assert (r.returncode, r.stdout) == (0, case["stdout"]), (case["name"], r.stderr)
        files = {f: (d / f).read_text() for f in os.listdir(d) if f.startswith("strategy_")}
        assert files == case["strategy"], case["name"]
        ran += 1
    assert ran >= 2


def test_cli_cuda_argv_form(pkg, golden, tmp_path):
    for case in golden[::4]:
        if case["returncode"] != 0:
            continue
        d = tmp_path / case["name"]
        d.mkdir()
        (d / "matrix.txt").write_text(case["matrix"])
        B, T = 512, 32
        r = _run(pkg, [str(B), str(T), "matrix.txt"] + case["args"], d)
        assert r.returncode == 0, (case["name"], r.stderr)
        argv = ["L_CUDA", str(B), str(T), "matrix.txt"] + case["args"]
        _, blocks, _, n0, stat, marg = ho.parse_args(argv, "cuda")
        p = ho.load(case["matrix"].encode(), n0, stat, marg, "cuda", T, blocks)
        ref_lines = case["stdout"].splitlines()
        want = p.lines + ref_lines[len(p.lines):]
        assert r.stdout.splitlines() == want, case["name"]
        files = {f: (d / f).read_text() for f in os.listdir(d) if f.startswith("strategy_")}
        assert files == case["strategy"], case["name"]


def test_cli_marginal_workload_preprocessing_invariance(pkg, tmp_path):
    """The 36 x 36 marginal workload (entries in [-100, 100], seed 102) with 4
    duplicated rows, 3 duplicated columns and 2 further rows zeroed (40 x 39
    file): L_HIP prints the same LM with preprocessing ('y m', about 2^33
    sign vectors after the merges) and without it ('n m', 2^36).  Merging
    duplicates is exact for this objective (convex in each row's sign), so
    the two runs must agree; they share no enumeration."""
    import numpy as np
    from lnorm_amd import gen
    M = gen.gen_matrix(36, 36, 100, 102)
    rng = np.random.default_rng(102)
    dup_r = rng.choice(np.arange(1, 36), 4, replace=False)
    dup_c = rng.choice(np.arange(1, 36), 3, replace=False)
    V = np.vstack([M, M[dup_r]])
    V = np.hstack([V, V[:, dup_c]])
    zero = [r for r in range(1, 36) if r not in set(dup_r.tolist())][:2]
    V[zero, :] = 0
    assert V.shape == (40, 39)
    gen.write_matrix(tmp_path / "v.txt", V)
    norms = {}
    for flag in ("y", "n"):
        print("cli marginal workload", flag, flush=True)
        r = _run(pkg, ["--omp", "16", "v.txt", "1", flag, "m"], tmp_path)
        assert r.returncode == 0, r.stderr
        norms[flag] = int(r.stdout.strip().splitlines()[-1].split()[-1])
        assert r.stdout.strip().splitlines()[-1].startswith("LM is: ")
    assert norms["y"] == norms["n"], norms


def test_timing_sweep_on_cli(pkg, tmp_path):
    """The timing harness over the real L_HIP: every run of every (threads,
    blocks) pair prints the oracle's norm, and the statistics table has one
    row per pair."""
    import oracle
    from lnorm_amd import gen, timing
    A = gen.gen_matrix(20, 20, 10, 101)
    gen.write_matrix(str(tmp_path / "m.txt"), A)
    want = oracle.calc_lnorm(A, 1, 0, 8)[0]   # L1 is the same for A and its transpose
    print("timing sweep", flush=True)
    grids = (((256, 64), (4, 16), ()), ((128,), (16,), ("n",)))   # grid files 1024 and 2048
    res = timing.sweep(str(tmp_path / "m.txt"), 1, 2, grids, str(tmp_path / "o"), pkg.CLI, log=print)
    assert {k: r.shape for k, r in res.items()} == {"grid_1024_order_1.txt": (2, 8), "grid_2048_order_1.txt": (1, 8)}
    for (t, b) in ((256, 4), (64, 16), (128, 16)):
        meas = (tmp_path / "o" / f"meas_thread_{t}_block_{b}_order_1.dat").read_text()
        assert meas.count(f"L1 is: {want}\n") == 2, meas[-400:]
