"""The C host library (reader, preprocessing, layout, work sizing, printer)
against the reference's golden stdout/strategy files and the Python host
oracle; CLI argument handling of L_HIP.  CPU only (norms come from the
oracle; the only L_HIP runs here never reach the GPU)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from oracle import host_oracle as ho


@pytest.fixture(scope="module")
def host(pkg):
    from lnorm_amd import host as h
    return h


def _args(case):
    argv = ["L", str(case["threads"]), "m.txt"] + case["args"]
    return ho.parse_args(argv, "omp")


def test_golden_stdout_and_strategy(host, golden, tmp_path):
    for case in golden:
        threads, _, _, n0, stat, marg = _args(case)
        p, out, exited = host.load(case["matrix"].encode(), n0, stat, marg, host.OMP, threads)
        if exited:
            assert case["returncode"] == 255 and out == case["stdout"], case["name"]
            host.free(p)
            continue
        if p.R == 0:
            out += "This is a zero matrix.\n"
            norm, strat = 0, []
        else:
            norm, strat, _ = oracle.calc_lnorm(p.matrix_laid_out(), p.n, p.marginal, threads)
        d = tmp_path / case["name"]
        d.mkdir()
        out += host.print_results(p, norm, strat, str(d))
        assert out == case["stdout"], case["name"]
        files = {f: open(os.path.join(d, f)).read() for f in os.listdir(d)}
        assert files == case["strategy"], case["name"]
        host.free(p)


def _text(M):
    return "".join(" ".join(str(int(v)) for v in r) + "\n" for r in M).encode()


def test_matches_host_oracle_on_random_inputs(host, tmp_path):
    rng = np.random.default_rng(23)
    for trial in range(300):
        R, C = int(rng.integers(1, 8)), int(rng.integers(1, 8))
        M = rng.integers(-3, 4, size=(R, C))
        if R > 2 and trial % 2:
            M[2] = M[0] * rng.choice([-2, -1, 1, 3])
        if C > 2 and trial % 3:
            M[:, 1] = M[:, 0] * rng.choice([-1, 2])
        if trial % 5 == 0:
            M[int(rng.integers(0, R))] = 0
        if trial % 7 == 0:
            M[:, int(rng.integers(0, C))] = 0
        d = int(rng.choice([1, 1, 2, 3, 4]))
        stat = str(rng.choice(list("nyrcYRCx")))
        marg = int(d == 1 and rng.integers(0, 2))
        flavor = int(rng.integers(0, 2))
        threads, blocks = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        text = _text(M)
        p, out, exited = host.load(text, d, stat, marg, flavor, threads, blocks)
        try:
            q = ho.load(text, d, stat, marg, "omp" if flavor == 0 else "cuda", threads, blocks)
            q_exit = False
        except ho.HostExit as e:
            q, q_exit = e, True
        assert exited == q_exit
        if exited:
            assert out.splitlines() == q.lines
            host.free(p)
            continue
        assert out.splitlines() == q.lines, (M.tolist(), d, stat, marg)
        assert (p.R, p.C, p.n, p.marginal, p.total, p.maxRows) == (q.R, q.C, q.n, q.marginal, q.total, q.maxRows)
        assert p.matrix_laid_out().flatten().tolist() == q.A
        strat = [int(v) for v in rng.integers(0, max(p.n, 2), size=max(p.R - 1, 0))]
        dd = tmp_path / str(trial)
        dd.mkdir()
        line = host.print_results(p, 17, strat, str(dd)).strip()
        qline, fname, qlines = ho.result_lines(q, 17, strat)
        assert line == qline
        assert open(dd / fname).read().splitlines() == qlines
        host.free(p)


def _run_cli(pkg, args, cwd):
    return subprocess.run([pkg.CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def test_cli_argument_errors(pkg, tmp_path):
    (tmp_path / "m.txt").write_text("1 2\n3 4\n")
    cases = [
        ([], "Incorrect number of input arguments."),
        (["0", "32", "m.txt", "1"], "number of blocks is a positive integer"),
        (["4", "x", "m.txt", "1"], "number of threads is a positive integer"),
        (["4", "32", "missing.txt", "1"], "file containig the matrix exists"),
        (["4", "32", "m.txt", "0"], "order of the L norm is a positive integer"),
        (["4", "32", "m.txt", "11"], "Please increase the RANK_OF_NORM variable in the code to 11"),
        (["4", "2048", "m.txt", "1"], "The maximum number of threads per block cannot be more than 1024."),
        (["--omp", "0", "m.txt", "1"], "number of threads is a positive integer"),
    ]
    for args, msg in cases:
        r = _run_cli(pkg, args, tmp_path)
        assert r.returncode == 255 and msg in r.stdout, (args, r.stdout, r.stderr)


def test_argv_restatement_matches_golden(golden_argv):
    # the Python restatement of arguments_OMP against the reference's own
    # output on every golden argv case (None where the run went on)
    for case in golden_argv:
        want = case["stdout"].rstrip("\n") if case["returncode"] != 0 else None
        assert ho.argv_error(["L_OpenMP"] + case["argv"], "omp") == want, case["name"]


def test_cli_argv_errors_match_reference(pkg, golden_argv, tmp_path):
    """L_HIP in both argv forms on the golden argv cases that exit before any
    computation: the L_OpenMP form prints the reference's bytes and exit
    status; the L_CUDA form (a block count before the thread count) prints
    arguments_CUDA's line (L_CUDA.cu:361-412, restated: no CUDA compiler
    here to build the reference)."""
    for case in golden_argv:
        if case["returncode"] == 0:
            continue
        d = tmp_path / case["name"]
        d.mkdir()
        (d / "matrix.txt").write_text(case["matrix"])
        r = _run_cli(pkg, ["--omp"] + case["argv"], d)
        assert (r.returncode, r.stdout) == (case["returncode"] & 0xFF, case["stdout"]), case["name"]
        cuda = ["8"] + case["argv"] if case["argv"] else []
        r = _run_cli(pkg, cuda, d)
        want = ho.argv_error(["L_CUDA"] + cuda, "cuda")
        assert (r.returncode, r.stdout) == (255, want + "\n"), case["name"]
        assert not [f for f in os.listdir(d) if f.startswith("strategy_")], case["name"]


def test_argv_restatement_cuda_checks():
    # arguments_CUDA's own checks: the block count first, the per-block
    # thread limit last (after the order), an EOF field keeps the value before it
    c = ho.argv_error
    assert c(["L_CUDA", "x", "32", "m", "1"], "cuda").startswith("Please make sure that the number of blocks")
    assert c(["L_CUDA", "4", "2048", "m", "11"], "cuda").startswith("The order of the L norm is too large")
    assert c(["L_CUDA", "4", "2048", "m", "1"], "cuda") == \
        "The maximum number of threads per block cannot be more than 1024."
    assert c(["L_CUDA", "4", "", "m", "1"], "cuda") is None          # threads keeps the block count
    assert c(["L_OpenMP", "", "m", "1"], "omp").startswith("Please make sure that the number of threads")
    assert c(["L_OpenMP", "3", "m", ""], "omp") is None             # the order keeps the thread count
    assert c(["L_OpenMP", "3", "m", "1"], "omp", file_exists=False).startswith(
        "Please make sure that the file containig the matrix exists")


def test_cli_host_only_paths(pkg, golden, tmp_path):
    # paths that never reach the GPU: the zero matrix and "Matrix is too large"
    for name in ("zero3x4_l1_y", "zero3x4_l2_y", "zero3x4_lm_y", "too_many_rows_l3"):
        case = next(c for c in golden if c["name"] == name)
        d = tmp_path / name
        d.mkdir()
        (d / "matrix.txt").write_text(case["matrix"])
        r = _run_cli(pkg, ["--omp", str(case["threads"]), "matrix.txt"] + case["args"], d)
        assert r.returncode == (case["returncode"] & 0xFF), (name, r.stderr)
        assert r.stdout == case["stdout"], name
        files = {f: (d / f).read_text() for f in os.listdir(d) if f.startswith("strategy_")}
        assert files == case["strategy"], name


def test_cuda_flavour_partition_line(host):
    # L_CUDA.cu:267-275: threads clamp to the work, blocks to work/threads
    for R, T, B in [(3, 32, 512), (12, 1024, 8), (20, 32, 512), (1, 64, 4)]:
        text = _text(np.arange(1, 3 * R + 1).reshape(R, 3) % 7 - 3)
        p, out, _ = host.load(text, 1, "n", 0, host.CUDA, T, B)
        q = ho.load(text, 1, "n", 0, "cuda", T, B)
        assert out.splitlines() == q.lines
        host.free(p)
