"""The timing harness (timing.py, the time_order_1.sh / calc_Statistics.m
replacement): quantile method, time-string parsing, file names and the
statistics table's columns.  CPU only: a stand-in command plays the CLI."""
import os
import stat

import numpy as np
import pytest


@pytest.fixture(scope="module")
def timing(pkg):
    from lnorm_amd import timing
    return timing


def test_quantiles_are_octaves_default(timing):
    # piecewise linear through ((k - 0.5) / n, x_(k)): numpy's default
    # ("linear", (k - 1) / (n - 1)) would give 1.75 / 3.25 for the first case
    assert timing.quantiles([4, 2, 3, 1]) == [1.0, 1.5, 2.5, 3.5, 4.0]
    assert timing.quantiles([3, 1, 2]) == [1.0, 1.25, 2.0, 2.75, 3.0]
    assert timing.quantiles([5.0]) == [5.0] * 5
    # constant beyond the first and last plotting positions
    x = [10, 20, 30, 40, 50, 60, 70, 80, 90, 100]
    q = timing.quantiles(x)
    assert q[0] == 10 and q[4] == 100 and q[2] == 55.0 and q[1] == 30.0 and q[3] == 80.0


def test_time_strings(timing):
    assert timing.parse_time("0m1.234s") == pytest.approx(1.234)
    assert timing.parse_time("2m3.500s\n") == pytest.approx(123.5)
    for v in (0.0, 0.0015, 59.9994, 61.25, 3725.125):
        assert timing.parse_time(timing.format_time(v)) == pytest.approx(v, abs=6e-4)
    assert timing.format_time(61.25) == "1m1.250s"


def _fake_cli(tmp_path, body='echo "L1 is: 2"'):
    p = tmp_path / "fake_cli"
    p.write_text("#!/bin/sh\n" + body + "\n")
    p.chmod(p.stat().st_mode | stat.S_IXUSR)
    return str(p)


def test_sweep_files_and_columns(timing, tmp_path):
    cli = _fake_cli(tmp_path)
    m = tmp_path / "m.txt"
    m.write_text("1 1\n1 -1\n")
    out = tmp_path / "out"
    grids = (((64, 32, 16), (8, 16, 32), ()), ((512,), (21,), ("n",)))
    res = timing.sweep(str(m), 1, 3, grids, str(out), cli, log=lambda s: None)
    assert sorted(res) == ["grid_10752_order_1.txt", "grid_512_order_1.txt"]
    for (t, b) in ((64, 8), (32, 16), (16, 32), (512, 21)):
        meas = (out / f"meas_thread_{t}_block_{b}_order_1.dat").read_text()
        assert meas.count("L1 is: 2") == 3 and meas.count("\nreal\t") == 3
        assert meas.startswith(f"Thread count: {t}, Iteration: 1\n")
        proc = (out / f"proc_thread_{t}_block_{b}_order_1.dat").read_text().split()
        assert len(proc) == 3 and all(s.endswith("s") and "m" in s for s in proc)
    rows = np.loadtxt(out / "grid_512_order_1.txt", ndmin=2)
    assert rows.shape == (3, 8)
    assert list(rows[:, 0]) == [8, 16, 32] and list(rows[:, 1]) == [64, 32, 16]
    # runTimes.gnu plots columns 5:4:8:7 = q25, q0, q100, q75
    assert np.all(rows[:, 3] <= rows[:, 4]) and np.all(rows[:, 4] <= rows[:, 5])
    assert np.all(rows[:, 5] <= rows[:, 6]) and np.all(rows[:, 6] <= rows[:, 7])
    assert np.all((rows[:, 2] >= rows[:, 3]) & (rows[:, 2] <= rows[:, 7]))
    # the statistics equal the proc file's values (3 decimals)
    x = [timing.parse_time(s) for s in (out / "proc_thread_32_block_16_order_1.dat").read_text().split()]
    assert rows[1, 2] == pytest.approx(np.mean(x), rel=1e-6)
    assert list(rows[1, 3:]) == pytest.approx(timing.quantiles(x), rel=1e-6)


def test_sweep_errors(timing, tmp_path):
    m = tmp_path / "m.txt"
    m.write_text("1\n")
    with pytest.raises(RuntimeError):
        timing.sweep(str(m), 1, 1, (((64,), (8,), ()),), str(tmp_path / "o1"), _fake_cli(tmp_path, "exit 3"),
                     log=lambda s: None)
    with pytest.raises(ValueError):
        timing.sweep(str(m), 1, 1, (((64, 32), (8,), ()),), str(tmp_path / "o2"), _fake_cli(tmp_path),
                     log=lambda s: None)
    msgs = []
    rows = timing.grid_stats(str(tmp_path / "o2"), (64,), (8,), 1, log=msgs.append)
    assert rows.shape == (0, 8) and msgs and msgs[0].startswith("Error: Could not open file")


def test_main_pairs(timing, tmp_path):
    m = tmp_path / "m.txt"
    m.write_text("1\n")
    cli = _fake_cli(tmp_path, 'echo "$@"')
    assert timing.main([str(m), "--pairs", "256x4,64x16", "--iterations", "2", "--order", "3", "--extra", "y m",
                        "--outdir", str(tmp_path / "o"), "--cli", cli]) == 0
    meas = (tmp_path / "o" / "meas_thread_256_block_4_order_3.dat").read_text()
    assert f"4 256 {m} 3 y m" in meas
    assert os.path.exists(tmp_path / "o" / "grid_1024_order_3.txt")
