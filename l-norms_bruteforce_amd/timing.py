"""Timing sweeps of the L_HIP command line and their statistics.

This replaces the reference's timing harness:
- time_order_1.sh:8-70 runs `L_CUDA B T file order [n]` many times for each
  (blocks, threads) pair of three grids, appends bash `time` output to
  meas_thread_T_block_B_order_O.dat and the `real` times to
  proc_thread_T_block_B_order_O.dat;
- calc_Statistics.m:1-55 reads the proc files ("%dm%f", :17) and writes, per
  grid, grid_<threads[0]*blocks[0]>_order_O.txt (:27-28) with one row per
  pair: blocks, threads, mean, and the 0/25/50/75/100 % quantiles (:24-25);
- runTimes.gnu:33-34 plots those columns as candlesticks (5:4:8:7).

Here one Python process does both steps, with the same file names and
columns, so runTimes.gnu plots the output unchanged.  L_HIP takes the CUDA
argv form `B T file d [p] [m]`; its engine autotunes, so B is ignored and T
sets the workgroup size of the plain walks when it is a multiple of 64.

Usage:
    python l-norms_bruteforce_amd/timing.py MATRIX [--order 1] [--iterations 50]
        [--grid 0,1,2] [--pairs 1024x8,512x16] [--extra n] [--outdir DIR] [--cli PATH]
"""
from __future__ import annotations

import argparse
import os
import resource
import subprocess
import sys
import time

import numpy as np

# (threads, blocks, extra argv) of time_order_1.sh:11-12, :30-31 and :53-54
# (the third grid runs with preprocessing off, :68)
GRIDS = (
    ((1024, 512, 256, 128, 64, 32, 16, 8), (8, 16, 32, 64, 128, 256, 512, 1024), ()),
    ((1024, 512, 256, 128, 64, 32, 16, 8), (16, 32, 64, 128, 256, 512, 1024, 2048), ()),
    ((512, 256, 128, 64, 32, 16, 8), (21, 42, 84, 168, 336, 672, 1344), ("n",)),
)
QUANTILES = (0.0, 0.25, 0.5, 0.75, 1.0)
PKG_DIR = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CLI = os.path.join(PKG_DIR, "bin", "L_HIP")


def quantiles(x) -> list:
    """Octave's and MATLAB's default `quantile` (calc_Statistics.m:24):
    piecewise linear through the points ((k - 0.5) / n, x_(k)) of the sorted
    sample, constant beyond the first and last of them."""
    return [float(v) for v in np.quantile(np.asarray(x, dtype=np.float64), QUANTILES, method="hazen")]


def format_time(seconds: float) -> str:
    """bash's `time` layout for real/user/sys: XmY.YYYs."""
    m = int(seconds // 60)
    return f"{m}m{seconds - 60 * m:.3f}s"


def parse_time(s: str) -> float:
    """calc_Statistics.m:17-20: textscan "%dm%f", minutes * 60 + seconds."""
    s = s.strip().rstrip("s")
    m, sec = s.split("m", 1)
    return int(m) * 60 + float(sec)


def meas_name(base: str, t: int, b: int, order: int) -> str:
    return f"{base}_thread_{t}_block_{b}_order_{order}.dat"


def proc_name(t: int, b: int, order: int) -> str:
    return f"proc_thread_{t}_block_{b}_order_{order}.dat"


def grid_name(threads, blocks, order: int) -> str:
    return f"grid_{threads[0] * blocks[0]}_order_{order}.txt"


def run_pair(cmd: list, cwd: str, iterations: int, meas_path: str, t: int, timeout: float = 3600) -> list:
    """Runs cmd `iterations` times.  Like time_order_1.sh:23-27: each run's
    output and its real/user/sys lines go to the meas file; returns the real
    (wall) times in seconds."""
    reals = []
    with open(meas_path, "a") as meas:
        for j in range(1, iterations + 1):
            meas.write(f"Thread count: {t}, Iteration: {j}\n")
            ru0 = resource.getrusage(resource.RUSAGE_CHILDREN)
            t0 = time.perf_counter()
            r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout)
            real = time.perf_counter() - t0
            ru1 = resource.getrusage(resource.RUSAGE_CHILDREN)
            meas.write(r.stdout.decode(errors="replace"))
            meas.write(f"\nreal\t{format_time(real)}\nuser\t{format_time(ru1.ru_utime - ru0.ru_utime)}\n"
                       f"sys\t{format_time(ru1.ru_stime - ru0.ru_stime)}\n")
            if r.returncode != 0:
                raise RuntimeError(f"{' '.join(cmd)} exited with {r.returncode}")
            reals.append(real)
    return reals


def read_proc(path: str) -> list:
    with open(path) as f:
        return [parse_time(line) for line in f if line.strip()]


def grid_stats(outdir: str, threads, blocks, order: int, log=print) -> np.ndarray:
    """calc_Statistics.m:6-26 for one grid: rows (blocks, threads, mean,
    q0, q25, q50, q75, q100); a missing proc file is reported and skipped."""
    rows = []
    for t, b in zip(threads, blocks):
        path = os.path.join(outdir, proc_name(t, b, order))
        if not os.path.exists(path):
            log(f"Error: Could not open file {path}")
            continue
        x = read_proc(path)
        rows.append([b, t, float(np.mean(x))] + quantiles(x))
    return np.array(rows, dtype=np.float64).reshape(-1, 3 + len(QUANTILES))


def save_ascii(path: str, rows: np.ndarray) -> None:
    """`save -ascii` layout (calc_Statistics.m:28): 8 significant digits,
    space-separated, one row per line."""
    with open(path, "w") as f:
        for r in rows:
            f.write("".join(f" {v:15.7e}" for v in r) + "\n")


def sweep(matrix: str, order: int = 1, iterations: int = 50, grids=GRIDS, outdir: str = ".",
          cli: str = DEFAULT_CLI, base: str = "meas", log=print) -> dict:
    """Runs every grid and writes its statistics file; returns {grid file:
    rows}."""
    if not os.path.exists(cli):
        raise FileNotFoundError(f"{cli} is not built")
    os.makedirs(outdir, exist_ok=True)
    matrix = os.path.abspath(matrix)
    out = {}
    for threads, blocks, extra in grids:
        if len(threads) != len(blocks):
            raise ValueError("The length of threads and blocks arrays must be the same.")
        for t, b in zip(threads, blocks):
            cmd = [cli, str(b), str(t), matrix, str(order)] + list(extra)
            reals = run_pair(cmd, outdir, iterations, os.path.join(outdir, meas_name(base, t, b, order)), t)
            with open(os.path.join(outdir, proc_name(t, b, order)), "a") as f:
                f.writelines(format_time(v) + "\n" for v in reals)
            log(f"threads {t:5d} blocks {b:5d}: median {np.median(reals):.4f} s over {len(reals)} runs")
        rows = grid_stats(outdir, threads, blocks, order, log)
        name = grid_name(threads, blocks, order)
        save_ascii(os.path.join(outdir, name), rows)
        out[name] = rows
    return out


def _pairs(text: str):
    threads, blocks = [], []
    for item in text.split(","):
        t, b = item.lower().split("x")
        threads.append(int(t))
        blocks.append(int(b))
    return tuple(threads), tuple(blocks)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("matrix")
    ap.add_argument("--order", type=int, default=1, help="d (the reference's `order`)")
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--grid", default="0,1,2", help="which of time_order_1.sh's three grids to run")
    ap.add_argument("--pairs", help="a grid of your own, THREADSxBLOCKS,... (replaces --grid)")
    ap.add_argument("--extra", default="", help="argv after the order for --pairs, e.g. 'n' or 'y m'")
    ap.add_argument("--outdir", default=".")
    ap.add_argument("--cli", default=DEFAULT_CLI)
    a = ap.parse_args(argv)
    if a.pairs:
        grids = (_pairs(a.pairs) + (tuple(a.extra.split()),),)
    else:
        grids = tuple(GRIDS[int(i)] for i in a.grid.split(","))
    sweep(a.matrix, a.order, a.iterations, grids, a.outdir, a.cli)
    return 0


if __name__ == "__main__":
    sys.exit(main())
