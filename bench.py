#!/usr/bin/env python3
"""Benchmark: exact L_1 sign-vectors/s on one MI355X.

Workload (one step = one full computation: enumeration of all 2^(k-1) sign
vectors, the maximum, and the strategy of the smallest maximising Gray rank):
  64 x 40 int8 matrix, entries uniform in [-127, 127] (seed 104).  L_1 is
  transpose-invariant and the reference lays it out as 40 x 64 (k = 40 rows
  enumerated, 64 partial sums): 2^39 sign vectors per step.
The blocked kernel (int8-MFMA-built low-row table + scalar Gray walk over the
high rows, v_sad_u16) is the headline; the plain Gray kernel runs the same
steps beside it, as do a 40 x 40 matrix (entries in [-100, 100], seed 106)
and the reference L_OpenMP on the host cores over a bounded sample.

Prints one JSON line.  --gpus above 1 is refused: the bench runs on one GPU.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "l-norms_bruteforce_amd")

METRIC = "exact L_1 sign-vectors/s on 1 MI355X (int8-MFMA vs plain Gray), bit-exact"
VALU_PEAK = 256 * 4 * 16 * 2.4e9        # int32 lane-ops/s: 256 CUs x 4 SIMD16 x 2.4 GHz
# packed 16-bit |.|-accumulates: one v_sad_u16 lane-op retires two row-updates
# (two biased 16-bit halves), so the row-update roof of the 16-bit paths is twice the int32 one
VALU16_PEAK = 2 * VALU_PEAK
MFMA_I8_PEAK = 5.0e15                   # dense int8 ops/s
# HBM bytes per launch of the headline instantiation, measured with
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE on `bench.py --case blocked`,
# each scaled by the counter's reading for a 1 GiB device-to-device copy (x2.0
# for FETCH_SIZE, x1.0 for WRITE_SIZE).  Each launch leaves one 64-bit key
# (its maximum and first wave); the matrix and the pattern table stay on chip.
#   P=2, s=11 (autotuned, one launch of 2^21 waves per call): FETCH_SIZE
#   136.94 KiB (x2.0), WRITE_SIZE 0.97 KiB (x1.0)
PROFILED_TRAFFIC = {"blocked16<np=32,P=2,marg=0,s=11>": 280448 + 992}
# VALU issue rate of the headline instantiation from the counters:
# SQ_INSTS_VALU x 64 / dispatch time / 39.32e12 (rocprofv3 --pmc on
# `bench.py --case blocked`: 2.79452e11 instructions in 464.14 ms)
PROFILED_ISSUE = {"blocked16<np=32,P=2,marg=0,s=11>": 0.980}


def load_pkg():
    if "lnorm_amd" in sys.modules:
        return sys.modules["lnorm_amd"]
    spec = importlib.util.spec_from_file_location(
        "lnorm_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lnorm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def valu_ops_per_unit(kernel: str, ny: int, marginal: int, P: int = 2) -> float:
    """Minimal VALU lane-operations per sign vector of each design:
    plain:   one add of the row delta and one |.|-accumulate (v_sad_u32) per
             row, plus the running maximum;
    blocked: one v_sad_u16 per two |.| columns (the high-row update runs on
             the scalar unit), plus the marginal add, and the running maximum
             shared by two patterns (one v_max3_i32; P = 1: one v_max)."""
    if kernel == "blocked":
        return (ny - marginal + 1) // 2 + marginal + (1.0 if P == 1 else 0.5)
    if kernel == "blocked32":        # one v_sad_u32 per |.| column
        return ny - marginal + 1 + marginal
    if kernel == "table16":          # tabulated walk: one v_sad_u16 per two |.| columns, half a v_max3,
        words = (ny - marginal + 1) // 2    # the marginal add, and the high step's packed add per 32
        return words + 0.5 + marginal + words / 32.0
    if kernel == "table16w":         # over W waves: the padded words' v_sad_u16, the high step per 2^M
        words = (ny - marginal + 1) // 2    # patterns, the W partials' v_add3 and the running maximum
        nwt = 256 if words <= 256 else 512 if words <= 512 else 1024
        w, m = (8, 5) if nwt == 256 else (16, 4)
        return nwt + nwt / 2 ** m + w / 2 + 1 + marginal
    if kernel == "table32":          # 32-bit words: one v_sad_u32 per padded |.| column, half a v_max3,
        words = next(w for w in TAB32_WORDS if w >= ny - marginal)   # the high step per 2^M patterns
        return words + 0.5 + marginal + words / 2 ** (5 if words <= 128 else 4)
    if kernel == "table32w":         # over W waves of 32 (64 at 1024) words, 2^M patterns per block;
        words = next(w for w in (256, 320, 384, 512, 1024) if w >= ny - marginal)   # at 64 words per wave
        w, m = {256: (8, 5), 320: (10, 4), 384: (12, 4), 512: (16, 4), 1024: (16, 4)}[words]   # two |.|
        return words + words / 2 ** m + w / 2 + 1 + marginal + (1 if words == 1024 else 0)   # chains joined
    return 2 * ny + 1


# instantiated widths (words = |.| columns) of the 32-bit tabulated L1 walk in one wave
TAB32_WORDS = (8, 16, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192)


AUTOTUNE = True    # bench.py --no-autotune clears it


_T0 = time.perf_counter()


def progress(what: str) -> None:
    """One line on stderr as each workload of the default run ends (stdout
    keeps the single JSON line)."""
    print("[bench %6.1f s] %s" % (time.perf_counter() - _T0, what), file=sys.stderr, flush=True)


def run_steps(pkg, A, kernel, steps, warmup, P=0, n=1):
    """`steps` timed calls (norm + strategy) after `warmup` untimed ones.
    With autotuning, the launch shape is tuned once before any of them
    (lnorm_hip_autotune, cached) and every call uses the cached shape."""
    res = None
    tuned = None
    if AUTOTUNE:
        best, _ = pkg.autotune(A, n, 0, kernel=kernel)
        tuned = {k: best[k] for k in ("patterns_per_lane", "wave_log2", "block")}
    kw = {"kernel": kernel, "patterns_per_lane": P, "autotune": 1 if AUTOTUNE else 0}
    for _ in range(warmup):
        pkg.calc_lnorm(A, n, 0, **kw)
    times, kms = [], []
    out = None
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        norm, strat, res = pkg.calc_lnorm(A, n, 0, **kw)
        times.append(time.perf_counter() - t)
        kms.append(res["ms_kernel"])
        if out is None:
            out = (norm, strat, res["best_rank"])
        elif out != (norm, strat, res["best_rank"]):
            raise RuntimeError("result changed between steps")
    wall = time.perf_counter() - t0
    return {"norm": out[0], "best_rank": out[2], "strategy": out[1], "wall_s": wall, "step_s": times,
            "kernel_ms": kms, "res": res, "tuned": tuned}


def alg_frac(kernel: str, row_updates_per_s: float) -> float:
    """Row-update rate over its roof by the algorithmic rule: 1 op per
    row-update on the blocked and tabulated paths (the update itself is
    tabulated or on the scalar unit), 2 on the plain walk (add and |.|), each
    against the VALU rate of its word size: 16-bit halves (v_sad_u16, two
    row-updates per lane-op) or 32-bit words."""
    if kernel in ("blocked", "table16", "table16w"):
        return row_updates_per_s / VALU16_PEAK
    if kernel == "plain":
        return 2 * row_updates_per_s / VALU_PEAK
    return row_updates_per_s / VALU_PEAK


def summarize(r, total, ny, kernel, marginal=0):
    steps = len(r["step_s"])
    kms = sum(r["kernel_ms"]) / steps
    P = int(r["res"].get("patterns_per_lane") or 2)
    ops = valu_ops_per_unit(kernel, ny, marginal, P) * total
    achieved = ops / (kms / 1e3)
    return {
        "alg_frac": alg_frac(kernel, total * ny / (kms / 1e3)),
        "value": total * steps / r["wall_s"],
        "ms_per_step": 1e3 * r["wall_s"] / steps,
        "kernel_ms": kms,
        "kernel_sign_vectors_per_s": total / (kms / 1e3),
        "row_updates_per_s": total * ny / (kms / 1e3),
        "valu_ops_per_sign_vector": valu_ops_per_unit(kernel, ny, marginal, P),
        "valu_TOPs": achieved / 1e12,
        "valu_frac": achieved / VALU_PEAK,
        "kernel_name": r["res"]["kernel_name"],
        "norm": r["norm"],
        "best_rank": r["best_rank"],
    }


def mfma_share(res, k: int, ny: int) -> dict:
    """int8 MFMA work of one blocked call: the low-row table of 2^kl patterns
    (kl = 6 + log2 P) over the |.| columns, built with 32x32x32 int8 tiles
    (K = the kl rows padded to 32; one tile pass per int8 plane).  It is built
    once per call; the enumeration itself runs on the VALU and SALU, so the
    MFMA fraction of peak is tiny by design."""
    P = max(1, int(res["patterns_per_lane"]))
    kl = 6 + P.bit_length() - 1
    cols = -(-ny // 32) * 32
    planes = 2 if "blocked32" in res["kernel_name"] else 1
    ops = 2 * (2 ** kl) * 32 * cols * planes
    secs = res["ms_kernel"] / 1e3
    return {"ops_per_call": ops, "frac": ops / secs / MFMA_I8_PEAK, "peak_TOPs": MFMA_I8_PEAK / 1e12,
            "note": "table build only; the walk over the %d high rows is VALU (v_sad_u16) + SALU" % (k - 1 - kl)}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_reference(gen, M, d: int, units: int, extra=(), runs: int = 3) -> dict:
    """The reference L_OpenMP (built from its sources into oracle/_ref) on the
    host cores, threads bound one per core (OMP_PLACES=cores,
    OMP_PROC_BIND=spread): `runs` runs of the file M with argv tail
    `d [extra]`, the median wall time (the time_order_1.sh method: whole
    process).  Falls back
    to the oracle's restatement when the reference binary is absent."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    ref = os.path.join(ROOT, "oracle", "_ref", "L_OpenMP")
    times, norm = [], None
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as td:
            gen.write_matrix(os.path.join(td, "m.txt"), M)
            for _ in range(runs):
                t = time.perf_counter()
                env = dict(os.environ, OMP_PLACES="cores", OMP_PROC_BIND="spread")
                p = subprocess.run([ref, str(threads), "m.txt", str(d), *extra], cwd=td, capture_output=True,
                                   text=True, timeout=600, env=env)
                times.append(time.perf_counter() - t)
                if p.returncode != 0:
                    raise RuntimeError("reference L_OpenMP failed: %s" % p.stderr)
                norm = int(p.stdout.strip().splitlines()[-1].split()[-1])
        kind = "reference"
    else:
        sys.path.insert(0, ROOT)
        import oracle
        from lnorm_amd import host
        pr, _, _ = host.load(gen.matrix_text(M).encode(), d, "n", 0)
        A = pr.matrix_laid_out()
        host.free(pr)
        for _ in range(runs):
            t = time.perf_counter()
            norm = oracle.calc_lnorm(A, d, 0, threads)[0]
            times.append(time.perf_counter() - t)
        kind = "port"
    dt = sorted(times)[len(times) // 2]
    return {"value": units / dt, "cores": threads, "kind": kind, "seconds": dt, "runs": len(times),
            "cpu": cpu_model(), "norm": norm}


def cpu_baseline(gen, rows: int = 64, r: int = 127, seed: int = 104, budget_cols: int = 34):
    """The headline workload's CPU baseline: the same entry distribution with
    `budget_cols` columns, 2^(budget_cols-1) sign vectors of `rows` partial
    sums (the reference transposes the file to budget_cols x rows)."""
    M = gen.gen_matrix(rows, budget_cols, r, seed)
    c = cpu_reference(gen, M, 1, 2 ** (budget_cols - 1))
    sample = ("%d x %d file, entries in [-%d, %d] (k = %d after the reference's transpose, %d partial sums): "
              "2^%d sign vectors; median of %d runs %.2f s; L1 = %d; %s"
              % (rows, budget_cols, r, r, budget_cols, rows, budget_cols - 1, c["runs"], c["seconds"], c["norm"],
                 c["cpu"]))
    return {"value": c["value"], "unit": "sign-vectors/s", "cores": c["cores"], "kind": c["kind"],
            "sample": sample, "seconds": c["seconds"]}


def single_line(pkg, A, n: int, total: int, ops, int64: bool = False, kernel: str = "auto") -> dict:
    """One call (after a warm-up call): labelings/s, the kernel that ran and,
    with an op count per unit, the VALU fraction by that model.  Used for the
    wide L1 matrices (36 x 256, 2^35 sign vectors; 34 x 1000, 2^33) on AUTO's
    tabulated walks and on the binary W-wave walk (2 ops per row-update), and
    for the 64-bit walk (L1 of
    a 34 x 40 matrix scaled by 2^28 + 3, beyond int32: 2^33 sign vectors)."""
    calc = pkg.calc_lnorm64 if int64 else pkg.calc_lnorm
    calc(A, n, 0, kernel=kernel)
    t = time.perf_counter()
    norm, strat, res = calc(A, n, 0, kernel=kernel)
    wall = time.perf_counter() - t
    kms = res["ms_kernel"]
    out = {"value": total / wall, "unit": "sign-vectors/s", "kernel_ms": kms, "kernel": res["kernel_name"],
           "kernel_sign_vectors_per_s": total / (kms / 1e3), "norm": norm}
    if ops:
        out["valu_frac"] = total * ops / (kms / 1e3) / VALU_PEAK
    return out


def cli_line(pkg, gen) -> dict:
    """The drop-in CLI end to end on the headline file (64 x 40, seed 104):
    `L_HIP --omp 16 m.txt 1`, a fresh process each time: reading,
    preprocessing, device initialisation, the 2^39 enumeration, the
    strategy and the printed output (the time_order_1.sh method)."""
    with tempfile.TemporaryDirectory() as td:
        gen.write_matrix(os.path.join(td, "m.txt"), gen.gen_matrix(64, 40, 127, 104))
        times = []
        for _ in range(2):
            t = time.perf_counter()
            p = subprocess.run([pkg.CLI, "--omp", "16", "m.txt", "1"], cwd=td, capture_output=True, text=True,
                               timeout=300)
            times.append(time.perf_counter() - t)
            if p.returncode != 0:
                raise RuntimeError("L_HIP failed: %s" % p.stderr)
        last = p.stdout.strip().splitlines()[-1]
    return {"wall_s": min(times), "sign_vectors_per_s": 2 ** 39 / min(times), "printed": last}


def l3_line(pkg, gen, steps: int, warmup: int, cpu: bool) -> dict:
    """L_3 of a 24 x 32 file (entries in [-50, 50], seed 103): 3^23 = 9.4e10
    labelings on the tabulated packed walk, beside the packed walk without
    tables, the 32-bit d-ary walk and the reference L_OpenMP on a 20-row
    sample of the same file (3^19)."""
    B = gen.gen_matrix(24, 32, 50, 103)
    total = 3 ** 23
    kw = {"autotune": 1 if AUTOTUNE else 0}
    if AUTOTUNE:
        pkg.autotune(B, 3, 0)
    for _ in range(warmup):
        pkg.calc_lnorm(B, 3, 0, **kw)
    t0 = time.perf_counter()
    kms, out = [], None
    for _ in range(max(1, steps)):
        norm, strat, res = pkg.calc_lnorm(B, 3, 0, **kw)
        kms.append(res["ms_kernel"])
        if out is None:
            out = (norm, strat, res["best_rank"])
        elif out != (norm, strat, res["best_rank"]):
            raise RuntimeError("result changed between steps")
    wall = (time.perf_counter() - t0) / max(1, steps)
    km = sum(kms) / len(kms)
    n2, s2, r2 = pkg.calc_lnorm(B, 3, 0, kernel="plain")
    n3, s3, r3 = pkg.calc_lnorm(B, 3, 0, kernel="plain16")
    if (n2, s2, r2["best_rank"]) != out or (n3, s3, r3["best_rank"]) != out:
        raise RuntimeError("the tabulated, packed and 32-bit d-ary walks disagree")
    # op model of the tabulated walk at 16 words (M = 5): 1.5 per labeling
    # for the sums and the maximum, plus (2 + 64) * 16 per block of 243
    ops = table16_ops(3, 16)
    line = {"value": total / wall, "unit": "labelings/s", "ms_per_step": 1e3 * wall, "kernel_ms": km,
            "kernel": res["kernel_name"], "valu_ops_per_labeling_model": ops,
            "valu_frac_model": total * ops / (km / 1e3) / VALU_PEAK,
            "plain16": {"value": total / (r3["ms_kernel"] / 1e3), "kernel": r3["kernel_name"]},
            "plain32": {"value": total / (r2["ms_total"] / 1e3), "kernel": r2["kernel_name"]},
            "norm": out[0]}
    if cpu:
        c = cpu_reference(gen, B[:20], 3, 3 ** 19)
        line["cpu_baseline"] = {"value": c["value"], "cores": c["cores"], "kind": c["kind"],
                                "sample": "first 20 rows of the file: 3^19 labelings; median of %d runs %.2f s; "
                                          "L3 = %d" % (c["runs"], c["seconds"], c["norm"])}
        line["speedup_vs_cpu"] = line["value"] / c["value"]
    return line


# the tabulated walk's width limit per d (16-bit words; batched items: 32)
# and the rows it keeps in its subset tables at a width
TABLE_MAX_WORDS = {2: 32, 3: 32, 4: 32, 5: 32, 6: 32, 7: 32, 8: 24, 9: 24, 10: 16}


def table16_rows(d: int, words: int) -> int:
    if d == 3 and words <= 16:
        return 6
    if d <= 3:
        return 5 if words <= 32 else 4
    return {4: 4 if words <= 32 else 3 if words <= 40 else 2, 5: 4 if words <= 16 else 3 if words <= 32 else 2,
            6: 4 if words <= 8 else 3 if words <= 24 else 2, 7: 3 if words <= 16 else 2 if words <= 24 else 1,
            8: 3 if words <= 16 else 2, 9: 3 if words <= 16 else 1, 10: 3 if words <= 8 else 2}[d]


def table16_ops(d: int, words: int) -> float:
    """VALU lane-operations per labeling of the tabulated walk: per block of
    d^M labelings, the high move (a v_pk_sub and a v_pk_add per word) and the
    two changed labels' subset tables (2^M v_sad_u16 per word each), plus per
    labeling one sum of table entries and half a v_max3 (d = 2, 3: one add,
    1.5 in all; d >= 4 the compiler shares partial sums between labelings,
    and 1.5 is an upper bound of what it issues)."""
    M = table16_rows(d, words)
    return 1.5 + (2 + 2 ** (M + 1)) * words / d ** M


def batch_valu_ops(d: int, rows: int, cols: int, marginal: int = 0) -> float:
    """VALU lane-operations per labeling of the batched walks by the op model,
    on the item's real width: L1/LM one add and one |.|-accumulate per column
    plus the maximum; d >= 2 with at least M + 1 rows the tabulated walk
    (table16_ops); other d >= 2 items the packed 16-bit walk: a v_pk_sub, a
    v_pk_add and two v_sad_u16 per two columns (the two label vectors that
    exchange a row), d - 1 adds of the label sums and the maximum (an upper
    bound: the unrolled row-0 runs need fewer)."""
    if d == 1:
        if rows - 1 >= 5:           # tabulated L1 / LM walk (16-bit sums: entries <= 100, <= 30 rows)
            return valu_ops_per_unit("table16", cols, marginal)
        return 2 * cols + 1
    words = (cols + 1) // 2
    if words <= TABLE_MAX_WORDS[d] and rows - 1 >= table16_rows(d, words):
        return table16_ops(d, words)
    return 4 * words + d


def batch_line(pkg, gen, steps: int) -> dict:
    """The mixed batch workload (4096 seeded files, m x n with m in [12, 30]
    rows, n in [8, 64] columns, d = 1/2/3 at 60/25/15 %, a quarter with
    duplicated rows/columns and 'y', a tenth of d = 1 with 'm'; d = 3 files
    keep at most 20 rows) through the host library's reader and
    preprocessing, then one lnorm_hip_calc_batch_ex call per step: one
    persistent launch per kernel instantiation, replayed as a HIP graph.
    Every replay must equal the eager (uncaptured) run, item for item."""
    from lnorm_amd import host
    work = gen.batch_workload(4096)
    probs, per_d, ops, row_updates = [], {}, 0, 0
    for w in work:
        p, _, exited = host.load(w["text"], w["d"], w["stat"], w["marginal"])
        if exited:
            raise RuntimeError("batch item %d rejected by the host pipeline" % w["seed"])
        probs.append((p.matrix_laid_out(), p.n, p.marginal))
        per_d[p.n] = per_d.get(p.n, 0) + p.total
        row_updates += p.total * p.C
        ops += p.total * batch_valu_ops(p.n, p.R, p.C, p.marginal)
        host.free(p)
    eager, st0 = pkg.calc_batch(probs, stats=True)
    if any(s != 0 for _, _, s in eager):
        raise RuntimeError("batch item failed")
    tuned = None
    if AUTOTUNE:
        best, _ = pkg.autotune_batch(probs)
        tuned = {k: best[k] for k in ("batch_streams", "batch_cap_log")}
    times, dev, call = [], [], []
    for _ in range(max(1, steps)):
        t = time.perf_counter()
        out, st = pkg.calc_batch(probs, stats=True, batch_graph=1, autotune=1 if AUTOTUNE else 0)
        times.append(time.perf_counter() - t)
        dev.append(st["ms_device"])
        call.append(st["ms_total"])
        if out != eager:
            raise RuntimeError("graph replay differs from the eager batch")
    lab = st["labelings"]
    wall = sum(times) / len(times)
    msd = sum(dev) / len(dev)
    return {"items": len(probs), "groups": st["groups"], "batched": st["batched"], "labelings": lab,
            "labelings_by_d": {str(k): v for k, v in sorted(per_d.items())},
            "value": lab / wall, "unit": "labelings/s", "ms_per_step": 1e3 * wall, "ms_device": msd,
            "ms_library_call": sum(call) / len(call),
            "device_labelings_per_s": lab / (msd / 1e3), "eager_ms_device": st0["ms_device"],
            "row_updates_per_s": row_updates / (msd / 1e3),
            "valu_TOPs_model": ops / (msd / 1e3) / 1e12, "valu_frac_model": ops / (msd / 1e3) / VALU_PEAK,
            "graph_replayed": st["graph_replayed"], "autotuned": tuned}


def cfg2_dup_matrix(gen):
    """The marginal workload's variant: the 36 x 36 matrix (entries in
    [-100, 100], seed 102) with 4 rows and 3 columns duplicated (sources
    among indices 1..35) and 2 further non-marginal rows zeroed: 40 x 39."""
    import numpy as np
    M = gen.gen_matrix(36, 36, 100, 102)
    rng = np.random.default_rng(102)
    dup_r = rng.choice(np.arange(1, 36), 4, replace=False)
    dup_c = rng.choice(np.arange(1, 36), 3, replace=False)
    V = np.vstack([M, M[dup_r]])
    V = np.hstack([V, V[:, dup_c]])
    zero = [r for r in range(1, 36) if r not in set(dup_r.tolist())][:2]
    V[zero, :] = 0
    return V


def configs_line(pkg, gen) -> dict:
    """The reference's configured workloads one by one, each file through
    the host library (reader, preprocessing, layout, parameters) and one
    timed call on AUTO's kernel after a warm-up call: L1 20 x 20 (2^19),
    LM 36 x 36 and its duplicated-rows variant with 'y' and without, L2 and
    L3 of the 18 x 32 file (2^17, 3^17).  The two marginal runs of the
    variant must print the same norm."""
    from lnorm_amd import host
    cases = [("cfg1_L1_20x20", gen.gen_matrix(20, 20, 10, 101), 1, "n", 0),
             ("cfg2_LM_36x36", gen.gen_matrix(36, 36, 100, 102), 1, "n", 1),
             ("cfg2_LM_40x39_dups_y", cfg2_dup_matrix(gen), 1, "y", 1),
             ("cfg2_LM_40x39_dups_n", cfg2_dup_matrix(gen), 1, "n", 1),
             ("cfg3_L2_18x32", gen.gen_matrix(18, 32, 50, 103), 2, "n", 0),
             ("cfg3_L3_18x32", gen.gen_matrix(18, 32, 50, 103), 3, "n", 0)]
    out = {}
    for name, M, d, stat, marg in cases:
        p, _, exited = host.load(gen.matrix_text(M).encode(), d, stat, marg)
        if exited:
            raise RuntimeError(name + ": rejected by the host pipeline")
        A, n, m, total = p.matrix_laid_out(), p.n, p.marginal, p.total
        host.free(p)
        pkg.calc_lnorm(A, n, m)
        t = time.perf_counter()
        norm, _, res = pkg.calc_lnorm(A, n, m)
        wall = time.perf_counter() - t
        out[name] = {"units": total, "laid_out": list(A.shape), "kernel": res["kernel_name"], "norm": norm,
                     "call_ms": 1e3 * wall, "kernel_ms": res["ms_kernel"], "units_per_s": total / wall}
    if out["cfg2_LM_40x39_dups_y"]["norm"] != out["cfg2_LM_40x39_dups_n"]["norm"]:
        raise RuntimeError("the marginal workload's preprocessed and plain runs disagree")
    return out


def batch_full_line(pkg, gen, chunk_labelings: float = 1.2e14) -> dict:
    """The mixed batch workload as drawn, without the 20-row cap on d = 3
    files (m up to 30 rows: 3^29 labelings for the largest item, about
    3.5e15 in all).  Every item goes through lnorm_hip_calc_batch_ex: first
    one call with the items the persistent launches take (up to 2^32
    labelings each), then the larger items, which the library runs one by
    one on AUTO's single-problem kernels, in calls of at most
    `chunk_labelings` labelings each, with a progress line after each call.
    Every result is checked on the host: the objective of the returned
    strategy (oracle.objective, one evaluation) equals the returned norm."""
    sys.path.insert(0, ROOT)
    import oracle
    from lnorm_amd import host
    work = gen.batch_workload(4096, d3_max_rows=30)
    probs, totals = [], []
    for w in work:
        p, _, exited = host.load(w["text"], w["d"], w["stat"], w["marginal"])
        if exited:
            raise RuntimeError("batch item %d rejected by the host pipeline" % w["seed"])
        probs.append((p.matrix_laid_out(), p.n, p.marginal))
        totals.append(p.total)
        host.free(p)
    small = [i for i, t in enumerate(totals) if t <= 2 ** 32]
    large = sorted((i for i, t in enumerate(totals) if t > 2 ** 32), key=lambda i: totals[i])
    calls = [small]
    cur, acc = [], 0
    for i in large:
        if cur and acc + totals[i] > chunk_labelings:
            calls.append(cur)
            cur, acc = [], 0
        cur.append(i)
        acc += totals[i]
    if cur:
        calls.append(cur)
    results = [None] * len(probs)
    by_d = {}
    wall_small = wall_large = 0.0
    st_small = None
    done = 0
    for k, idx in enumerate(calls):
        t = time.perf_counter()
        # the persistent call uses the batch autotuner; the larger items run
        # on AUTO's round-aware default launch shapes
        out, st = pkg.calc_batch([probs[i] for i in idx], stats=True, autotune=1 if (AUTOTUNE and k == 0) else 0)
        dt = time.perf_counter() - t
        if k == 0:
            wall_small, st_small = dt, st
        else:
            wall_large += dt
        for i, r in zip(idx, out):
            if r[2] != 0:
                raise RuntimeError("batch item %d failed (status %d)" % (i, r[2]))
            results[i] = r
            A, n, m = probs[i]
            if A.shape[0] > 1 and oracle.objective(A, n, m, r[1]) != r[0]:
                raise RuntimeError("item %d: the strategy's objective differs from the norm" % i)
            by_d.setdefault(n, [0, 0])
            by_d[n][0] += totals[i]
            by_d[n][1] += 1
        done += sum(totals[i] for i in idx)
        print("batch_full: call %d/%d, %d items, %.3g labelings in %.1f s (%.3g of %.3g done)"
              % (k + 1, len(calls), len(idx), sum(totals[i] for i in idx), dt, done, sum(totals)),
              file=sys.stderr, flush=True)
    lab_small = sum(totals[i] for i in small)
    lab_large = sum(totals[i] for i in large)
    wall = wall_small + wall_large
    return {"items": len(probs), "labelings": sum(totals),
            "labelings_by_d": {str(k): v[0] for k, v in sorted(by_d.items())},
            "items_by_d": {str(k): v[1] for k, v in sorted(by_d.items())},
            "value": sum(totals) / wall, "unit": "labelings/s", "wall_s": wall,
            "persistent_call": {"items": len(small), "labelings": lab_small, "wall_s": wall_small,
                                "ms_device": st_small["ms_device"], "groups": st_small["groups"],
                                "labelings_per_s": lab_small / wall_small},
            "large_items": {"items": len(large), "labelings": lab_large, "calls": len(calls) - 1,
                            "wall_s": wall_large, "labelings_per_s": lab_large / max(wall_large, 1e-9),
                            "rows": sorted({int(probs[i][0].shape[0]) for i in large})},
            "checked": "objective(strategy) == norm for every item"}


def tune_sweep(pkg, gen) -> dict:
    """Every autotuner candidate on the bench workloads: labelings/s of each
    probe and the shape the tuner keeps (lnorm_hip_autotune and
    lnorm_hip_autotune_batch)."""
    from lnorm_amd import host
    out = {}
    cases = {
        "L1_64x40_int8": (gen.gen_matrix(64, 40, 127, 104).T.copy(), 1, 0, {}),
        "L1_40x40": (gen.gen_matrix(40, 40, 100, 106), 1, 0, {}),
        "L1_128x40_int16": (gen.gen_matrix(128, 40, 30000, 105).T.copy(), 1, 0, {}),
        "LM_36x36": (gen.gen_matrix(36, 36, 100, 102), 1, 1, {}),
        "L1_64x40_plain": (gen.gen_matrix(64, 40, 127, 104).T.copy(), 1, 0, {"kernel": "plain"}),
        "L3_18x32": (gen.gen_matrix(18, 32, 50, 103), 3, 0, {}),
    }
    for name, (A, n, m, kw) in cases.items():
        best, entries = pkg.autotune(A, n, m, **kw)
        out[name] = {"best": {k: best[k] for k in ("kernel", "patterns_per_lane", "wave_log2", "block")},
                     "candidates": entries}
    probs = []
    for w in gen.batch_workload(4096):
        p, _, exited = host.load(w["text"], w["d"], w["stat"], w["marginal"])
        probs.append((p.matrix_laid_out(), p.n, p.marginal))
        host.free(p)
    best, entries = pkg.autotune_batch(probs)
    out["batch_4096"] = {"best": {k: best[k] for k in ("batch_streams", "batch_cap_log")}, "candidates": entries}
    return out


def wide_lines(pkg, gen) -> dict:
    """The wide L1 matrices on AUTO's tabulated walks beside the binary
    W-wave walk: 36 x 256 (2^35 sign vectors, one wave per segment) and
    34 x 1000 (2^33, 16 waves per segment)."""
    out = {}
    W = gen.gen_matrix(36, 256, 100, 107)
    wt = single_line(pkg, W, 1, 2 ** 35, valu_ops_per_unit("table16", 256, 0) + 128 / 16 - 128 / 32)
    ww = single_line(pkg, W, 1, 2 ** 35, 2 * 256 + 1, kernel="wide")
    if wt["norm"] != ww["norm"]:
        raise RuntimeError("tabulated and W-wave walks disagree on 36x256")
    out["wide_36x256"] = {**wt, "wwave": {k2: ww[k2] for k2 in ("value", "kernel", "kernel_ms", "valu_frac")},
                          "speedup_vs_wwave": wt["value"] / ww["value"]}
    # L1 of a 34 x 1000 matrix (2^33 sign vectors): the tabulated walk over
    # 16 waves beside the binary W-wave walk
    X = gen.gen_matrix(34, 1000, 30, 109)
    xt = single_line(pkg, X, 1, 2 ** 33, valu_ops_per_unit("table16w", 1000, 0))
    xw = single_line(pkg, X, 1, 2 ** 33, 2 * 1024 + 1, kernel="wide")
    if xt["norm"] != xw["norm"]:
        raise RuntimeError("tabulated and binary W-wave walks disagree on 34x1000")
    out["wide_34x1000"] = {**xt, "wwave": {k2: xw[k2] for k2 in ("value", "kernel", "kernel_ms", "valu_frac")},
                           "speedup_vs_wwave": xt["value"] / xw["value"]}
    return out


# Wide int32 workloads beyond the tabulated and W-wave walks: the column walk
# (one wave per segment, the columns over the lanes) beside the general
# kernel it replaces, which the test suite checks it against
COLUMN_CASES = [("L3_14x300_int32", 14, 300, 20000, 112, 3), ("L3_16x512_int32", 16, 512, 20000, 113, 3),
                ("L6_10x300_int32", 10, 300, 20000, 114, 6), ("L1_24x1500_int16", 24, 1500, 20000, 115, 1),
                ("L10_7x700_int32", 7, 700, 20000, 118, 10)]


def columns_lines(pkg, gen) -> dict:
    """Each COLUMN_CASES matrix on AUTO (the column walk) and on the general
    kernel (the 16x512 one on AUTO only): labelings/s, and the same norm and
    rank from both."""
    out = {}
    for name, R, C, r, seed, d in COLUMN_CASES:
        A = gen.gen_matrix(R, C, r, seed)
        total = 2 ** (R - 1) if d == 1 else d ** (R - 1)
        line = {}
        for k in ("auto", "general") if C != 512 else ("auto",):
            pkg.calc_lnorm(A, d, 0, kernel=k)
            norm, _, res = pkg.calc_lnorm(A, d, 0, kernel=k)
            line[k] = {"kernel": res["kernel_name"], "kernel_ms": res["ms_kernel"],
                       "labelings_per_s": total / (res["ms_kernel"] / 1e3), "norm": norm, "rank": res["best_rank"]}
        if len({(v["norm"], v["rank"]) for v in line.values()}) != 1:
            raise RuntimeError("the column walk and the general kernel disagree on " + name)
        if "general" in line:
            line["speedup_vs_general"] = line["auto"]["labelings_per_s"] / line["general"]["labelings_per_s"]
        out[name] = line
    return out


# int16-range workloads: column sums beyond 16 bits, so the 32-bit words
#   int16:      L1 of the 128 x 40 int16 matrix (entries in [-30000, 30000],
#               seed 105), laid out 40 x 128: 2^39 sign vectors
#   wide_int16: L1 of a 34 x 1000 matrix, entries in [-30000, 30000] (seed
#               110): 2^33 sign vectors over W waves
#   l3_int16:   L3 of a 20 x 200 matrix, entries in [-20000, 20000] (seed
#               111): 3^19 labelings over W waves
INT16_CASES = {
    "int16": lambda gen: (gen.gen_matrix(128, 40, 30000, 105).T.copy(), 1, 2 ** 39),
    "wide_int16": lambda gen: (gen.gen_matrix(34, 1000, 30000, 110), 1, 2 ** 33),
    "l3_int16": lambda gen: (gen.gen_matrix(20, 200, 20000, 111), 3, 3 ** 19),
}


def int16_line(pkg, gen, case: str, kernel: str, steps: int, warmup: int) -> dict:
    """One int16-range workload on one kernel: rates, the kernel that ran
    and, for the L1 walks, the VALU fraction by the kernel's op model."""
    A, n, total = INT16_CASES[case](gen)
    r = run_steps(pkg, A, kernel, steps, warmup, n=n)
    name = r["res"]["kernel_name"]
    model = next((k for k in ("table32w", "table32", "blocked32", "plain") if name.startswith(k)), None)
    if n == 1 and model:
        s = summarize(r, total, A.shape[1], model)
    else:
        steps_ = len(r["step_s"])
        kms = sum(r["kernel_ms"]) / steps_
        s = {"value": total * steps_ / r["wall_s"], "ms_per_step": 1e3 * r["wall_s"] / steps_, "kernel_ms": kms,
             "kernel_labelings_per_s": total / (kms / 1e3), "kernel_name": name, "norm": r["norm"],
             "best_rank": r["best_rank"]}
    return {"workload": case, "units": total, **s}


def run_case(pkg, gen, case: str, steps: int, warmup: int) -> int:
    """One workload alone: the headline matrix on the blocked, plain or
    tabulated kernel, the L3 workload, the wide L1 matrices, the 4096-item
    batch, or an int16-range workload (int16, wide_int16, l3_int16; with
    the suffix _blocked, _plain, _wide or _general on another kernel)."""
    base = case.rsplit("_", 1)[0] if case.endswith(("_blocked", "_plain", "_wide", "_general")) else case
    if base in INT16_CASES:
        kernel = case[len(base) + 1:] if case != base else "auto"
        print(json.dumps({"case": case, **int16_line(pkg, gen, base, kernel, steps, warmup)}), flush=True)
        return 0
    if case == "batch":
        print(json.dumps({"case": case, **batch_line(pkg, gen, steps)}), flush=True)
        return 0
    if case == "batch_full":
        print(json.dumps({"case": case, **batch_full_line(pkg, gen)}), flush=True)
        return 0
    if case == "l3":
        print(json.dumps({"case": case, **l3_line(pkg, gen, steps, warmup, False)}), flush=True)
        return 0
    if case == "wide":
        print(json.dumps({"case": case, **wide_lines(pkg, gen)}), flush=True)
        return 0
    if case == "configs":
        print(json.dumps({"case": case, **configs_line(pkg, gen)}), flush=True)
        return 0
    if case == "columns":
        print(json.dumps({"case": case, **columns_lines(pkg, gen)}), flush=True)
        return 0
    if case == "int64":
        # L1 of a 34 x 40 matrix scaled by 2^28 + 3 (beyond int32): the 64-bit walk
        A64 = gen.gen_matrix(34, 40, 100, 108).astype("int64") * (2 ** 28 + 3)
        print(json.dumps({"case": case, **single_line(pkg, A64, 1, 2 ** 33, None, int64=True)}), flush=True)
        return 0
    A = gen.gen_matrix(64, 40, 127, 104).T.copy()
    k, ny = A.shape
    r = run_steps(pkg, A, case, steps, warmup)
    print(json.dumps({"case": case, **summarize(r, 2 ** (k - 1), ny, case)}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the plain kernel and the 40 x 40 line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the default launch shapes instead of the autotuner's (tuned before timing)")
    ap.add_argument("--tune", action="store_true",
                    help="sweep the autotuner's candidates on the bench workloads and print them as JSON")
    ap.add_argument("--case", choices=["blocked", "plain", "table16", "batch", "batch_full", "l3", "int64", "columns",
                                       "configs", "wide", "int16", "int16_blocked",
                                       "int16_plain", "wide_int16", "wide_int16_wide", "l3_int16",
                                       "l3_int16_general"],
                    help="run only this workload (for profiler passes) and print its figures")
    ap.add_argument("--copy-calibration", type=int, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB (torch) and exit: the reference "
                         "kernel for scaling the profiler's FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only.  Over several GPUs the enumeration shards by Gray-rank ranges "
              "(rank_begin / rank_end, one call per GPU, then one max / smallest-rank exchange).", file=sys.stderr)
        return 2
    if a.copy_calibration:
        import torch
        n = a.copy_calibration * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                          "ok": bool(torch.equal(dst[-16:].cpu(), src[-16:].cpu()))}), flush=True)
        return 0
    global AUTOTUNE
    AUTOTUNE = not a.no_autotune
    pkg = load_pkg()
    from lnorm_amd import gen
    if a.case:
        return run_case(pkg, gen, a.case, a.steps, a.warmup)
    if a.tune:
        print(json.dumps(tune_sweep(pkg, gen)), flush=True)
        return 0

    A = gen.gen_matrix(64, 40, 127, 104).T.copy()       # laid out 40 x 64 as the reference does
    k, ny = A.shape
    total = 2 ** (k - 1)
    blk = run_steps(pkg, A, "blocked", a.steps, a.warmup)
    sb = summarize(blk, total, ny, "blocked")
    progress("headline (blocked) done")
    # the sustained rate of the headline's instruction (v_sad_u16) and of the
    # 32-bit walks' (v_sad_u32) on this GPU, for the published peak beside it
    r16, ms16 = pkg.valu_rate("sad_u16")
    r32, _ = pkg.valu_rate("sad_u32")
    measured = {"sad_u16_lane_ops_per_s": r16, "sad_u32_lane_ops_per_s": r32,
                "sustained_clock_ghz": r16 / (256 * 64) / 1e9,
                "row_update_peak_TOPs": 2 * r16 / 1e12,
                "frac_of_measured": sb["row_updates_per_s"] / (2 * r16),
                "probe": "lnorm_hip_valu_rate: 8 independent chains per lane, 8 waves per SIMD, SGPR operand, "
                         "no memory traffic, one %.1f ms launch after a warm-up launch of half that" % ms16}
    line = {
        "metric": METRIC,
        "value": sb["value"],
        "unit": "sign-vectors/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sb["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "L1 of a 64x40 int8 matrix U[-127,127] seed 104, laid out 40x64: 2^39 sign vectors "
                        "per step, norm + strategy",
            "k": k, "n_y": ny, "sign_vectors_per_step": total,
            "kernel": sb["kernel_name"],
            "autotuned": blk["tuned"],
            "arith": "int8 MFMA (int32 accumulate) builds the low-row table; v_sad_u16 on biased 16-bit halves "
                     "into uint32 sums; exact",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "achieved": sb["row_updates_per_s"] / 1e12,
            "peak": VALU16_PEAK / 1e12,
            "unit": "TOP/s",
            "frac": sb["row_updates_per_s"] / VALU16_PEAK,
            "work": "algorithmic: 1 op per row-update (2^39 sign vectors x 64 partial sums per launch: the "
                    "MFMA path's VALU rule) over the kernel's HIP-event time; peak: the packed 16-bit VALU rate, "
                    "2 row-updates per lane-op (v_sad_u16 on two biased halves), 2 x 39.32e12",
            "issue_frac": sb["valu_frac"],
            "issue_frac_basis": "the kernel's VALU lane-ops by its op model (%.1f per sign vector, equal to "
                                "SQ_INSTS_VALU x 64 per sign vector in the counter profile) over the int32 "
                                "lane-op peak 39.32e12" % sb["valu_ops_per_sign_vector"],
            "issue_frac_counters": PROFILED_ISSUE.get(sb["kernel_name"]),
            "traffic": PROFILED_TRAFFIC.get(sb["kernel_name"]),
            "traffic_per": "launch (2^22 waves at most per launch; one per call at the tuned shape)",
            "kernel_ms": sb["kernel_ms"],
            "peak_measured": measured,
        },
        "row_updates_per_s": sb["row_updates_per_s"],
        "mfma": mfma_share(blk["res"], k, ny),
        "norm": sb["norm"],
        "best_rank": sb["best_rank"],
    }
    if not a.no_extra:
        pl = run_steps(pkg, A, "plain", a.steps, a.warmup)
        if (pl["norm"], pl["best_rank"]) != (blk["norm"], blk["best_rank"]):
            raise RuntimeError("plain and blocked kernels disagree")
        sp = summarize(pl, total, ny, "plain")
        line["plain_gray"] = {k2: sp[k2] for k2 in ("value", "ms_per_step", "kernel_ms", "row_updates_per_s",
                                                    "valu_TOPs", "valu_frac", "alg_frac", "kernel_name")}
        line["speedup_vs_plain"] = sb["value"] / sp["value"]
        line["kernel_speedup_vs_plain"] = sp["kernel_ms"] / sb["kernel_ms"]
        tb = run_steps(pkg, A, "table16", a.steps, a.warmup)
        if (tb["norm"], tb["best_rank"]) != (blk["norm"], blk["best_rank"]):
            raise RuntimeError("tabulated and blocked kernels disagree")
        st16 = summarize(tb, total, ny, "table16")
        line["table16_same_matrix"] = {k2: st16[k2] for k2 in ("value", "ms_per_step", "kernel_ms", "valu_TOPs",
                                                              "valu_frac", "alg_frac", "kernel_name")}
        progress("plain and table16 on the headline matrix done")
        B = gen.gen_matrix(40, 40, 100, 106)
        nb = run_steps(pkg, B, "blocked", a.steps, a.warmup)
        npl = run_steps(pkg, B, "plain", max(1, a.steps // 2), a.warmup)
        if (nb["norm"], nb["best_rank"]) != (npl["norm"], npl["best_rank"]):
            raise RuntimeError("plain and blocked kernels disagree on 40x40")
        ntb = run_steps(pkg, B, "table16", a.steps, a.warmup)
        if (ntb["norm"], ntb["best_rank"]) != (npl["norm"], npl["best_rank"]):
            raise RuntimeError("plain and tabulated kernels disagree on 40x40")
        s1 = summarize(nb, 2 ** 39, 40, "blocked")
        s2 = summarize(npl, 2 ** 39, 40, "plain")
        s5 = summarize(ntb, 2 ** 39, 40, "table16")
        line["n40_m40"] = {"blocked": s1["value"], "plain": s2["value"], "table16": s5["value"],
                           "speedup": s1["value"] / s2["value"],
                           "blocked_valu_frac": s1["valu_frac"], "plain_valu_frac": s2["valu_frac"],
                           "table16_valu_frac": s5["valu_frac"], "norm": s1["norm"]}
        progress("40 x 40 done")
        # 128 x 40 int16 (entries in [-30000, 30000], seed 105), laid out 40 x 128:
        # column sums beyond 2^15, so the blocked path uses 32-bit words
        # (hi/lo int8 MFMA planes for the table, one v_sad_u32 per column)
        C16 = gen.gen_matrix(128, 40, 30000, 105).T.copy()
        a16 = run_steps(pkg, C16, "auto", 2, 1)
        b16 = run_steps(pkg, C16, "blocked", 1, 0)
        p16 = run_steps(pkg, C16, "plain", 1, 0)
        if not ((a16["norm"], a16["best_rank"]) == (b16["norm"], b16["best_rank"]) == (p16["norm"], p16["best_rank"])):
            raise RuntimeError("the kernels disagree on the int16 matrix")
        s6 = summarize(a16, 2 ** 39, 128, "table32")
        s3 = summarize(b16, 2 ** 39, 128, "blocked32")
        s4 = summarize(p16, 2 ** 39, 128, "plain")
        line["int16_128x40"] = {"auto": s6["value"], "auto_kernel": s6["kernel_name"],
                                "auto_speedup": s6["value"] / s4["value"],
                                "auto_valu_frac": s6["valu_frac"], "auto_alg_frac": s6["alg_frac"],
                                "blocked": s3["value"], "plain": s4["value"], "speedup": s3["value"] / s4["value"],
                                "blocked_valu_frac": s3["valu_frac"], "plain_valu_frac": s4["valu_frac"],
                                "kernel": s3["kernel_name"], "norm": s3["norm"]}
        progress("int16 128 x 40 done")
        line["cli_end_to_end"] = cli_line(pkg, gen)
        line.update(wide_lines(pkg, gen))
        progress("CLI and wide matrices done")
        line["int64_34x40"] = single_line(pkg, gen.gen_matrix(34, 40, 100, 108).astype("int64") * (2 ** 28 + 3),
                                          1, 2 ** 33, None, int64=True)
        line["L3_24x32"] = l3_line(pkg, gen, a.steps, a.warmup, not a.no_cpu)
        progress("int64 and L3 done")
        line["configs"] = configs_line(pkg, gen)
        line["batch_4096"] = batch_line(pkg, gen, max(1, a.steps // 2))
        progress("configs and batch done")
    if not a.no_cpu:
        cb = cpu_baseline(gen)
        line["cpu_baseline"] = {k2: cb[k2] for k2 in ("value", "unit", "cores", "kind", "sample")}
        line["speedup_vs_cpu"] = sb["value"] / cb["value"]
        if "n40_m40" in line:
            c40 = cpu_baseline(gen, rows=40, r=100, seed=106, budget_cols=34)
            line["n40_m40"]["cpu_baseline"] = {k2: c40[k2] for k2 in ("value", "cores", "kind", "sample")}
            line["n40_m40"]["speedup_vs_cpu"] = line["n40_m40"]["blocked"] / c40["value"]
        progress("CPU baselines done")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
