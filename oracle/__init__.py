"""The oracle: a CPU restatement of the reference L_d enumeration.

TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker; the product never imports it.

  lnorm_oracle.c  -- the Gray walks of L_OpenMP.c (L1, LM, L2, L3, Ln), the
                     calc_jmin_jmax partition and first-maximum argmax, and an
                     independent brute force (build/liblnorm_oracle.so).
  host_oracle.py  -- the host pipeline of functions.h / L_OpenMP.c /
                     L_CUDA.cu: reader, preprocessing, layout, parameters,
                     printer.
  _ref/L_OpenMP   -- the reference itself, compiled from /root/reference by
                     `make -C oracle ref` when the sources are present.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liblnorm_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "L_OpenMP")

_lib = None


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        i32p = ctypes.POINTER(ctypes.c_int32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.lo_total.restype = ctypes.c_uint64
        L.lo_total.argtypes = [ctypes.c_int, ctypes.c_int]
        L.lo_max_rows.restype = ctypes.c_int
        L.lo_max_rows.argtypes = [ctypes.c_int]
        L.lo_digit.restype = ctypes.c_int
        L.lo_digit.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        L.lo_walk.restype = ctypes.c_int
        L.lo_walk.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_uint64, ctypes.c_uint64, i32p, u64p, i32p]
        L.lo_calc_lnorm.restype = ctypes.c_int
        L.lo_calc_lnorm.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, i32p, i32p, u64p]
        L.lo_brute.restype = ctypes.c_int
        L.lo_brute.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, u64p]
        L.lo_objective.restype = ctypes.c_int32
        L.lo_objective.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.lo_walk64.restype = ctypes.c_int
        L.lo_walk64.argtypes = [i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_uint64, ctypes.c_uint64, i64p, u64p, i32p]
        L.lo_calc_lnorm64.restype = ctypes.c_int
        L.lo_calc_lnorm64.argtypes = [i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, i64p, i32p, u64p]
        L.lo_brute64.restype = ctypes.c_int
        L.lo_brute64.argtypes = [i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i64p, u64p]
        L.lo_objective64.restype = ctypes.c_int64
        L.lo_objective64.argtypes = [i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        _lib = L
    return _lib


def _arr(A):
    a = np.ascontiguousarray(np.asarray(A, dtype=np.int32))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def total(R: int, n: int) -> int:
    return int(lib().lo_total(R, n))


def max_rows(n: int) -> int:
    return int(lib().lo_max_rows(n))


def digit(N: int, i: int, n: int) -> int:
    return int(lib().lo_digit(N, i, n))


def labels_of_rank(N: int, R: int, n: int):
    """Strategy (R-1 labels) of Gray rank N, as the reference's vect[]."""
    return [digit(N, i, n) for i in range(R - 1)]


def calc_lnorm(A, n: int, marginal: int = 0, threads: int = 8):
    """(norm, strategy, rank) for a laid-out R x C matrix (mtx_as_vec)."""
    a, p = _arr(A)
    R, C = a.shape
    norm = ctypes.c_int32()
    rank = ctypes.c_uint64()
    strat = np.zeros(max(R - 1, 1), dtype=np.int32)
    rc = lib().lo_calc_lnorm(p, R, C, n, marginal, threads, ctypes.byref(norm),
                             strat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(rank))
    if rc != 0:
        raise RuntimeError("lo_calc_lnorm failed")
    return int(norm.value), [int(v) for v in strat[: R - 1]], int(rank.value)


def walk(A, n: int, marginal: int, jmin: int, jmax: int):
    """(best, first rank of best, strategy) over the inclusive rank range."""
    a, p = _arr(A)
    R, C = a.shape
    best = ctypes.c_int32()
    rank = ctypes.c_uint64()
    strat = np.zeros(max(R - 1, 1), dtype=np.int32)
    rc = lib().lo_walk(p, R, C, n, marginal, jmin, jmax, ctypes.byref(best), ctypes.byref(rank),
                       strat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc != 0:
        raise RuntimeError("lo_walk failed")
    return int(best.value), int(rank.value), [int(v) for v in strat[: R - 1]]


def brute(A, n: int, marginal: int = 0):
    """(norm, smallest maximising rank) by direct decoding of every rank."""
    a, p = _arr(A)
    R, C = a.shape
    norm = ctypes.c_int32()
    rank = ctypes.c_uint64()
    if lib().lo_brute(p, R, C, n, marginal, ctypes.byref(norm), ctypes.byref(rank)) != 0:
        raise RuntimeError("lo_brute failed")
    return int(norm.value), int(rank.value)


def objective(A, n: int, marginal: int, labels) -> int:
    a, p = _arr(A)
    R, C = a.shape
    lab = np.ascontiguousarray(np.asarray(list(labels) + [0], dtype=np.int32))
    return int(lib().lo_objective(p, R, C, n, marginal, lab.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))


# ---- the int_type = long long variant (README.md:100-105) -------------------

def _arr64(A):
    a = np.ascontiguousarray(np.asarray(A, dtype=np.int64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def calc_lnorm64(A, n: int, marginal: int = 0, threads: int = 8):
    """(norm, strategy, rank) in 64-bit arithmetic for an int64 matrix."""
    a, p = _arr64(A)
    R, C = a.shape
    norm = ctypes.c_int64()
    rank = ctypes.c_uint64()
    strat = np.zeros(max(R - 1, 1), dtype=np.int32)
    rc = lib().lo_calc_lnorm64(p, R, C, n, marginal, threads, ctypes.byref(norm),
                               strat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(rank))
    if rc != 0:
        raise RuntimeError("lo_calc_lnorm64 failed")
    return int(norm.value), [int(v) for v in strat[: R - 1]], int(rank.value)


def brute64(A, n: int, marginal: int = 0):
    a, p = _arr64(A)
    R, C = a.shape
    norm = ctypes.c_int64()
    rank = ctypes.c_uint64()
    if lib().lo_brute64(p, R, C, n, marginal, ctypes.byref(norm), ctypes.byref(rank)) != 0:
        raise RuntimeError("lo_brute64 failed")
    return int(norm.value), int(rank.value)


def walk64(A, n: int, marginal: int, jmin: int, jmax: int):
    """(best, first rank of best, strategy) over the inclusive rank range,
    long long variant."""
    a, p = _arr64(A)
    R, C = a.shape
    best = ctypes.c_int64()
    rank = ctypes.c_uint64()
    strat = np.zeros(max(R - 1, 1), dtype=np.int32)
    rc = lib().lo_walk64(p, R, C, n, marginal, jmin, jmax, ctypes.byref(best), ctypes.byref(rank),
                         strat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc != 0:
        raise RuntimeError("lo_walk64 failed")
    return int(best.value), int(rank.value), [int(v) for v in strat[: R - 1]]


def objective64(A, n: int, marginal: int, labels) -> int:
    a, p = _arr64(A)
    R, C = a.shape
    lab = np.ascontiguousarray(np.asarray(list(labels) + [0], dtype=np.int32))
    return int(lib().lo_objective64(p, R, C, n, marginal, lab.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
