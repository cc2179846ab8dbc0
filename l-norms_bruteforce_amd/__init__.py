"""lnorm_amd -- MI355X (gfx950) exact L_d matrix-norm engine.

Python view of the C ABI in include/lnorm_hip.h (the drop-in for the
reference's calc_Lnorm, L_CUDA.cu:281-359) and include/lnorm_host.h (the
reference host pipeline: reader, preprocessing, layout, printer).  The
compute path is the HIP library lib/liblnorm_hip.so; there is no CPU
fallback: loading fails loudly when the library is missing.

This directory is imported under the name `lnorm_amd` (its folder name is not
a Python identifier); see load() in tests/conftest.py and bench.py.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
HIP_LIB = os.path.join(LIB_DIR, "liblnorm_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "liblnorm_host.so")
CLI = os.path.join(PKG_DIR, "bin", "L_HIP")

(KERNEL_AUTO, KERNEL_PLAIN, KERNEL_BLOCKED, KERNEL_GENERAL, KERNEL_PLAIN16, KERNEL_WIDE, KERNEL_TABLE16,
 KERNEL_TABLE32, KERNEL_COLUMNS) = range(9)
KERNELS = {"auto": 0, "plain": 1, "blocked": 2, "general": 3, "plain16": 4, "wide": 5, "table16": 6, "table32": 7,
           "columns": 8}
ERRORS = {0: "OK", -1: "E_ARG", -2: "E_BOUND", -3: "E_SIZE", -4: "E_HIP", -5: "E_UNSUPPORTED"}


class Opts(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("want_strategy", ctypes.c_int32),
                ("seg_log", ctypes.c_int32), ("patterns_per_lane", ctypes.c_int32),
                ("wave_log2", ctypes.c_int32), ("block", ctypes.c_int32),
                ("device", ctypes.c_int32), ("word_bits", ctypes.c_int32),
                ("stream", ctypes.c_void_p), ("batch_graph", ctypes.c_int32), ("autotune", ctypes.c_int32),
                ("batch_streams", ctypes.c_int32), ("batch_cap_log", ctypes.c_int32),
                ("rank_begin", ctypes.c_uint64), ("rank_end", ctypes.c_uint64)]


class Result(ctypes.Structure):
    _fields_ = [("Lnorm", ctypes.c_int32), ("kernel", ctypes.c_int32),
                ("total", ctypes.c_uint64), ("best_rank", ctypes.c_uint64),
                ("ms_enumerate", ctypes.c_double), ("ms_kernel", ctypes.c_double),
                ("ms_strategy", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("units", ctypes.c_uint64), ("unit_log", ctypes.c_int32),
                ("patterns_per_lane", ctypes.c_int32), ("kernel_name", ctypes.c_char * 64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "kernel_name"}
        d["kernel_name"] = self.kernel_name.decode()
        return d


class BatchStats(ctypes.Structure):
    _fields_ = [("items", ctypes.c_int32), ("batched", ctypes.c_int32), ("fallback", ctypes.c_int32),
                ("groups", ctypes.c_int32), ("graph_replayed", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("labelings", ctypes.c_uint64), ("ms_device", ctypes.c_double), ("ms_total", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class TuneEntry(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("patterns_per_lane", ctypes.c_int32),
                ("wave_log2", ctypes.c_int32), ("block", ctypes.c_int32),
                ("batch_streams", ctypes.c_int32), ("batch_cap_log", ctypes.c_int32),
                ("ms", ctypes.c_double), ("labelings_per_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Item(ctypes.Structure):
    _fields_ = [("mtx_as_vec", ctypes.POINTER(ctypes.c_int32)),
                ("iRows", ctypes.c_int32), ("iCols", ctypes.c_int32),
                ("n", ctypes.c_int32), ("marginal", ctypes.c_int32),
                ("strategy", ctypes.POINTER(ctypes.c_int32)),
                ("Lnorm", ctypes.c_int32), ("status", ctypes.c_int32)]


# symbols declared in include/lnorm_hip.h
HIP_SYMBOLS = ["lnorm_hip_calc_Lnorm", "lnorm_hip_calc_ex", "lnorm_hip_calc_batch",
               "lnorm_hip_default_opts", "lnorm_hip_last_error", "lnorm_hip_abi_version",
               "lnorm_hip_device_count", "lnorm_hip_release", "lnorm_hip_selftest",
               "lnorm_hip_debug_table", "lnorm_hip_calc_batch_ex", "lnorm_hip_autotune",
               "lnorm_hip_autotune_batch", "lnorm_hip_calc_Lnorm64", "lnorm_hip_calc_ex64",
               "lnorm_hip_calc_noalloc", "lnorm_hip_valu_rate"]

_hip = None


class LnormError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


def hip_lib():
    """The HIP engine.  Raises if it was not built (no CPU fallback)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError("HIP engine not built: %s is missing (run __graft_entry__.build())" % HIP_LIB)
        L = ctypes.CDLL(HIP_LIB)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.lnorm_hip_calc_Lnorm.restype = ctypes.c_int
        L.lnorm_hip_calc_Lnorm.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, i32p, i32p]
        L.lnorm_hip_calc_ex.restype = ctypes.c_int
        L.lnorm_hip_calc_ex.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(Opts), i32p, ctypes.POINTER(Result)]
        L.lnorm_hip_calc_noalloc.restype = ctypes.c_int
        L.lnorm_hip_calc_noalloc.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(Opts), i32p, ctypes.POINTER(Result)]
        L.lnorm_hip_calc_batch.restype = ctypes.c_int
        L.lnorm_hip_calc_batch.argtypes = [ctypes.POINTER(Item), ctypes.c_int, ctypes.POINTER(Opts)]
        L.lnorm_hip_calc_batch_ex.restype = ctypes.c_int
        L.lnorm_hip_calc_batch_ex.argtypes = [ctypes.POINTER(Item), ctypes.c_int, ctypes.POINTER(Opts),
                                              ctypes.POINTER(BatchStats)]
        L.lnorm_hip_autotune.restype = ctypes.c_int
        L.lnorm_hip_autotune.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(Opts), ctypes.POINTER(Opts), ctypes.POINTER(TuneEntry),
                                         ctypes.c_int]
        L.lnorm_hip_autotune_batch.restype = ctypes.c_int
        L.lnorm_hip_autotune_batch.argtypes = [ctypes.POINTER(Item), ctypes.c_int, ctypes.POINTER(Opts),
                                               ctypes.POINTER(Opts), ctypes.POINTER(TuneEntry), ctypes.c_int]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.lnorm_hip_calc_ex64.restype = ctypes.c_int
        L.lnorm_hip_calc_ex64.argtypes = [i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(Opts), i32p, i64p, ctypes.POINTER(Result)]
        L.lnorm_hip_calc_Lnorm64.restype = ctypes.c_int
        L.lnorm_hip_calc_Lnorm64.argtypes = [i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, i64p, i32p]
        L.lnorm_hip_default_opts.restype = None
        L.lnorm_hip_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.lnorm_hip_last_error.restype = ctypes.c_char_p
        L.lnorm_hip_last_error.argtypes = []
        L.lnorm_hip_abi_version.restype = ctypes.c_int
        L.lnorm_hip_device_count.restype = ctypes.c_int
        L.lnorm_hip_release.restype = None
        L.lnorm_hip_selftest.restype = ctypes.c_int
        L.lnorm_hip_valu_rate.restype = ctypes.c_int
        L.lnorm_hip_valu_rate.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
        L.lnorm_hip_debug_table.restype = ctypes.c_int
        L.lnorm_hip_debug_table.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
        _hip = L
    return _hip


def last_error() -> str:
    return hip_lib().lnorm_hip_last_error().decode()


def _check(rc):
    if rc < 0:
        raise LnormError(rc, last_error())
    return rc


def _i32(A):
    a = np.ascontiguousarray(np.asarray(A, dtype=np.int32))
    if a.ndim != 2:
        raise ValueError("matrix must be 2-D")
    return a


def default_opts(**kw) -> Opts:
    o = Opts()
    hip_lib().lnorm_hip_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if k == "kernel" and isinstance(v, str):
            v = KERNELS[v]
        setattr(o, k, v)
    return o


def calc_lnorm(A, n: int, marginal: int = 0, want_strategy: bool = True, noalloc: bool = False, **opts):
    """The reference's calc_Lnorm on the GPU for a laid-out matrix
    (mtx_as_vec, R x C).  Returns (Lnorm, strategy list, result dict).
    noalloc: through lnorm_hip_calc_noalloc (device buffers carved from the
    library's static per-device scratch array, no device allocation)."""
    a = _i32(A)
    R, C = a.shape
    o = default_opts(want_strategy=int(want_strategy), **opts)
    strat = np.zeros(max(R - 1, 1), dtype=np.int32)
    res = Result()
    i32p = ctypes.POINTER(ctypes.c_int32)
    fn = hip_lib().lnorm_hip_calc_noalloc if noalloc else hip_lib().lnorm_hip_calc_ex
    _check(fn(a.ctypes.data_as(i32p), R, C, n, marginal, ctypes.byref(o), strat.ctypes.data_as(i32p),
              ctypes.byref(res)))
    return int(res.Lnorm), [int(v) for v in strat[: R - 1]] if want_strategy else None, res.as_dict()


def calc_lnorm64(A, n: int, marginal: int = 0, want_strategy: bool = True, **opts):
    """The int_type = long long variant (lnorm_hip_calc_ex64): int64 entries
    and norm.  Returns (Lnorm, strategy list, result dict)."""
    a = np.ascontiguousarray(np.asarray(A, dtype=np.int64))
    if a.ndim != 2:
        raise ValueError("matrix must be 2-D")
    R, C = a.shape
    o = default_opts(want_strategy=int(want_strategy), **opts)
    strat = np.zeros(max(R - 1, 1), dtype=np.int32)
    res = Result()
    norm = ctypes.c_int64()
    _check(hip_lib().lnorm_hip_calc_ex64(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), R, C, n, marginal,
                                         ctypes.byref(o), strat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         ctypes.byref(norm), ctypes.byref(res)))
    return int(norm.value), [int(v) for v in strat[: R - 1]] if want_strategy else None, res.as_dict()


def calc_lnorm_simple(A, n: int, marginal: int = 0, blocks: int = 512, threads: int = 32):
    """lnorm_hip_calc_Lnorm: the calc_Lnorm(item_calc*, B, T) signature."""
    a = _i32(A)
    R, C = a.shape
    norm = ctypes.c_int32()
    strat = np.zeros(max(R - 1, 1), dtype=np.int32)
    i32p = ctypes.POINTER(ctypes.c_int32)
    _check(hip_lib().lnorm_hip_calc_Lnorm(a.ctypes.data_as(i32p), R, C, n, marginal, blocks, threads,
                                          ctypes.byref(norm), strat.ctypes.data_as(i32p)))
    return int(norm.value), [int(v) for v in strat[: R - 1]]


def calc_batch(problems, want_strategy=True, stats=False, **opts):
    """problems: iterable of (A, n, marginal).  Returns [(Lnorm, strategy, status)]
    (and the batch statistics when stats=True)."""
    b = _Batch(problems, want_strategy)
    o = default_opts(**opts)
    st = BatchStats()
    rc = hip_lib().lnorm_hip_calc_batch_ex(b.ptr, b.count, ctypes.byref(o), ctypes.byref(st))
    if rc == -4:
        _check(rc)
    out = b.results(want_strategy)
    return (out, st.as_dict()) if stats else out


def _opts_dict(o: Opts) -> dict:
    return {k: getattr(o, k) for k, _ in o._fields_ if k != "stream"}


def autotune(A, n: int, marginal: int = 0, **opts):
    """lnorm_hip_autotune: times the launch-shape candidates of the kernel this
    problem runs on and caches the fastest.  Returns (best options dict,
    [candidate dicts])."""
    a = _i32(A)
    R, C = a.shape
    o = default_opts(**opts)
    best = Opts()
    ent = (TuneEntry * 64)()
    i32p = ctypes.POINTER(ctypes.c_int32)
    n_ent = _check(hip_lib().lnorm_hip_autotune(a.ctypes.data_as(i32p), R, C, n, marginal, ctypes.byref(o),
                                                ctypes.byref(best), ent, 64))
    return _opts_dict(best), [ent[i].as_dict() for i in range(min(n_ent, 64))]


# lnorm_hip_item as a numpy record (the same offsets and size as Item)
_ITEM_DTYPE = np.dtype({"names": [f for f, _ in Item._fields_],
                        "formats": [np.uintp, np.int32, np.int32, np.int32, np.int32, np.uintp, np.int32, np.int32],
                        "offsets": [getattr(Item, f).offset for f, _ in Item._fields_],
                        "itemsize": ctypes.sizeof(Item)})


class _Batch:
    """The items of one batch call: an array of lnorm_hip_item records built
    with numpy (no per-item ctypes objects), every matrix in one int32
    buffer and every strategy (max(R - 1, 1) ints each) in another."""

    def __init__(self, problems, want_strategy: bool):
        problems = list(problems)
        mats = [_i32(A) for A, _, _ in problems]
        k = self.count = len(mats)
        shape = np.array([m.shape for m in mats], dtype=np.int64).reshape(k, 2)
        size = shape[:, 0] * shape[:, 1]
        self.buf = np.concatenate([m.ravel() for m in mats] + [np.zeros(1, np.int32)])
        moff = np.concatenate(([0], np.cumsum(size)[:-1])) if k else np.zeros(0, np.int64)
        slen = np.maximum(shape[:, 0] - 1, 1)
        self.soff = np.concatenate(([0], np.cumsum(slen))).astype(np.int64)
        self.strat = np.zeros(int(self.soff[-1]) + 1, np.int32)
        a = np.zeros(k, _ITEM_DTYPE)
        a["mtx_as_vec"] = self.buf.ctypes.data + 4 * moff
        a["iRows"], a["iCols"] = shape[:, 0], shape[:, 1]
        a["n"] = np.fromiter((p[1] for p in problems), np.int32, k)
        a["marginal"] = np.fromiter((p[2] for p in problems), np.int32, k)
        if want_strategy:
            a["strategy"] = self.strat.ctypes.data + 4 * self.soff[:-1]
        self.a = a
        self.ptr = ctypes.cast(a.ctypes.data, ctypes.POINTER(Item))

    def results(self, want_strategy: bool) -> list:
        """[(Lnorm, strategy (R - 1 ints) or None, status)] per item."""
        L, S = self.a["Lnorm"].tolist(), self.a["status"].tolist()
        if not want_strategy:
            return [(l, None, s) for l, s in zip(L, S)]
        full = self.strat.tolist()
        n = np.maximum(self.a["iRows"].astype(np.int64) - 1, 0).tolist()
        return [(l, full[o:o + c], s) for l, o, c, s in zip(L, self.soff[:-1].tolist(), n, S)]


def autotune_batch(problems, **opts):
    """lnorm_hip_autotune_batch: times batch_streams x batch_cap_log on a
    sample of the batch and caches the fastest for this batch shape.
    Returns (best options dict, [candidate dicts])."""
    b = _Batch(problems, False)
    o = default_opts(**opts)
    best = Opts()
    ent = (TuneEntry * 64)()
    n_ent = _check(hip_lib().lnorm_hip_autotune_batch(b.ptr, b.count, ctypes.byref(o), ctypes.byref(best),
                                                      ent, 64))
    return _opts_dict(best), [ent[i].as_dict() for i in range(min(n_ent, 64))]


def selftest() -> int:
    return _check(hip_lib().lnorm_hip_selftest())


def debug_table(A, marginal: int, P: int, word_bits: int = 0):
    """The blocked kernel's low-pattern table, flat [P][words][64]."""
    a = _i32(A)
    R, C = a.shape
    words = P * (C + 40) * 64
    out = np.zeros(words, dtype=np.uint32)
    i32p = ctypes.POINTER(ctypes.c_int32)
    n = _check(hip_lib().lnorm_hip_debug_table(a.ctypes.data_as(i32p), R, C, marginal, P, word_bits,
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), words))
    return out[:n]


VALU_RATE_OPS = {"sad_u16": 0, "sad_u32": 1, "add_u32": 2}


def valu_rate(op: str = "sad_u16"):
    """lnorm_hip_valu_rate: the measured sustained rate of one VALU
    instruction on the current device.  Returns (lane-ops per second, ms of
    the timed launch)."""
    rate, ms = ctypes.c_double(), ctypes.c_double()
    _check(hip_lib().lnorm_hip_valu_rate(VALU_RATE_OPS[op], ctypes.byref(rate), ctypes.byref(ms)))
    return rate.value, ms.value


def device_count() -> int:
    return hip_lib().lnorm_hip_device_count()
