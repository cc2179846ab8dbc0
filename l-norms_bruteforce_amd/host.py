"""ctypes view of the C host library (include/lnorm_host.h): the reader,
preprocessing, layout, work sizing and printer of the reference host code."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import HOST_LIB

OMP, CUDA = 0, 1


class Problem(ctypes.Structure):
    _fields_ = [("n_original", ctypes.c_int), ("stat", ctypes.c_char), ("marginal_req", ctypes.c_int),
                ("iRows", ctypes.c_int), ("iCols", ctypes.c_int),
                ("matrix", ctypes.POINTER(ctypes.c_int32)),
                ("original_r", ctypes.POINTER(ctypes.c_int)), ("original_c", ctypes.POINTER(ctypes.c_int)),
                ("R", ctypes.c_int), ("C", ctypes.c_int), ("n", ctypes.c_int), ("marginal", ctypes.c_int),
                ("maxRows", ctypes.c_int), ("total", ctypes.c_uint64), ("copyNum", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("steps_remainder", ctypes.c_uint64),
                ("mtx_as_vec", ctypes.POINTER(ctypes.c_int32)),
                ("original", ctypes.POINTER(ctypes.c_int)), ("original_length", ctypes.c_int),
                ("transposed", ctypes.c_int)]

    def matrix_laid_out(self) -> np.ndarray:
        if self.R == 0 or self.C == 0:
            return np.zeros((self.R, self.C), np.int32)
        return np.ctypeslib.as_array(self.mtx_as_vec, shape=(self.R * self.C,)).reshape(self.R, self.C).copy()


_lib = None
_libc = ctypes.CDLL(None)
_libc.open_memstream.restype = ctypes.c_void_p
_libc.open_memstream.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_size_t)]
_libc.fclose.argtypes = [ctypes.c_void_p]
_libc.fflush.argtypes = [ctypes.c_void_p]
_libc.free.argtypes = [ctypes.c_void_p]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("host library not built: %s" % HOST_LIB)
        L = ctypes.CDLL(HOST_LIB)
        P = ctypes.POINTER(Problem)
        L.lnorm_host_load_text.restype = ctypes.c_int
        L.lnorm_host_load_text.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_char,
                                           ctypes.c_int, ctypes.c_void_p]
        L.lnorm_host_params.restype = ctypes.c_int
        L.lnorm_host_params.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.c_void_p]
        L.lnorm_host_print_results.restype = ctypes.c_int
        L.lnorm_host_print_results.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p,
                                               ctypes.c_void_p]
        L.lnorm_host_strategy_name.restype = ctypes.c_char_p
        L.lnorm_host_strategy_name.argtypes = [P]
        L.lnorm_host_free.restype = None
        L.lnorm_host_free.argtypes = [P]
        _lib = L
    return _lib


class _Capture:
    """A FILE* writing into memory (open_memstream)."""

    def __enter__(self):
        self.buf = ctypes.c_char_p()
        self.size = ctypes.c_size_t()
        self.fp = _libc.open_memstream(ctypes.byref(self.buf), ctypes.byref(self.size))
        return self

    def text(self):
        _libc.fflush(self.fp)
        return ctypes.string_at(self.buf, self.size.value).decode() if self.size.value else ""

    def __exit__(self, *a):
        _libc.fclose(self.fp)
        _libc.free(ctypes.cast(self.buf, ctypes.c_void_p))


def load(text: bytes, n_original: int, stat: str, marginal: int, flavor: int = OMP, threads: int = 1,
         blocks: int = 1):
    """load_parameters + calc_Parameters.  Returns (problem, stdout text,
    exited) where `exited` tells whether the reference would exit(-1)."""
    p = Problem()
    with _Capture() as cap:
        if lib().lnorm_host_load_text(ctypes.byref(p), text, len(text), n_original, stat.encode()[:1] or b"n",
                                      marginal, cap.fp):
            raise RuntimeError("lnorm_host_load_text failed")
        t, b = ctypes.c_int(threads), ctypes.c_int(blocks)
        rc = lib().lnorm_host_params(ctypes.byref(p), flavor, ctypes.byref(t), ctypes.byref(b), cap.fp)
        out = cap.text()
    return p, out, rc != 0


def print_results(p: Problem, norm: int, strategy, directory: str) -> str:
    s = np.ascontiguousarray(np.asarray(list(strategy) + [0], dtype=np.int32))
    with _Capture() as cap:
        rc = lib().lnorm_host_print_results(ctypes.byref(p), norm, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            directory.encode(), cap.fp)
        out = cap.text()
    if rc:
        raise RuntimeError("lnorm_host_print_results failed")
    return out


def strategy_name(p: Problem) -> str:
    return lib().lnorm_host_strategy_name(ctypes.byref(p)).decode()


def free(p: Problem) -> None:
    lib().lnorm_host_free(ctypes.byref(p))
