"""Pure-Python restatement of the reference host pipeline (functions.h,
L_OpenMP.c, L_CUDA.cu argument and parameter handling).

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker for the product's C
host library (reader, preprocessing, layout, parameters, printer).  The
product never imports it.

Pinned by the golden vectors in tests/golden/ (stdout and strategy files of
the reference L_OpenMP built from source) and by the norm-preservation
invariant of preprocessing (tests/test_host_oracle.py).

Integers follow the reference's int32 (int_type, functions.h:5): merged
rows/columns wrap modulo 2^32 like the C code does on x86.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

LENGTH = 4096          # functions.h:2
RANK_OF_NORM = 10      # functions.h:3
NUM_OF_BITS = 64       # functions.h:4


def _double_to_u64(d: float) -> int:
    """(unsigned long long) of a double as gcc emits it on x86-64: values
    from 2^63 go through cvttsd2si(d - 2^63) ^ 2^63, which is 0 from 2^64 up."""
    if d < 9223372036854775808.0:
        return int(d)
    d -= 9223372036854775808.0
    if d >= 9223372036854775808.0:
        return 0
    return int(d) ^ (1 << 63)


def _wrap32(v: int) -> int:
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v & 0x80000000 else v


class HostExit(Exception):
    """The reference calls exit(-1); carries the lines printed so far."""

    def __init__(self, lines):
        super().__init__("exit(-1)")
        self.lines = lines


@dataclass
class Problem:
    lines: list = field(default_factory=list)   # stdout so far
    n_original: int = 1
    stat: str = "n"
    marginal_req: int = 0        # first->marginal
    iRows: int = 0
    iCols: int = 0
    matrix: list = field(default_factory=list)
    original_r: list = field(default_factory=list)
    original_c: list = field(default_factory=list)
    R: int = 0                   # second->iRows_reduced
    C: int = 0                   # second->iCols_reduced
    marginal: int = 0            # second->marginal
    n: int = 1                   # second->n
    maxRows: int = 63
    total: int = 0
    copyNum: int = 0
    A: list = field(default_factory=list)       # R*C row-major (mtx_as_vec)
    original: list = field(default_factory=list)
    original_length: int = 0
    transposed: bool = False


# ---- reader (functions.h:40-79) -------------------------------------------
_INT_RE = re.compile(rb"[+-]?[0-9]+")
_FLT_RE = re.compile(rb"[+-]?(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?")


def _scan_int(tok: bytes):
    m = _INT_RE.match(tok)
    if not m:
        return None
    v = int(m.group(0))
    v = max(min(v, (1 << 63) - 1), -(1 << 63))   # strtol saturation, then (int)
    return _wrap32(v)


def _scan_double(tok: bytes):
    m = _FLT_RE.match(tok)
    if not m:
        return None
    return float(m.group(0))


def _double_to_int(d: float) -> int:
    # C (int) conversion truncates toward zero; x86 yields INT_MIN out of range.
    if d != d or d >= 2147483648.0 or d <= -2147483649.0:
        return -(1 << 31)
    return int(d)


def matrix_read(data: bytes):
    """Returns (rows as lists of int, iCols).  A 0xFF byte ends the file, as
    `char g = fgetc(fp)` compares equal to EOF there (functions.h:52,74)."""
    rows, row, tok = [], [], bytearray()
    has_e = False
    value, dvalue = 0, 0.0
    icols = 0
    n = len(data)
    i = 0
    while True:
        g = data[i] if i < n else None
        i += 1
        if g == 0xFF:
            g = None
        if g is not None and (48 <= g <= 57 or g in b"eE.+-"):
            tok.append(g)
            if g in b"eE":
                has_e = True
        else:
            if tok:
                if has_e:
                    d = _scan_double(bytes(tok))
                    if d is not None:
                        dvalue = d
                    value = _double_to_int(dvalue)
                    has_e = False
                else:
                    v = _scan_int(bytes(tok))
                    if v is not None:
                        value = v
                row.append(value)
                tok = bytearray()
            if (g is None or g == 10) and row:
                icols = len(row)
                rows.append(row)
                row = []
        if g is None:
            break
    return rows, icols


# ---- preprocessing (functions.h:81-289) -----------------------------------
def _flag(stat: str, chars: str) -> bool:
    return stat in chars


def eliminate_zero_rows_cols(p: Problem):
    M = p.matrix
    p.original_r = [0] * p.iRows
    p.original_c = [0] * p.iCols
    p.R, p.C, p.marginal = p.iRows, p.iCols, p.marginal_req
    if not _flag(p.stat, "yYrRcC"):
        return
    cr = 0
    for i in range(p.iRows):
        if all(M[i][k] == 0 for k in range(p.iCols)):
            p.original_r[i] = p.iRows
            if p.marginal_req == 0 or i > 0:
                cr += 1
                p.lines.append("row %d entries are zeros." % (i + 1))
    p.R = p.iRows - cr
    cc = 0
    for i in range(p.iCols):
        if all(M[k][i] == 0 for k in range(p.iRows)):
            p.original_c[i] = p.iCols
            if p.marginal_req == 0 or i > 0:
                cc += 1
                p.lines.append("col %d entries are zeros." % (i + 1))
    p.C = p.iCols - cc
    if p.marginal_req == 1:
        if p.original_r[0] == p.iRows and p.original_c[0] == p.iCols:
            p.marginal = 0
            p.R -= 1
            p.C -= 1
            p.lines.append("The entries of the first row and column are zeros.")
        else:
            p.original_r[0] = 0
            p.original_c[0] = 0


def _proportional(u, v):
    """(mult, index) as the nested loops of functions.h:248-255 compute them."""
    L = len(u)
    k = 0
    while k < L and u[k] == 0 and v[k] == 0:
        k += 1
    if k >= L:          # cannot happen for non-zero vectors
        return 0, k
    mult = 1
    if (u[k] == 0) != (v[k] == 0):
        mult = 0
    index = k
    while k < L - 1 and mult:
        mult = 1 if u[index] * v[k + 1] == v[index] * u[k + 1] else 0
        k += 1
    return mult, index


def calc_reduce_matrix_rows(p: Problem):
    M = p.matrix
    cr = 0
    for i in range(p.marginal_req, p.iRows):
        if p.original_r[i] != 0:
            continue
        for j in range(i + 1, p.iRows):
            if p.original_r[j] != 0:
                continue
            mult, idx = _proportional(M[i], M[j])
            if mult == 0:
                p.original_r[j] = 0
            elif M[i][idx] * M[j][idx] < 0:
                if p.n_original == 1:
                    p.original_r[j] = -(i + 1)
                    cr += 1
                    M[i] = [_wrap32(a - b) for a, b in zip(M[i], M[j])]
                    p.lines.append("row %d was subtracted from %d" % (j + 1, i + 1))
            else:
                p.original_r[j] = i + 1
                cr += 1
                M[i] = [_wrap32(a + b) for a, b in zip(M[i], M[j])]
                p.lines.append("row %d was added to %d" % (j + 1, i + 1))
    p.R -= cr


def _col(M, j, nrows):
    return [M[r][j] for r in range(nrows)]


def _merge_col(p: Problem, i, j, sign):
    for r in range(p.iRows):
        p.matrix[r][i] = _wrap32(p.matrix[r][i] + sign * p.matrix[r][j])


def calc_reduce_matrix_cols(p: Problem):
    M = p.matrix
    cc = 0
    for i in range(p.marginal_req, p.iCols):
        if p.original_c[i] != 0:
            continue
        for j in range(i + 1, p.iCols):
            if p.original_c[j] != 0:
                continue
            mult, idx = _proportional(_col(M, i, p.iRows), _col(M, j, p.iRows))
            if mult == 0:
                p.original_c[j] = 0
            elif M[idx][i] * M[idx][j] < 0:
                p.original_c[j] = -(i + 1)
                cc += 1
                _merge_col(p, i, j, -1)
                p.lines.append("col %d was subtracted from %d" % (j + 1, i + 1))
            else:
                p.original_c[j] = i + 1
                cc += 1
                _merge_col(p, i, j, +1)
                p.lines.append("col %d was added to %d" % (j + 1, i + 1))
    p.C -= cc


def calc_reduce_matrix_cols_sign(p: Problem):
    M = p.matrix
    sign = [0] * p.iCols
    cnt = 0
    for i in range(p.iCols):
        if p.original_c[i] != 0:
            continue
        pos = any(M[k][i] > 0 for k in range(p.iRows))
        neg = any(M[k][i] < 0 for k in range(p.iRows))
        if pos and neg:
            sign[i] = 0
        elif neg:
            sign[i] = -1
            cnt += 1
        else:
            sign[i] = 1
            cnt += 1
    if cnt > 1:
        p.C = p.C - cnt + 1
        index = next(i for i in range(p.iCols) if sign[i] != 0)
        for i in range(index + 1, p.iCols):
            if p.original_c[i] != 0:
                continue
            mult = sign[index] * sign[i]
            if mult < 0:
                p.original_c[i] = -(index + 1)
                _merge_col(p, index, i, -1)
                p.lines.append("col %d was subtracted from %d" % (i + 1, index + 1))
            if mult > 0:
                p.original_c[i] = index + 1
                _merge_col(p, index, i, +1)
                p.lines.append("col %d was added to %d" % (i + 1, index + 1))


def compact(p: Problem):
    """delete_rows/delete_cols (functions.h:153-191): a stable compaction of
    the rows/columns whose original_r/original_c entry is zero."""
    keep_r = [i for i in range(p.iRows) if p.original_r[i] == 0][: p.R]
    keep_c = [j for j in range(p.iCols) if p.original_c[j] == 0][: p.C]
    return keep_r, keep_c


def convert_mtx_to_vec(p: Problem, keep_r, keep_c):
    """functions.h:291-329: row-major layout, L_1 transpose when rows > cols,
    marginal row swapped to the last position."""
    M = p.matrix
    sub = [[M[i][j] for j in keep_c] for i in keep_r]
    if p.n_original == 1 and p.R > p.C:
        sub = [list(col) for col in zip(*sub)] if sub else []
        p.R, p.C = p.C, p.R
        p.original, p.original_length = p.original_c, p.iCols
        p.transposed = True
        p.lines.append("Matrix was transposed.")
    else:
        p.original, p.original_length = p.original_r, p.iRows
    if p.marginal == 1 and p.R > 0:
        sub[0], sub[p.R - 1] = sub[p.R - 1], sub[0]
    p.A = [v for row in sub for v in row]


def calc_parameters(p: Problem, flavor: str, threads: int = 1, blocks: int = 1):
    """calc_Parameters: L_OpenMP.c:344-376 ('omp') or L_CUDA.cu:242-279 ('cuda')."""
    if p.n_original < p.R or p.n_original == 1:
        p.n = p.n_original
    else:
        p.n = p.R if p.R > 1 else 2
        p.lines.append("The preprocessed matrix has a number of rows (%d) less than or equal "
                       "to the order of the L norm (%d)." % (p.R, p.n_original))
    if p.n <= 2:
        p.total = 0 if p.R == 0 else 1 << (p.R - 1)
        p.maxRows = NUM_OF_BITS - 1
    else:
        p.total = 0 if p.R == 0 else _double_to_u64(math.pow(p.n, p.R - 1))
        p.maxRows = int(math.floor(NUM_OF_BITS / math.log2(p.n)) + 1)
    if flavor == "omp":
        p.copyNum = min(threads, p.total)
        p.lines.append("number of threads: %d" % p.copyNum)
    else:
        t = p.total if threads > p.total else threads
        b = 0 if t == 0 else (blocks if p.total // t > blocks else p.total // t)
        p.copyNum = b * t
        p.lines.append("num_ofBlock: %d, num_ofThread in block: %d" % (b, t))
    p.lines.append("maximum length of strategy vector: %d" % p.maxRows)
    if p.maxRows < p.R:
        p.lines.append("Matrix is too large. The matrix has %d rows after reduction. The maximum "
                       "number of rows cannot be more than %d." % (p.R, p.maxRows))
        raise HostExit(p.lines)
    if p.C > LENGTH:
        p.lines.append("Matrix is too large. The length variable %d should be greater or equal "
                       "than %d." % (LENGTH, p.C))
        raise HostExit(p.lines)


def load(data: bytes, n_original: int, stat: str, marginal: int, flavor="omp",
         threads=1, blocks=1) -> Problem:
    """load_parameters (L_OpenMP.c:378-387 / L_CUDA.cu:414-423)."""
    p = Problem(n_original=n_original, stat=stat, marginal_req=marginal)
    rows, icols = matrix_read(data)
    p.iRows, p.iCols = len(rows), icols
    p.matrix = [(r + [0] * icols)[:icols] if len(r) < icols else list(r) for r in rows]
    p.lines.append("rows: %d, cols: %d" % (p.iRows, p.iCols))
    eliminate_zero_rows_cols(p)
    if _flag(stat, "yYrR"):
        calc_reduce_matrix_rows(p)
    if _flag(stat, "yYcC"):
        calc_reduce_matrix_cols(p)
        if n_original > 1:
            calc_reduce_matrix_cols_sign(p)
    keep_r, keep_c = compact(p) if _flag(stat, "yYrRcC") else (list(range(p.R)), list(range(p.C)))
    convert_mtx_to_vec(p, keep_r, keep_c)
    calc_parameters(p, flavor, threads, blocks)
    return p


_USAGE = {"omp": "Use the following command: ./L_OpenMP number_of_threads filename_of_matrix order_of_the_L_norm",
          "cuda": "Use the following command: ./L_CUDA number_of_blocks number_of_threads_per_block "
                  "filename_of_matrix order_of_the_L_norm"}


def _sscanf_d(text, prev):
    """sscanf(text, "%d", &t): (items converted, t).  Leading white space,
    then an optional sign and decimal digits; nothing converted leaves t as
    it was; an empty or all-blank field is EOF (-1)."""
    m = re.match(r"\s*([+-]?\d+)", text)
    if m:
        v = max(min(int(m.group(1)), (1 << 63) - 1), -(1 << 63))   # strtol saturation, then (int)
        return 1, _wrap32(v)
    return (-1 if text.strip() == "" else 0), prev


def argv_error(argv, flavor="omp", file_exists=True, max_threads_per_block=1024):
    """The line arguments_OMP (L_OpenMP.c:299-342) or arguments_CUDA
    (L_CUDA.cu:361-412) prints before exit(-1), or None when the argv passes
    every check.  A field is refused when sscanf converts nothing (sd == 0)
    or the value is below 1; sd == EOF keeps the previous field's value."""
    msg = _USAGE[flavor]
    off = 0 if flavor == "omp" else 1
    if len(argv) < 4 + off:
        return "Incorrect number of input arguments. %s" % msg
    t = 0
    if flavor == "cuda":
        sd, t = _sscanf_d(argv[1], t)
        if sd == 0 or t < 1:
            return "Please make sure that the number of blocks is a positive integer. %s" % msg
    sd, t = _sscanf_d(argv[1 + off], t)
    if sd == 0 or t < 1:
        return "Please make sure that the number of threads is a positive integer. %s" % msg
    threads = t
    if not file_exists:
        return "Please make sure that the file containig the matrix exists within this directory. %s" % msg
    sd, t = _sscanf_d(argv[3 + off], t)
    if sd == 0 or t < 1:
        return "Please make sure that the order of the L norm is a positive integer. %s" % msg
    if t > RANK_OF_NORM:
        return ("The order of the L norm is too large. Please increase the RANK_OF_NORM variable in the code to %d "
                "and compile and run it again." % t)
    if flavor == "cuda" and threads > max_threads_per_block:
        return "The maximum number of threads per block cannot be more than %d." % max_threads_per_block
    return None


def parse_args(argv, flavor="omp"):
    """arguments_OMP (L_OpenMP.c:299-342) / arguments_CUDA (L_CUDA.cu:361-412)
    for well-formed argv: returns (threads, blocks, file, n_original, stat, marginal)."""
    off = 1 if flavor == "omp" else 2
    if flavor == "omp":
        threads, blocks = int(argv[1]), 1
    else:
        blocks, threads = int(argv[1]), int(argv[2])
    fname = argv[off + 1]
    n_original = int(argv[off + 2])
    rest = argv[off + 3:]
    stat = rest[0][0] if rest else "n"
    if not rest or n_original > 1:
        marginal = 0
    elif len(rest) == 1:
        marginal = 1 if rest[0][0] in "mM" else 0
    else:
        marginal = 1 if rest[1][0] in "mM" else 0
    return threads, blocks, fname, n_original, stat, marginal


def result_lines(p: Problem, norm: int, strategy):
    """print_results (functions.h:371-412): the norm line, the strategy file
    name and its lines.  `strategy` is second->strategy (R-1 labels)."""
    if p.n_original == 1 and p.marginal_req == 1:
        line, fname = "LM is: %d" % norm, "strategy_LM.txt"
    else:
        line, fname = "L%d is: %d" % (p.n_original, norm), "strategy_L%d.txt" % p.n_original
    out = []
    vals = [0] * p.original_length
    j = 0
    for i in range(p.original_length):
        o = p.original[i]
        if o == 0:
            if p.marginal == 1 and p.n == 1:
                val = 1 if j == 0 else (strategy[0] if j == p.R - 1 else strategy[j])
            else:
                val = (1 if p.n == 1 else 0) if j == p.R - 1 else strategy[j]
            j += 1
        elif o == p.original_length:
            val = 1 if p.n == 1 else 0
        elif o < 0:
            val = (vals[-o - 1] + 1) & 1
        else:
            val = vals[o - 1]
        vals[i] = val
        out.append("%d" % (2 * val - 1 if p.n == 1 else val))
    return line, fname, out
