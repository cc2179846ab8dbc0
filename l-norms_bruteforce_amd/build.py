"""Builds the engine in-tree:

  l-norms_bruteforce_amd/lib/liblnorm_hip.so   HIP kernels + C ABI (include/lnorm_hip.h)
  l-norms_bruteforce_amd/lib/liblnorm_host.so  reader / preprocessing / layout / printer
                                                 (include/lnorm_host.h)
  l-norms_bruteforce_amd/bin/L_HIP             the CLI (L_CUDA.cu's argv), linked to both

hipcc cross-compiles gfx950 code without a GPU.  Every HIP translation unit
is compiled with packed-fp32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
OBJ = os.path.join(PKG, "_build")
INC = os.path.join(ROOT, "include")

ARCH = os.environ.get("LNORM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
# the units in decreasing compile time (the pool starts them in this order, so
# the longest ones do not wait for a free worker)
HIP_SOURCES = ["lnorm_wide_tab32.hip", "lnorm_dary32.hip", "lnorm_dary_hi.hip", "lnorm_wide_tab32_hi.hip",
               "lnorm_dary32_hi.hip", "lnorm_dary.hip", "lnorm_wide_tab.hip", "lnorm_plain.hip",
               "lnorm_blocked.hip", "lnorm_general.hip", "lnorm_batch.hip", "lnorm_batch_tab.hip",
               "lnorm_batch_tab_hi.hip", "lnorm_int64.hip", "lnorm_wide.hip", "lnorm_tab_bin.hip",
               "lnorm_tab32_bin.hip", "lnorm_wide_tb32.hip", "lnorm_colwalk.hip", "lnorm_peak.hip", "lnorm_api.hip"]
# -simplifycfg-sink-common=false: the d-ary walks branch (uniformly) into one
# move per (from, to) label pair; sinking those near-identical bodies into a
# shared tail would index the label-vector array at run time and push it into
# scratch memory (seen for d >= 4).
# -structurizecfg-skip-uniform-regions=true: those branches are uniform (the
# moved row and its labels are the same for the whole wave), so they need no
# exec-mask structurisation; structurised, the merge of the per-move paths cost
# one v_mov per state register per step (187 copies per step in the d = 3,
# 16-word packed walk, against 254 useful operations).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-simplifycfg-sink-common=false",
             "-mllvm", "-structurizecfg-skip-uniform-regions=true",
             "-I" + CSRC, "-I" + INC]
HOST_SOURCES = ["lnorm_host.c"]
C_FLAGS = ["-O2", "-std=c11", "-fPIC", "-Wall", "-Wextra", "-I" + INC]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s%s" % (p.returncode, " ".join(cmd), p.stdout, p.stderr))
    return p


def _includes(path: str, seen: set) -> set:
    """path and every file it includes with #include "...", transitively."""
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("#include") and '"' in s:
                _includes(os.path.join(os.path.dirname(path), s.split('"')[1]), seen)
    return seen


def _stale(obj: str, src: str) -> bool:
    """True when obj is missing or older than src or a header it includes."""
    if not os.path.exists(obj):
        return True
    deps = _includes(src, set())
    return os.path.getmtime(obj) < max(os.path.getmtime(d) for d in deps)


def build_hip(jobs: int = 8, incremental: bool = False) -> str:
    """Compiles every HIP translation unit (incremental: only those older
    than their source or a header) and links lib/liblnorm_hip.so."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        if incremental and not _stale(obj, os.path.join(CSRC, src)):
            continue
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    out = os.path.join(LIB, "liblnorm_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host() -> tuple:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    lib = os.path.join(LIB, "liblnorm_host.so")
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES]
    _run(["gcc", *C_FLAGS, "-shared", *srcs, "-o", lib + ".tmp", "-lm"])
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(BIN, "L_HIP")
    _run(["gcc", *C_FLAGS, os.path.join(HOST, "L_HIP.c"), *srcs, "-o", exe + ".tmp",
          "-L" + LIB, "-llnorm_hip", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    os.replace(exe + ".tmp", exe)
    return lib, exe


def build(jobs: int = 8, incremental: bool = False) -> None:
    build_hip(jobs, incremental)
    if os.path.exists(os.path.join(HOST, "lnorm_host.c")):
        build_host()


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "--incremental"]
    build(int(args[0]) if args else 8, incremental="--incremental" in sys.argv[1:])
    print("built", LIB, BIN)
