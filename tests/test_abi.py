"""CPU-side checks of the C ABI: the built libraries load and export every
function their headers declare.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lnorm_\w+)\s*\(", text)))


def test_hip_library_exports_header(pkg):
    names = _declared("lnorm_hip.h")
    assert "lnorm_hip_calc_Lnorm" in names and len(names) >= 10
    lib = pkg.hip_lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(pkg.HIP_SYMBOLS)
    assert lib.lnorm_hip_abi_version() == 1


def test_host_library_exports_header(pkg):
    if not os.path.exists(os.path.join(ROOT, "include", "lnorm_host.h")):
        pytest.skip("host header not present")
    names = _declared("lnorm_host.h")
    lib = ctypes.CDLL(pkg.HOST_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_engine_fails_loudly_without_library(pkg, monkeypatch):
    monkeypatch.setattr(pkg, "_hip", None)
    monkeypatch.setattr(pkg, "HIP_LIB", "/nonexistent/liblnorm_hip.so")
    with pytest.raises(RuntimeError, match="not built"):
        pkg.hip_lib()
