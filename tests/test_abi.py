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
    version = int(re.search(r"#define LNORM_HIP_ABI_VERSION (\d+)",
                            open(os.path.join(ROOT, "include", "lnorm_hip.h")).read()).group(1))
    assert lib.lnorm_hip_abi_version() == version


def _struct_fields(header, name):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1] if not decl.startswith("const") else decl.split(None, 2)[2]
        for nm in names.split(","):
            fields.append(re.sub(r"\[.*?\]|[*\s]", "", nm))
    return fields


@pytest.mark.parametrize("cname,pyname", [("lnorm_hip_opts", "Opts"), ("lnorm_hip_result", "Result"),
                                          ("lnorm_hip_item", "Item"), ("lnorm_hip_batch_stats", "BatchStats"),
                                          ("lnorm_hip_tune_entry", "TuneEntry")])
def test_ctypes_structs_match_header(pkg, cname, pyname):
    """The Python views of the ABI structs list the header's fields in order."""
    assert [f for f, _ in getattr(pkg, pyname)._fields_] == _struct_fields("lnorm_hip.h", cname)


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
