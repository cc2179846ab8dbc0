import importlib.util
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
PKG_DIR = os.path.join(ROOT, "l-norms_bruteforce_amd")


def load_pkg():
    """Imports l-norms_bruteforce_amd/ as the package `lnorm_amd`."""
    if "lnorm_amd" in sys.modules:
        return sys.modules["lnorm_amd"]
    spec = importlib.util.spec_from_file_location(
        "lnorm_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lnorm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="session")
def golden_argv():
    """The reference's output for malformed and edge-case argv lines."""
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)["argv_cases"]


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()
