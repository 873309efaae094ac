"""Shared fixtures: the product module (rbtp_amd, the HIP library's Python view) and the oracle.

GPU tests are marked ``gpu``; everything else runs on the CPU (oracle self-tests, host logic,
C-ABI symbol checks).
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_rbtp():
    if "rbtp_amd" in sys.modules:
        return sys.modules["rbtp_amd"]
    spec = importlib.util.spec_from_file_location(
        "rbtp_amd", os.path.join(ROOT, "rb-thermoparticles_amd", "__init__.py"),
        submodule_search_locations=[os.path.dirname(os.path.join(ROOT, "rb-thermoparticles_amd", "__init__.py"))])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rbtp_amd"] = mod
    try:
        spec.loader.exec_module(mod)
    except Exception:
        del sys.modules["rbtp_amd"]
        raise
    return mod


@pytest.fixture(scope="session")
def rb():
    return load_rbtp()


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle
