import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs HIP kernels)")


def _load_pkg():
    path = os.path.join(ROOT, "casm-hella_amd", "hella_amd.py")
    spec = importlib.util.spec_from_file_location("hella_amd", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hella_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def hamd():
    return _load_pkg()


@pytest.fixture(scope="session")
def orc():
    import oracle_bind
    return oracle_bind
