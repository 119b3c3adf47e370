"""Shared test set-up: markers, and loaders for the engine binding and the oracle."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_engine():
    """The ctypes binding of libmce.so (package directory mce-gpu_amd)."""
    if "mce_gpu_amd" in sys.modules:
        return sys.modules["mce_gpu_amd"]
    spec = importlib.util.spec_from_file_location("mce_gpu_amd", os.path.join(ROOT, "mce-gpu_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mce_gpu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    if "mce_oracle" in sys.modules:
        return sys.modules["mce_oracle"]
    spec = importlib.util.spec_from_file_location("mce_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mce_oracle"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def mce():
    return load_engine()


@pytest.fixture(scope="session")
def oracle():
    return load_oracle()
