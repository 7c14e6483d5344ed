import importlib.util
import os
import sys

import pytest

# Load torch's HIP runtime before the propagator library: the library then
# binds to the same libamdhip64 (one runtime per process), so tests that pass
# torch device tensors to the device entry points can use both.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sc17-born_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
if os.path.join(ROOT, "tests") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_pkg_module(name):
    """Import a module of the sc17-born_amd package directory by path."""
    key = "sc17born_" + name
    if key in sys.modules:
        return sys.modules[key]
    spec = importlib.util.spec_from_file_location(key, os.path.join(PKG, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[key] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pyborn():
    return load_pkg_module("pyborn")
