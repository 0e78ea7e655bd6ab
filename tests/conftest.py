"""Shared test setup: the gpu marker, the package and oracle loaders.

``-m "not gpu"`` runs on any host (oracle vs golden vectors, host logic, the
C ABI exports and device-free hipRTC compiles); ``-m gpu`` runs the parity
tests through the C ABI on a ROCm device.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "sunny-batched-geneig_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE_DATA = "/root/reference/tests/data"

sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    """Import sunny-batched-geneig_amd/ (hyphenated directory) as a package."""
    name = "sunny_batched_geneig_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); runs the HIP kernels")


@pytest.fixture(scope="session")
def geneig():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test selected but no ROCm device is visible")
    return torch.device("cuda:0")
