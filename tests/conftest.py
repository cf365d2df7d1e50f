"""Shared test setup: import paths, the `gpu` marker, seeded workloads."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "mo-dmft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# libmoed links the system ROCm HIP runtime and PyTorch-ROCm bundles its own under the same soname;
# whichever is loaded first serves the whole process, and torch fails to initialise its GPU ("No
# HIP GPUs are available") when libmoed's copy came first.  Tests that use torch tensors beside
# libmoed therefore need torch loaded before the first libmoed call, whatever subset of the suite
# runs.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libmoed kernels)")


@pytest.fixture(scope="session")
def configs():
    from moed.model import load_configs
    return {c.name: c for c in load_configs()}


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
