"""Shared fixtures and markers.

``gpu`` tests need a HIP device and the built native library; everything
else runs on the CPU (oracle checks, host logic, ABI/export checks,
hiprtc compilation).
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running test")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available() or os.environ.get("CUBIE_COMPILE_ONLY") == "1":
        return
    skip = pytest.mark.skip(reason="no HIP device in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def julia_golden():
    return dict(np.load(os.path.join(GOLDEN, "julia_lorenz_subset.npz")))


@pytest.fixture(scope="session")
def oracle_module():
    import oracle as orc
    orc.build()
    return orc


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_call(item):
    """In compile-only mode a GPU test stops once its kernels are cached."""
    outcome = yield
    if os.environ.get("CUBIE_COMPILE_ONLY") == "1" and outcome.excinfo is not None:
        from cubie_amd.batchsolving.kernel import CompileOnly
        if isinstance(outcome.excinfo[1], CompileOnly):
            outcome.force_exception(pytest.skip.Exception("kernel compiled (compile-only mode)"))
