"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU oracle work")


@pytest.fixture(scope="session", autouse=True)
def _gpu_runtime(request):
    """When the session runs GPU tests, initialise the HIP runtime (through torch,
    whose runtime libpbrt.so then shares) before any test loads the library or starts
    a GPU child process."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
    yield


@pytest.fixture(scope="session")
def pbrt_lib():
    from vulkanpbrt_amd import build, capi
    if not capi.LIB_PATH.exists():
        build.build()
    return capi.lib()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    if not oracle.LIB_PATH.exists():
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch
