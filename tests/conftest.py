"""Test configuration: the ``gpu`` marker, the repository on the path, the oracle built once."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) GPU")
    if "LUMEN_CACHE" not in os.environ:  # the navigator build cache of this session only
        import tempfile

        os.environ["LUMEN_CACHE"] = tempfile.mkdtemp(prefix="lumen-cache-")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip where no GPU is visible (the GPU tier runs them on an MI355X)."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    return True
