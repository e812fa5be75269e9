"""Shared test setup: import paths, the gpu marker and mesh fixtures."""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

# PyTorch's HIP runtime must be the process's runtime before libhbem loads
# (the library then binds to it and shares device pointers with tensors).
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "hornlab-metal-bem_amd"
for p in (str(PKG), str(ROOT / "oracle"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libhbem")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def gpu_available() -> bool:
    try:
        from hornlab_metal_bem.native import discover_native_runtime

        return discover_native_runtime().available
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    """GPU tests fail loudly (no silent fallback) when the device or the
    library is missing on a GPU run."""
    from hornlab_metal_bem.native import discover_native_runtime

    rt = discover_native_runtime()
    if not rt.available:
        pytest.fail("HIP backend unavailable: " + "; ".join(rt.unavailable_reasons))
    return rt
