"""Shared pytest configuration.

Markers: ``gpu`` tests need an AMD GPU and the built libminirecast_hip.so; everything
else runs on the CPU.  The repository root (for ``oracle``) and the package
directory (for ``minirecast``) are put on sys.path.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tomography-local-demo_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_composition.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


def gpu_available():
    try:
        from minirecast import _hip
        return _hip.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN)


def rel_l2(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))
