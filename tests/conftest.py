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


def entropy_score(img):
    """estimate_cor's score of one trial image (recon.py:509-511), numpy in float64."""
    img = np.asarray(img, np.float64)
    hist, _ = np.histogram(img[img != 0], bins=128)
    prob = hist[hist > 0] / max(hist.sum(), 1)
    return float(np.sum(prob * np.log(prob)))


def cor_trial_images(rec, shifts, n, z=0.0):
    """The trial images of rec.estimate_cor(shifts, n, z): the same task bytes that its
    one scoring request back-projects (axial planes, outputs one after another), so
    the images are the ones its device scores were computed from, bit for bit."""
    from minirecast import _hip, recon
    T = np.array([recon.plane_task(rec.box, (0.0, 0.0, z), (0.0, 0.0, 1.0), n,
                                   cor_shift=float(s), out_offset=i * n * n)
                  for i, s in enumerate(shifts)], dtype=_hip.TASK_DTYPE)
    flat = rec._ctx.bp_tasks(T, rec.n, len(T) * n * n)
    return flat.reshape(len(T), n, n)
