"""Per-projection preprocessing applied on arrival (preprocess.py of the reference),
computed on the GPU.

Blocks are [rows, k, cols] float32.  ``apply`` is the identity when both options are
off (every benchmark workload).  Inside SliceReconstructor the preprocessing runs on
the device-resident block before filtering (mr_set_preprocess); these module-level
functions serve callers that preprocess host arrays themselves.
"""
import numpy as np


def _ctx():
    from .recon import _default_ctx
    return _default_ctx()


def gaussian_blur(block, sigma):
    """Edge-padded separable Gaussian low-pass, sigma in detector pixels
    (preprocess.py:8-28): rows, then columns, float32 between the passes."""
    if sigma <= 0:
        return block
    return _ctx().preprocess(np.asarray(block, np.float32), sigma, 0.0)


def rotate_projections(block, degrees):
    """Bilinear rotation by ``degrees`` about the detector centre, zero outside
    (preprocess.py:31-56)."""
    if abs(degrees) < 1e-9:
        return block
    return _ctx().preprocess(np.asarray(block, np.float32), 0.0, -float(degrees))


def apply(block, gaussian=0.0, tilt=0.0):
    """Tilt correction (rotation by -tilt), then Gaussian (preprocess.py:59-64)."""
    if not tilt and not gaussian:
        return block
    return _ctx().preprocess(np.asarray(block, np.float32), gaussian if gaussian else 0.0,
                             tilt if tilt else 0.0)
