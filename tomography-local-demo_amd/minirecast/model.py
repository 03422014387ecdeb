"""Threshold of a reconstructed volume (model.py:18-29 otsu_threshold), with the passes
over the volume on the GPU.

The reference computes, with numpy on the host,
    v = vol[vol != 0]; lo, hi = np.percentile(v, [0.1, 99.95])
    hist = np.histogram(np.clip(v, lo, hi), bins, range=(lo, hi))
and then Otsu's between-class variance over the bins.  Here the percentiles (exact
order statistics by radix select) and the histogram come from libminirecast_hip.so;
the 512-bin Otsu scan runs in numpy exactly as in the reference.  The surface
extraction that follows (VTK / scikit-image) stays on the CPU and is not part of this
package.
"""
import numpy as np

from . import recon


def nonzero_percentiles(vol, percent, ctx=None):
    """np.percentile(vol[vol != 0], percent) computed on the GPU (float64 values)."""
    ctx = ctx or recon._default_ctx()
    q = np.true_divide(np.asarray(percent, dtype=np.float64), np.float32(100))
    vals, _ = ctx.nonzero_percentiles(vol, q)
    return vals


def otsu_threshold(vol, bins=512, ctx=None):
    """Otsu's threshold on the non-zero voxels (separates air from bone)."""
    ctx = ctx or recon._default_ctx()
    q = np.true_divide(np.asarray([0.1, 99.95], dtype=np.float64), np.float32(100))
    (lo, hi), m = ctx.nonzero_percentiles(vol, q)
    if lo == hi:
        # every non-zero voxel has one value: numpy widens the range by 0.5 each side
        # and the clipped values all fall in the middle bin
        edges = np.linspace(lo - 0.5, hi + 0.5, bins + 1)
        hist = np.zeros(bins, dtype=np.int64)
        hist[np.searchsorted(edges, lo, side="right") - 1] = m
    else:
        edges = np.linspace(lo, hi, bins + 1)
        hist = ctx.clip_histogram(vol, lo, hi, edges)
    centers = 0.5 * (edges[1:] + edges[:-1])
    w0 = np.cumsum(hist)
    w1 = w0[-1] - w0
    m0 = np.cumsum(hist * centers) / np.maximum(w0, 1)
    m1 = (np.sum(hist * centers) - np.cumsum(hist * centers)) / np.maximum(w1, 1)
    between = w0 * w1 * (m0 - m1) ** 2
    return float(centers[int(np.argmax(between))])
