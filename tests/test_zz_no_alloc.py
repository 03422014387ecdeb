"""GPU test of the entry point that allocates nothing (mr_clip_histogram_static): the
same counts as mr_clip_histogram on the same device volume, on two calls in a row whose
second has a smaller input (fewer voxels and fewer bins, so counters the first call
left in the shared static scratch would show), and its size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _static_hist(ctx, dptr, n, lo, hi, edges):
    from minirecast import _hip
    e = np.ascontiguousarray(edges, dtype=np.float64)
    hist = np.full(len(e) - 1, -7, dtype=np.int64)
    rc = ctx.lib.mr_clip_histogram_static(ctx.h, ctypes.c_void_p(dptr), int(n), float(lo),
                                          float(hi), len(hist), _hip.ptr(e), _hip.ptr(hist))
    _hip.check(rc)
    return hist


def test_static_clip_histogram_matches_and_allocates_nothing():
    from minirecast import _hip
    print("[no_alloc] mr_clip_histogram_static vs mr_clip_histogram", flush=True)
    rng = np.random.default_rng(77)
    ctx = _hip.Context(0)
    n1 = 3_000_017
    vol = rng.normal(0.3, 1.0, n1).astype(np.float32)
    vol[rng.random(n1) < 0.25] = 0.0                       # zeros are not counted
    vol[:5] = [np.nan, 1e30, -1e30, 0.0, 2.5]              # NaN dropped, far values clipped
    dptr = ctx.alloc(vol.nbytes)
    try:
        ctx.h2d(dptr, vol)
        cases = [(n1, -1.5, 2.0, 8192), (250_003, -0.5, 1.25, 37)]   # second: smaller
        for n, lo, hi, bins in cases:
            edges = np.linspace(lo, hi, bins + 1)
            want = ctx.clip_histogram(None, lo, hi, edges, dptr=dptr, n=n)
            got = _static_hist(ctx, dptr, n, lo, hi, edges)
            np.testing.assert_array_equal(got, want)
            x = vol[:n][vol[:n] != 0].astype(np.float64)
            ref, _ = np.histogram(np.clip(x[~np.isnan(x)], lo, hi), bins, range=(lo, hi))
            np.testing.assert_array_equal(got, ref)
        # an empty volume gives zero counts (the counters are cleared every call)
        edges = np.linspace(0.0, 1.0, 9)
        np.testing.assert_array_equal(_static_hist(ctx, dptr, 0, 0.0, 1.0, edges), 0)
        # beyond the static scratch: refused with a clear error, nothing launched
        big = np.linspace(0.0, 1.0, 8194)
        with pytest.raises(ValueError, match="static scratch"):
            _static_hist(ctx, dptr, 10, 0.0, 1.0, big)
        with pytest.raises(ValueError):
            _static_hist(ctx, dptr, 10, 1.0, 1.0, edges)
        # a second context on the same GPU shares the scratch and gets the same result
        other = _hip.Context(0)
        lo, hi, bins = cases[1][1:]
        edges = np.linspace(lo, hi, bins + 1)
        np.testing.assert_array_equal(_static_hist(other, dptr, cases[1][0], lo, hi, edges),
                                      ctx.clip_histogram(None, lo, hi, edges, dptr=dptr,
                                                         n=cases[1][0]))
        other.close()
    finally:
        ctx.free(dptr)
