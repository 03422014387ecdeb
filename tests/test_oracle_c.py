"""Pins the C restatement of the oracle against the NumPy float64 oracle."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
from oracle import oracle_c as OC


@pytest.mark.parametrize("kind", ["exact", "fast"])
def test_c_filter_rows_matches_numpy(kind):
    rng = np.random.default_rng(0)
    for n in (1, 7, 96, 486):
        x = rng.normal(size=(5, n)).astype(np.float32)
        for w in O.WINDOWS:
            np.testing.assert_allclose(OC.filter_rows(x, 0.3, w, kind), O.filter_rows(x, 0.3, w),
                                       rtol=1e-10, atol=1e-10 * np.abs(x).max() / 0.3)


def test_c_fdk_filter_matches_numpy():
    rng = np.random.default_rng(1)
    V = O.circular_cone_vectors(4, 400.0, 600.0, 0.3, det_shift_u=1.3)
    K = O.cone_constants(V)
    block = rng.uniform(0, 3, size=(9, 4, 40)).astype(np.float32)
    want = O.fdk_filter_block(block, K, "shepp-logan")
    for v in range(4):
        got = OC.fdk_filter_view(block, v, V[v], "shepp-logan")
        np.testing.assert_allclose(got, want[v], rtol=1e-11, atol=1e-11 * np.abs(want).max())


@pytest.mark.parametrize("kind", ["exact", "fast"])
def test_c_bp_matches_numpy(kind):
    rng = np.random.default_rng(2)
    V = O.circular_cone_vectors(7, 40.0, 70.0, 1.1, det_shift_u=0.7)
    K = O.cone_constants(V)
    q = rng.normal(size=(7, 6, 9)).astype(np.float32)
    pts = rng.uniform(-4, 4, size=(300, 3))
    args = (K['s'], K['d'], K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, 0.3)
    want = O.bp_cone(q.astype(np.float64), *args)
    got = OC.bp_cone(q, *args, kind=kind)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-14)
    th = rng.uniform(0, np.pi, 7)
    pargs = (np.cos(th), np.sin(th), 0.9, 1.2, pts, 0.5, 0.3)
    assert rel_l2(OC.bp_parallel(q, *pargs, kind=kind), O.bp_parallel(q.astype(np.float64), *pargs)) < 1e-12
    # per-view float64 accumulation with a COR shift equals bp_cone with D - s*u
    acc = np.zeros(len(pts))
    for k in range(7):
        OC.bp_cone_acc_view(q[k].astype(np.float64), V[k], 0.6, pts, acc)
    D = K['d'] - 0.6 * K['u']
    want = O.bp_cone(q.astype(np.float64), K['s'], D, K['uu'], K['vv'], K['n'], K['sdd'],
                     K['sod'], pts, 1.0)
    np.testing.assert_allclose(acc, want, rtol=1e-12, atol=1e-12)


def test_c_bp_parallel_acc_matches_numpy():
    """The strided float64 parallel-beam accumulation equals the NumPy oracle's
    bp_parallel (without its scale), on a [rows, views, cols] array read as
    [views, rows, cols] through its strides, with a column shift and points whose
    taps leave the detector on every side."""
    rng = np.random.default_rng(5)
    R, A, C = 6, 11, 13
    q = rng.normal(size=(R, A, C))
    th = rng.uniform(0, np.pi, A)
    pts = rng.uniform(-9, 9, size=(400, 3))
    pts[:, 2] = rng.uniform(-4.5, 4.5, 400)
    view = np.transpose(q, (1, 0, 2))                 # [A, R, C], not contiguous
    assert not view.flags.c_contiguous
    for shift in (0.0, -1.75, 2.5):
        acc = np.full(len(pts), 0.5)
        OC.bp_parallel_acc(view, np.cos(th), np.sin(th), 0.9, 1.2, shift, pts, acc)
        want = O.bp_parallel(np.ascontiguousarray(view), np.cos(th), np.sin(th), 0.9, 1.2, pts,
                             1.0, shift)
        np.testing.assert_allclose(acc - 0.5, want, rtol=1e-12, atol=1e-12)
