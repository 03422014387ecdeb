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
