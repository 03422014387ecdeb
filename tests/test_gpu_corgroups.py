"""GPU tests of the shift-group kernel (kernel 10) of rotation-centre searches.

Requests of a parallel geometry whose tasks are axial planes that differ only in their
cor_shift, on a grid of 1/q px (q = 1, 2, 4), run one kernel per group of up to 6q + 1
shifts: per point and view the column is formed once, 8 detector values of the slice's
row (interpolated between its two ring rows beforehand) are loaded once, and every
shift's lerp comes from them.  Checked, for f32 and f16 rings:
- such requests run kernel 10 (mr_last_bp_kernel), other requests (shifts off a grid,
  oblique planes, an explicit kernel choice) the plane tiles;
- every output equals the plane tiles' (kernel 6, the default of other parallel
  requests) to float32 rounding, and sampled pixels the float64 oracle (f32 <= 1e-5,
  f16 <= 1e-3 relative L2), recon.py:175-191 with the shift added to the column;
- grids of 1, 0.5 and 0.25 px, 41 shifts (two groups), gaps in a grid, planes of sizes
  that are not multiples of the tile, a plane above the detector (every tap zero), a view
  count below the stored views;
- a plan of the trial images scores them on the device (one graph) like numpy's
  histogram entropy of the plan's own images (estimate_cor's formula, recon.py:509-511),
  also after the stored views were reset and fewer pushed again;
- a bf16 ring; a request mixing a search with an oblique plane runs the plane tiles.
"""
import numpy as np
import pytest

from conftest import entropy_score, rel_l2

pytestmark = pytest.mark.gpu

A, R, C, PU, PV = 180, 20, 300, 1.0, 1.0
WINDOW = "hann"
# axial planes: size n, height z (rows at z / pv + (R - 1) / 2: fractional, and one past
# the detector's last row)
PLANES = [dict(n=100, z=2.3), dict(n=96, z=-4.6), dict(n=64, z=0.0)]
SHIFTS = np.arange(-3.0, 3.01, 0.25)


def _tasks(box, planes, shifts, size=None):
    from minirecast import _hip, recon
    T, off = [], 0
    for p in planes:
        for s in shifts:
            kw = dict(size=size) if size else {}
            T.append(recon.plane_task(box, (0.0, 0.0, p['z']), (0.0, 0.0, 1.0), p['n'],
                                      cor_shift=float(s), out_offset=off, **kw))
            off += p['n'] ** 2
    return np.array(T, dtype=_hip.TASK_DTYPE), off


@pytest.fixture(scope="module")
def cg():
    from minirecast import _hip, phantom
    from minirecast.geometry import ParallelGeometry
    from oracle import oracle_c as OC
    g = ParallelGeometry(np.arange(A) * np.pi / A, R, C, PU, PV)
    box = g.volume_box()
    ph = phantom.random_ellipsoids(8, 9, box, (0.1, 1.0))
    proj = phantom.project_parallel(ph, g)              # [R, A, C]
    ctx = {}
    for st in ("f32", "f16"):
        for mode in ("dflt", "tiles"):
            c = _hip.Context(0)
            c.set_parallel(g.angles, R, C, PU, PV, WINDOW, st)
            c.push(np.arange(A), proj)
            if mode == "tiles":
                c.set_bp_config(6, 4, 0, 2)
            ctx[st, mode] = c
    q64 = OC.filter_rows(proj, PU, WINDOW)
    views = np.ascontiguousarray(np.transpose(q64, (1, 0, 2)))
    return dict(g=g, box=box, ctx=ctx, views=views, proj=proj, OC=OC)


def _both(cg, st, T, n_views, out_len):
    got = cg["ctx"][st, "dflt"].bp_tasks(T, n_views, out_len)
    kern = cg["ctx"][st, "dflt"].last_bp_kernel()
    ref = cg["ctx"][st, "tiles"].bp_tasks(T, n_views, out_len)
    assert cg["ctx"][st, "tiles"].last_bp_kernel() == 6
    return got, ref, kern


def _close(got, ref):
    d = rel_l2(got.astype(np.float64), ref.astype(np.float64))
    dmax = float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))
    np.testing.assert_array_equal(got == 0, ref == 0)
    assert d < 1e-6 and dmax < 1e-5, (d, dmax)
    return d, dmax


@pytest.mark.parametrize("st,tol", [("f32", 1e-5), ("f16", 1e-3)])
def test_cor_groups_vs_tiles_and_oracle(cg, st, tol):
    from minirecast import recon
    T, out_len = _tasks(cg["box"], PLANES, SHIFTS)
    got, ref, kern = _both(cg, st, T, A, out_len)
    assert kern == 10, kern
    d, dmax = _close(got, ref)
    # the float64 oracle at sampled in-box pixels of shifts -3, 0, +1.25, +3
    g, OC = cg["g"], cg["OC"]
    rng = np.random.default_rng(10)
    errs = []
    off = 0
    for p in PLANES:
        pts, mask = recon.slice_points(cg["box"], (0.0, 0.0, p['z']), (0.0, 0.0, 1.0), p['n'])
        sel = np.sort(rng.choice(np.flatnonzero(mask.ravel()), size=200, replace=False))
        for s in (-3.0, 0.0, 1.25, 3.0):
            k = int(np.flatnonzero(np.isclose(SHIFTS, s))[0])
            acc = np.zeros(len(sel))
            OC.bp_parallel_acc(cg["views"], np.cos(g.angles), np.sin(g.angles), PU, PV, s,
                               pts.reshape(-1, 3)[sel], acc)
            img = got[off + k * p['n'] ** 2: off + (k + 1) * p['n'] ** 2]
            errs.append(rel_l2(img[sel].astype(np.float64), acc * (np.pi / A)))
        off += len(SHIFTS) * p['n'] ** 2
    print(f"{st}: shift groups vs plane tiles rel-L2 {d:.2e} (max {dmax:.2e}); vs float64 "
          f"worst {max(errs):.2e}")
    assert max(errs) < tol, errs


@pytest.mark.parametrize("shifts,grouped", [
    (np.arange(-5.0, 5.01, 1.0), True),            # q = 1: 11 shifts, groups of <= 7
    (np.arange(-5.0, 5.01, 0.5), True),            # q = 2: 21 shifts, groups of <= 13
    (np.arange(-5.0, 5.01, 0.25), True),           # q = 4: 41 shifts, groups of 25 + 16
    (np.array([-1.0, -0.75, 0.0, 0.25, 1.0, 2.5, 2.75, 3.0]), True),   # gaps in a grid
    (np.array([0.0, 0.3, 0.7, 1.1, 1.6]), False),  # no grid of 1, 1/2 or 1/4 px
])
def test_cor_groups_shift_grids(cg, shifts, grouped):
    T, out_len = _tasks(cg["box"], PLANES[:2], shifts)
    got, ref, kern = _both(cg, "f32", T, A, out_len)
    assert (kern == 10) == grouped, kern
    _close(got, ref)


def test_cor_groups_partial_views_plane_above_detector_and_oblique(cg):
    from minirecast import _hip, recon
    # a view count below the stored views; a plane above the detector rows
    planes = PLANES[:1] + [dict(n=40, z=40.0)]
    T, out_len = _tasks(cg["box"], planes, SHIFTS)
    T['mask_box'][len(SHIFTS):] = 0                     # every point computed: rows clamped
    got, ref, kern = _both(cg, "f32", T, 77, out_len)
    assert kern == 10
    _close(got, ref)
    n0 = planes[0]['n'] ** 2 * len(SHIFTS)
    assert not got[n0:].any()
    # oblique planes are not grouped (their row changes along the plane)
    box = cg["box"]
    T = np.array([recon.plane_task(box, (0.0, 0.0, 0.0), (0.1, 0.0, 1.0), 64, cor_shift=float(s),
                                   out_offset=i * 64 * 64) for i, s in enumerate(SHIFTS)],
                 dtype=_hip.TASK_DTYPE)
    got, ref, kern = _both(cg, "f32", T, A, len(T) * 64 * 64)
    assert kern != 10
    _close(got, ref)


@pytest.mark.parametrize("st", ["f32", "f16"])
def test_cor_groups_plan_scores(cg, st):
    """A plan of the trial images of two slices: the captured graph runs the shift-group
    kernel and scores the images; the scores equal numpy's entropy of the plan's own
    images, the images equal one request's, and a second replay gives the same."""
    from minirecast import _hip
    T, out_len = _tasks(cg["box"], PLANES[:2], SHIFTS)
    ctx = cg["ctx"][st, "dflt"]
    plan = _hip.Plan(ctx, T, out_len)
    try:
        scores = plan.cor_scores(A)
        assert ctx.last_bp_kernel() == 10
        imgs = ctx.d2h(np.empty(out_len, np.float32), plan.output_ptr())
        again = plan.cor_scores(A)
    finally:
        plan.close()
    np.testing.assert_array_equal(scores, again)
    one = ctx.bp_tasks(T, A, out_len)
    assert rel_l2(imgs.astype(np.float64), one.astype(np.float64)) < 1e-7
    offs = np.concatenate([[0], np.cumsum([t['np'] * t['ns'] for t in T])])
    host = np.array([entropy_score(imgs[offs[i]:offs[i + 1]]) for i in range(len(T))])
    np.testing.assert_allclose(scores, host, rtol=1e-12, atol=0)


def test_cor_groups_bf16_and_mixed_requests(cg):
    """A bf16 ring through the shift groups; a request mixing a shift search with an
    oblique plane runs the plane tiles for all of it (one kernel per request)."""
    from minirecast import _hip, recon
    g = cg["g"]
    c16 = _hip.Context(0)
    c16.set_parallel(g.angles, R, C, PU, PV, WINDOW, "bf16")
    c16.push(np.arange(A), cg["proj"])
    T, out_len = _tasks(cg["box"], PLANES[:1], SHIFTS)
    got = c16.bp_tasks(T, A, out_len)
    assert c16.last_bp_kernel() == 10
    c16.set_bp_config(6, 4, 0, 2)
    ref = c16.bp_tasks(T, A, out_len)
    _close(got, ref)
    # mixed: the search plus one oblique plane
    ob = recon.plane_task(cg["box"], (0.0, 0.0, 0.0), (0.2, 0.1, 1.0), 48, out_offset=out_len)
    M = np.concatenate([T, np.array([ob], dtype=_hip.TASK_DTYPE)])
    got, ref, kern = _both(cg, "f32", M, A, out_len + 48 * 48)
    assert kern != 10
    _close(got, ref)


def test_cor_groups_plan_after_reset(cg):
    """A COR plan replayed after the stored views were reset and fewer pushed again:
    its graph reads the device view count and rebuilds the detector rows, equal to a
    request over the new views."""
    from minirecast import _hip
    g = cg["g"]
    ctx = _hip.Context(0)
    ctx.set_parallel(g.angles, R, C, PU, PV, WINDOW, "f32")
    ctx.push(np.arange(A), cg["proj"])
    T, out_len = _tasks(cg["box"], PLANES[:1], SHIFTS)
    plan = _hip.Plan(ctx, T, out_len)
    try:
        s_all = plan.cor_scores(A)
        ctx.reset_views()
        half = np.arange(0, A, 2)
        ctx.push(half, np.ascontiguousarray(cg["proj"][:, half, :]))
        s_half = plan.cor_scores(len(half))
        assert ctx.last_bp_kernel() == 10
        imgs = ctx.d2h(np.empty(out_len, np.float32), plan.output_ptr())
    finally:
        plan.close()
    one = ctx.bp_tasks(T, len(half), out_len)
    assert rel_l2(imgs.astype(np.float64), one.astype(np.float64)) < 1e-7
    assert not np.array_equal(s_all, s_half)
    host = np.array([entropy_score(imgs[i * 100 * 100:(i + 1) * 100 * 100]) for i in range(len(T))])
    np.testing.assert_allclose(s_half, host, rtol=1e-12, atol=0)
