"""GPU parity tests: the HIP path against the float64 oracle and the reference's
composition fixtures, on the same inputs.

Tolerance (floating point): relative L2 <= 1e-5 for the plain FP32 path against the
reference-equivalent float64 result; FP16 (per-view scale) and BF16 <= 1e-3
against the float64 oracle.
"""
import numpy as np
import pytest

from conftest import cor_trial_images, entropy_score, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _box_cone(V, R, C):
    s, d, u, v = np.split(np.asarray(V[0], float), 4)
    mag = np.linalg.norm(d - s) / np.linalg.norm(s)
    return np.array([C * np.linalg.norm(u) / mag / 2] * 2 + [R * np.linalg.norm(v) / mag / 2])


# ---------------------------------------------------------------- filtering
@pytest.mark.parametrize("n", [1, 2, 3, 5, 9, 17, 96, 486, 1025, 1944, 2048, 3001, 4096])
def test_filter_rows_vs_oracle(n):
    from minirecast import recon
    rng = np.random.default_rng(n)
    x = rng.normal(size=(7, n)).astype(np.float32)
    for window in O.WINDOWS:
        got = recon.filter_rows(x, 0.3, window)
        want = O.filter_rows(x, 0.3, window)
        assert got.dtype == np.float32 and got.shape == x.shape
        assert rel_l2(got, want) < TOL, (n, window, rel_l2(got, want))


@pytest.mark.parametrize("C", [48, 486, 1944, 4096])
def test_f16_ingest_vs_oracle_and_f32(C):
    """f16 rings store each view with a power-of-two scale chosen before filtering from
    the bound max|q| <= max|x| gain sum|h|: the stored values against the float64 oracle
    (<= 1e-3, one f16 rounding each), the scale an exact power of two with every stored
    value in range, and the same views as the f32 ring to f16 rounding; the f16 ring's
    slices against the f32 ring's."""
    from minirecast import _hip, recon
    from minirecast.geometry import circular_cone_vectors
    R, A = 9, 6
    g = circular_cone_vectors(A, 400.0, 600.0, R, C, 0.15)
    rng = np.random.default_rng(C)
    block = rng.uniform(0.0, 3.0, size=(R, A, C)).astype(np.float32)
    block[:, 2, :] *= 1e-3                               # views of very different scales
    block[:, 4, :] *= 50.0
    r32 = recon.SliceReconstructor(block, g, window="hann", verbose=False)
    r16 = recon.SliceReconstructor(block, g, window="hann", verbose=False, storage="f16")
    q32 = r32.filtered_views()
    q16 = r16.filtered_views()
    K = O.cone_constants(g.vectors)
    q64 = O.fdk_filter_block(block, K, "hann")
    assert rel_l2(q32, q64) < TOL
    for v in range(A):
        assert rel_l2(q16[v], q64[v]) < 1e-3, (v, rel_l2(q16[v], q64[v]))
        assert rel_l2(q16[v], q32[v]) < 1e-3
    assert np.isfinite(q16).all()
    img32, _ = r32.reconstruct_plane((0, 0, 0), (1, 0.5, 0.8), n=40)
    img16, _ = r16.reconstruct_plane((0, 0, 0), (1, 0.5, 0.8), n=40)
    assert rel_l2(img16, img32) < 1e-3


def _round_half_products(q):
    """q rounded to half precision (11 significant bits, nearest even) per view, through
    the power-of-two scale that puts each view's largest |q| in [2^14, 2^15)."""
    out = np.empty_like(q)
    for v in range(q.shape[0]):
        m = float(np.max(np.abs(q[v])))
        s = np.float32(2.0 ** (np.frexp(m)[1] - 15)) if m > 0 else np.float32(1.0)
        out[v] = (q[v] / s).astype(np.float16).astype(np.float32) * s
    return out


def test_f16_provisional_scale_and_fixups():
    """f16 ingest picks a provisional scale from about 96 sampled rows per view (here every
    3rd of 200), records the exact max |q| of each filtered view and filters again the
    views whose largest value would overflow a half or fall below 2^6.  Views: ordinary
    data; zeros on the sampled rows and values ~1e-6 elsewhere (the provisional scale 1
    would leave them subnormal); sampled rows ~1e-3 and one unsampled row of 1e4 (the
    provisional scale would overflow); all zeros.  Every stored product equals the f32
    ring's value rounded to half precision (any admissible power-of-two scale stores the
    same products) wherever it is normal under every admissible scale, and the slices
    equal the f32 ring's to f16 rounding."""
    from minirecast import recon
    from minirecast.geometry import circular_cone_vectors
    R, A, C = 200, 6, 130
    g = circular_cone_vectors(A, 400.0, 600.0, R, C, 0.15)
    rng = np.random.default_rng(96)
    block = rng.uniform(0.0, 3.0, size=(R, A, C)).astype(np.float32)
    sampled = np.arange(R) % 3 == 0
    block[sampled, 1, :] = 0.0
    block[:, 1, :] *= 1e-6
    block[:, 2, :] *= 1e-3
    block[100, 2, :] = 1e4
    block[:, 3, :] = 0.0
    r32 = recon.SliceReconstructor(block, g, window="hann", verbose=False)
    r16 = recon.SliceReconstructor(block, g, window="hann", verbose=False, storage="f16")
    q32, q16 = r32.filtered_views(), r16.filtered_views()
    assert np.isfinite(q16).all()
    want = _round_half_products(q32)
    for v in range(A):
        m = float(np.max(np.abs(q32[v])))
        if m == 0.0:
            assert not q16[v].any(), v
            continue
        e = np.frexp(m)[1]
        normal = np.abs(q32[v]) >= 2.0 ** (e - 20)
        np.testing.assert_array_equal(q16[v][normal], want[v][normal], err_msg=f"view {v}")
        assert np.max(np.abs(q16[v] - q32[v])) <= 2.0 ** (e - 11), v
    img32, _ = r32.reconstruct_plane((0, 0, 0), (1, 0.5, 0.8), n=40)
    img16, _ = r16.reconstruct_plane((0, 0, 0), (1, 0.5, 0.8), n=40)
    assert rel_l2(img16, img32) < 1e-3


def test_fdk_filter_block_vs_oracle():
    from minirecast import recon
    V = O.circular_cone_vectors(5, 400.0, 600.0, 0.3, det_shift_u=1.5)
    K = O.cone_constants(V)
    rng = np.random.default_rng(7)
    block = rng.uniform(0, 2, size=(33, 5, 96)).astype(np.float32)   # odd row count
    got = recon.fdk_filter_block(block, K, "shepp-logan")
    want = O.fdk_filter_block(block, K, "shepp-logan")
    assert got.shape == (5, 33, 96)
    assert rel_l2(got, want) < TOL


# ---------------------------------------------------------------- reference composition
@pytest.fixture(scope="module")
def golden_cone(golden):
    from minirecast import recon
    from minirecast.geometry import ConeVecGeometry
    R, C = (int(x) for x in golden['cone_shape'])
    g = ConeVecGeometry(golden['cone_vectors'], R, C)
    return recon.SliceReconstructor(golden['cone_proj'], g, verbose=False)


def test_reference_planes_cone(golden, golden_cone):
    rec = golden_cone
    for i in range(4):
        a = golden[f'cone_plane{i}_args']
        c = None if np.isnan(a[9]) else a[9:12]
        img, pts = rec.reconstruct_plane(a[0:3], a[3:6], n=int(a[6]), size=a[7] or None,
                                         zoom=a[8], center=c)
        ref = golden[f'cone_plane{i}']
        assert img.dtype == np.float64 and img.shape == ref.shape
        np.testing.assert_array_equal(img == 0, ref == 0)            # same in-box mask
        assert rel_l2(img, ref) < TOL, (i, rel_l2(img, ref))
    img, _ = rec.reconstruct_plane((0, 0, 0), (1, 0.5, 0.8), n=32, cor_shift=1.25)
    assert rel_l2(img, golden['cone_plane_cor']) < TOL


def test_reference_preview_and_cor_cone(golden, golden_cone):
    rec = golden_cone
    vol, origin, h = rec.preview_volume(12)
    assert vol.dtype == np.float32 and vol.shape == golden['cone_preview'].shape
    np.testing.assert_array_equal(origin, golden['cone_preview_origin'])
    assert h == golden['cone_preview_h']
    assert rel_l2(vol, golden['cone_preview']) < TOL
    shifts = golden['cone_cor_shifts']
    best, scores = rec.estimate_cor(shifts, n=32, verbose=False)
    # the device histogram entropy equals numpy's on the very same trial images
    host = [entropy_score(im) for im in cor_trial_images(rec, shifts, 32)]
    np.testing.assert_allclose(scores, host, rtol=1e-12, atol=0)
    # against the reference's images (float64 back-projection): ~820 in-box pixels, so
    # one pixel changing bin moves a score by ~(ln p + 1) / 820, i.e. a few 1e-3
    np.testing.assert_allclose(scores, golden['cone_cor_scores'], rtol=2e-3, atol=2e-3)
    assert best == golden['cone_cor_best']
    rec.set_cor_shift(0.0)


def test_reference_stream_order_cone(golden):
    from minirecast import recon
    from minirecast.geometry import ConeVecGeometry
    from minirecast.stream import Feeder
    R, C = (int(x) for x in golden['cone_shape'])
    g = ConeVecGeometry(golden['cone_vectors'], R, C)
    live = recon.SliceReconstructor(None, g, verbose=False)
    Feeder(live, golden['cone_proj'], order="interleaved", group=7, rate=0, verbose=False).run()
    assert live.n == len(golden['cone_vectors'])
    img, _ = live.reconstruct_plane((0, 0, 0), (0.2, 0.1, 1.0), n=32)
    assert rel_l2(img, golden['cone_stream_plane']) < TOL


def test_reference_planes_parallel(golden):
    from minirecast import recon
    from minirecast.geometry import ParallelGeometry
    R, C = (int(x) for x in golden['par_shape'])
    pu, pv = golden['par_pixel']
    g = ParallelGeometry(golden['par_angles'], R, C, pu, pv)
    rec = recon.SliceReconstructor(golden['par_proj'], g, window="hann", verbose=False)
    img, _ = rec.reconstruct_plane((0, 0, 2.0), (0, 0, 1), n=40)
    assert rel_l2(img, golden['par_plane_axial']) < TOL
    img, _ = rec.reconstruct_plane((1.0, 0, 0), (1, 0.5, 0.8), n=36, size=100.0)
    assert rel_l2(img, golden['par_plane_oblique']) < TOL
    img, _ = rec.reconstruct_plane((0, 0, 0), (0, 0, 1), n=32, cor_shift=-0.75)
    assert rel_l2(img, golden['par_plane_cor']) < TOL
    best, scores = rec.estimate_cor(np.arange(-1.0, 1.01, 0.5), n=32, verbose=False)
    host = [entropy_score(im) for im in cor_trial_images(rec, np.arange(-1.0, 1.01, 0.5), 32)]
    np.testing.assert_allclose(scores, host, rtol=1e-12, atol=0)
    np.testing.assert_allclose(scores, golden['par_cor_scores'], rtol=2e-3, atol=2e-3)
    assert best == golden['par_cor_best']


# ---------------------------------------------------------------- float64 oracle, seeded workload
@pytest.fixture(scope="module")
def small_cone():
    """The 96 x 96, 180-view workload: GPU projections, GPU reconstructor and the
    float64 oracle on the same float32 projections."""
    from minirecast import phantom, recon, workloads
    w = workloads.small_cone()
    proj = phantom.project_cone(w.phantom, w.geom)
    rec = recon.SliceReconstructor(proj, w.geom, window=w.window, verbose=False)
    orc = O.OracleReconstructor("cone", w.geom.rows, w.geom.cols, w.geom.volume_box(),
                                window=w.window, vectors=w.geom.vectors)
    orc.add_projections(np.arange(w.geom.n_angles), proj)
    return w, proj, rec, orc


def test_gpu_phantom_matches_oracle_line_integrals(small_cone):
    w, proj, _, _ = small_cone
    ph = w.phantom
    ks = [0, 45, 179]
    want = O.project_cone(ph.centers, ph.mats, ph.vals, w.geom.vectors[ks], w.geom.rows,
                          w.geom.cols)
    np.testing.assert_allclose(proj[:, ks, :], want, rtol=1e-6, atol=1e-6 * want.max())


def test_small_cone_slices_vs_float64_oracle(small_cone):
    w, _, rec, orc = small_cone
    for p in w.planes:
        img, _ = rec.reconstruct_plane(p['origin'], p['normal'], n=p['n'])
        ref, _ = orc.reconstruct_plane(p['origin'], p['normal'], n=p['n'])
        assert rel_l2(img, ref) < TOL, (p['normal'], rel_l2(img, ref))


def test_points_vs_oracle_and_volume_equals_points(small_cone):
    w, _, rec, orc = small_cone
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1, 1, size=(1000, 3)) * w.geom.volume_box()
    got = rec.reconstruct_points(pts)
    assert rel_l2(got, orc.reconstruct_points(pts)) < TOL
    assert rec.reconstruct_points(np.zeros((0, 3))).shape == (0,)
    vol, origin, h = rec.reconstruct_volume(voxel=0.7)
    assert vol.shape == tuple(np.maximum(np.ceil(2 * w.geom.volume_box() / 0.7).astype(int), 2))
    for frac in ([0.3, 0.6, 0.5], [0.0, 0.0, 0.0], [0.99, 0.2, 0.7]):
        i, j, k = (np.array(vol.shape) * frac).astype(int)
        p = origin + h * np.array([i, j, k])
        assert abs(vol[i, j, k] - rec.reconstruct_points(p[None])[0]) < 1e-5
    ref, _, _ = orc.reconstruct_volume(voxel=0.7)
    assert rel_l2(vol, ref) < TOL


def test_stream_equals_batch(small_cone):
    from minirecast import recon
    from minirecast.stream import Feeder
    w, proj, rec, _ = small_cone
    live = recon.SliceReconstructor(None, w.geom, window=w.window, verbose=False)
    Feeder(live, proj, order="interleaved", group=17, rate=0, verbose=False).run()
    a, _ = rec.reconstruct_plane((0, 0, 3.0), (0.2, 0.1, 1.0), n=64)
    b, _ = live.reconstruct_plane((0, 0, 3.0), (0.2, 0.1, 1.0), n=64)
    assert live.n == w.geom.n_angles
    assert np.max(np.abs(a - b)) < 1e-4 * max(1.0, np.abs(a).max())
    # re-sent views are ignored
    v0 = live.version
    assert live.add_projections([0, 1], proj[:, :2, :]) == live.n
    assert live.version == v0


def test_batched_planes_equal_single_planes(small_cone):
    w, _, rec, _ = small_cone
    imgs = rec.reconstruct_planes(w.planes + [dict(origin=(1, 2, 3), normal=(1, 0.5, 0.8), n=47,
                                                   zoom=2.0, center=(1, 0, 0))])
    for p, img in zip(w.planes, imgs):
        one, _ = rec.reconstruct_plane(p['origin'], p['normal'], n=p['n'])
        np.testing.assert_array_equal(img, one)


def test_device_output_fully_written(small_cone):
    """mr_bp_tasks and mr_bp_tasks_dshift write every element of every task into a
    device buffer (tiles outside the volume box included) and nothing else."""
    from minirecast import _hip, recon
    w, _, rec, _ = small_cone
    box = w.geom.volume_box()
    # a zoomed-out plane: most tiles lie outside the box and are culled
    tasks = np.array([recon.plane_task(box, (0, 0, 0), (1.0, 0.5, 0.8), 70, zoom=0.3,
                                       out_offset=5),
                      recon.plane_task(box, (0, 0, 2.0), (0, 0, 1), 40, out_offset=5 + 4900)],
                     dtype=_hip.TASK_DTYPE)
    out_len = 5 + 4900 + 1600 + 7
    ctx = rec._ctx
    want = ctx.bp_tasks(tasks, rec.n, out_len)
    dwant = ctx.bp_tasks_dshift(tasks, rec.n, out_len)
    sentinel = np.full(out_len, np.nan, np.float32)
    d = ctx.alloc(4 * out_len)
    try:
        for call, ref in ((ctx.bp_tasks_device, want), (ctx.bp_tasks_dshift_device, dwant)):
            ctx.h2d(d, sentinel)
            call(tasks, rec.n, d, out_len)
            got = np.empty(out_len, np.float32)
            ctx.d2h(got, d)
            assert np.isnan(got[:5]).all() and np.isnan(got[-7:]).all()
            np.testing.assert_array_equal(got[5:-7], ref[5:-7])
    finally:
        ctx.free(d)
    assert np.count_nonzero(want[5:5 + 4900]) < 4900 // 2


# ---------------------------------------------------------------- edge cases
def test_edge_cases_and_errors():
    from minirecast import recon
    from minirecast.geometry import circular_cone_vectors
    g = circular_cone_vectors(12, 300.0, 600.0, 9, 13, 1.0)
    rec = recon.SliceReconstructor(None, g, verbose=False)
    img, _ = rec.reconstruct_plane((0, 0, 0), (0, 0, 1), n=8)          # no views yet
    assert not img.any()
    assert not rec.reconstruct_points(np.ones((3, 3))).any()
    vol, _, _ = rec.preview_volume(6)
    assert not vol.any()
    with pytest.raises(ValueError):
        rec.add_projections([0, 1], np.zeros((9, 3, 13), np.float32))
    rng = np.random.default_rng(0)
    proj = rng.normal(size=(9, 12, 13)).astype(np.float32)
    rec.add_projections([3, 3, 5], proj[:, [3, 3, 5], :])               # duplicate within a batch
    assert rec.n == 3
    # plane far outside the box: all zeros, no device fault
    img, _ = rec.reconstruct_plane((0, 0, 1e4), (0, 0, 1), n=16)
    assert not img.any()
    # partial tiles
    orc = O.OracleReconstructor("cone", 9, 13, g.volume_box(), vectors=g.vectors, store_f32=True)
    orc.add_projections([3, 3, 5], proj[:, [3, 3, 5], :])
    got, want = [], []
    for n in (1, 31, 33, 70):
        a, _ = rec.reconstruct_plane((0.1, 0.2, 0.3), (0.3, -1, 0.4), n=n)
        b, _ = orc.reconstruct_plane((0.1, 0.2, 0.3), (0.3, -1, 0.4), n=n)
        np.testing.assert_array_equal(a == 0, b == 0)
        got.append(a.ravel())
        want.append(b.ravel())
    # three views of filtered noise: single pixels cancel, so compare all sizes together
    err = rel_l2(np.concatenate(got), np.concatenate(want))
    print(f"partial tiles rel-L2 {err:.2e}")
    assert err < TOL


def test_module_level_bp_functions_vs_oracle():
    from minirecast import recon
    rng = np.random.default_rng(4)
    V = O.circular_cone_vectors(6, 50.0, 90.0, 0.8, det_shift_u=0.5)
    K = O.cone_constants(V)
    q = rng.normal(size=(6, 11, 17)).astype(np.float32)
    pts = rng.uniform(-5, 5, size=(200, 3))
    D = K['d'] - 0.4 * K['u']
    args = (K['s'], D, K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, 0.37)
    assert rel_l2(recon.bp_cone_numba(q, *args), O.bp_cone(q.astype(np.float64), *args)) < TOL
    th = rng.uniform(0, np.pi, 6)
    pargs = (np.cos(th), np.sin(th), 0.9, 1.1, pts, 0.25, -0.6)
    assert rel_l2(recon.bp_parallel_numba(q, *pargs), O.bp_parallel(q.astype(np.float64), *pargs)) < TOL


# ---------------------------------------------------------------- COR search known answer
def _shepp_logan():
    from minirecast.phantom import EllipsoidPhantom
    return EllipsoidPhantom(scale=50.0)


def test_cor_search_finds_true_shift():
    """selftest.py:78-101: data with the axis 3 px off is fixed by estimate_cor."""
    from minirecast import phantom, recon
    from minirecast.geometry import ParallelGeometry, circular_cone_vectors
    ph = _shepp_logan()
    for kind in ("parallel", "cone"):
        if kind == "parallel":
            g = ParallelGeometry(np.arange(180) * np.pi / 180, 32, 160, 0.9, 0.9)
            proj = np.roll(phantom.project_parallel(ph, g), 3, axis=2)
        else:
            moved = circular_cone_vectors(180, 300.0, 600.0, 32, 160, 1.5, det_shift_u=-3.0)
            proj = phantom.project_cone(ph, moved)
            g = circular_cone_vectors(180, 300.0, 600.0, 32, 160, 1.5)
        rec = recon.SliceReconstructor(proj, g, verbose=False)
        img0, pts = rec.reconstruct_plane((0, 0, 0), (0, 0, 1), n=96, size=110)
        best, scores = rec.estimate_cor(np.arange(-5.0, 5.01, 0.5), n=96, verbose=False)
        img1, _ = rec.reconstruct_plane((0, 0, 0), (0, 0, 1), n=96, size=110)
        truth = ph.evaluate(pts)
        m = np.linalg.norm(pts, axis=-1) < 45.0
        e0 = np.median(np.abs(img0[m] - truth[m]))
        e1 = np.median(np.abs(img1[m] - truth[m]))
        assert abs(best - 3.0) <= 0.5 and e1 < e0, (kind, best, e0, e1)


# ---------------------------------------------------------------- storage forms
def test_storage_precision_vs_float64_oracle(small_cone):
    from minirecast import recon
    w, proj, _, orc = small_cone
    p = w.planes[0]
    ref, _ = orc.reconstruct_plane(p['origin'], (1.0, 0.5, 0.8), n=p['n'])
    errs = {}
    for storage, tol in (("f32", TOL), ("f16", 1e-3), ("bf16", 1e-3)):
        rec = recon.SliceReconstructor(proj, w.geom, window=w.window, verbose=False,
                                       storage=storage)
        img, _ = rec.reconstruct_plane(p['origin'], (1.0, 0.5, 0.8), n=p['n'])
        errs[storage] = rel_l2(img, ref)
        assert errs[storage] < tol, (storage, errs[storage])
        q = rec.filtered_views(0, 4)
        qref = orc.q[:4]
        # one stored value: unit roundoff 2^-24 (f32), 2^-11 (f16), 2^-8 = 3.9e-3 (bf16)
        assert rel_l2(q, qref) < {"f32": TOL, "f16": 1e-3, "bf16": 4e-3}[storage]
    assert errs["f32"] < errs["f16"] < errs["bf16"]


@pytest.mark.parametrize("variant", ["2", "4", "5", "6", "7"])
def test_bp_kernel_variants_vs_oracle(small_cone, variant, monkeypatch):
    """Every tile kernel (MINIRECAST_BP_VARIANT) against the float64 oracle,
    for every storage form, cone beam; and parallel beam on the reference fixture."""
    from minirecast import recon
    monkeypatch.setenv("MINIRECAST_BP_VARIANT", variant)
    w, proj, _, orc = small_cone
    p = dict(origin=(0.5, -0.3, 0.2), normal=(1.0, 0.5, 0.8), n=77)
    ref, _ = orc.reconstruct_plane(p['origin'], p['normal'], n=p['n'])
    for storage, tol in (("f32", TOL), ("f16", 1e-3), ("bf16", 1e-3)):
        rec = recon.SliceReconstructor(proj, w.geom, window=w.window, verbose=False,
                                       storage=storage)
        img, _ = rec.reconstruct_plane(p['origin'], p['normal'], n=p['n'])
        assert rel_l2(img, ref) < tol, (variant, storage, rel_l2(img, ref))


@pytest.mark.parametrize("variant", ["2", "4", "5", "6", "7"])
def test_bp_kernel_variants_parallel(golden, variant, monkeypatch):
    from minirecast import recon
    from minirecast.geometry import ParallelGeometry
    monkeypatch.setenv("MINIRECAST_BP_VARIANT", variant)
    R, C = (int(x) for x in golden['par_shape'])
    pu, pv = golden['par_pixel']
    g = ParallelGeometry(golden['par_angles'], R, C, pu, pv)
    rec = recon.SliceReconstructor(golden['par_proj'], g, window="hann", verbose=False)
    img, _ = rec.reconstruct_plane((1.0, 0, 0), (1, 0.5, 0.8), n=36, size=100.0)
    assert rel_l2(img, golden['par_plane_oblique']) < TOL
    img, _ = rec.reconstruct_plane((0, 0, 0), (0, 0, 1), n=32, cor_shift=-0.75)
    assert rel_l2(img, golden['par_plane_cor']) < TOL


# ---------------------------------------------------------------- preprocessing
@pytest.mark.parametrize("sigma", [0.7, 2.0, 9.0])
def test_gaussian_blur_vs_oracle(sigma):
    from minirecast import preprocess
    rng = np.random.default_rng(1)
    block = rng.normal(1.0, 0.3, (40, 3, 50)).astype(np.float32)
    got = preprocess.gaussian_blur(block, sigma)
    assert got.dtype == np.float32 and got.shape == block.shape
    assert rel_l2(got, O.gaussian_blur(block, sigma)) < 1e-6
    if sigma == 2.0:        # selftest.py:104-111
        assert abs(got.mean() - block.mean()) < 0.01 and got.std() < 0.5 * block.std()


def test_rotate_projections_vs_oracle():
    from minirecast import preprocess
    rng = np.random.default_rng(2)
    block = rng.normal(size=(31, 2, 45)).astype(np.float32)
    for deg in (90.0, -7.5, 0.3):
        got = preprocess.rotate_projections(block, deg)
        assert rel_l2(got, O.rotate_projections(block, deg)) < 1e-6
    img = np.zeros((41, 1, 41), np.float32)
    img[20, 0, 30] = 1.0                                   # selftest.py:112-116
    rot = preprocess.rotate_projections(img, 90.0)
    r, c = np.unravel_index(np.argmax(rot[:, 0, :]), (41, 41))
    assert (r, c) in ((10, 20), (30, 20))
    np.testing.assert_array_equal(preprocess.apply(block), block)


def test_reconstructor_with_preprocessing_vs_oracle(small_cone):
    from minirecast import recon
    w, proj, _, _ = small_cone
    rec = recon.SliceReconstructor(proj, w.geom, window=w.window, verbose=False, gaussian=1.5,
                                   tilt=0.7)
    orc = O.OracleReconstructor("cone", w.geom.rows, w.geom.cols, w.geom.volume_box(),
                                window=w.window, vectors=w.geom.vectors, gaussian=1.5, tilt=0.7)
    orc.add_projections(np.arange(w.geom.n_angles), proj)
    img, _ = rec.reconstruct_plane((0, 0, 0), (1.0, 0.5, 0.8), n=64)
    ref, _ = orc.reconstruct_plane((0, 0, 0), (1.0, 0.5, 0.8), n=64)
    assert rel_l2(img, ref) < TOL


# ---------------------------------------------------------------- raw-count ingest
@pytest.mark.parametrize("style,sample", [("flexray", 2), ("hdf5", 1)])
def test_raw_ingest_vs_host_correction(style, sample):
    """add_raw_projections (uint16 frames corrected on the device) against
    add_projections of the same frames corrected by the oracle restatement of the
    loaders: filtered views and a slice agree to float32 log rounding."""
    from minirecast import phantom, recon
    from minirecast.geometry import circular_cone_vectors
    from minirecast.rawcorr import RawCorrection
    R, C, A = 40, 56, 48
    g = circular_cone_vectors(A, 300.0, 500.0, R, C, 0.4)
    ph = phantom.random_ellipsoids(6, 8, g.volume_box(), (0.02, 0.1))
    p = phantom.project_cone(ph, g)                          # [R, A, C] line integrals
    rng = np.random.default_rng(7)
    r0, c0 = 3, 5
    hb, wb = R + r0 + 2, C + c0 + 1
    H, W = hb * sample + (sample - 1), wb * sample + (sample - 1)
    dark = rng.uniform(95, 105, (hb, wb)).astype(np.float32)
    flat = rng.uniform(9.5e3, 1.05e4, (hb, wb)).astype(np.float32)
    flip = style == "flexray"
    # raw counts whose binned, flipped, cropped transmission follows the projections
    trans = np.ones((A, hb, wb), np.float32)
    trans[:, r0:r0 + R, c0:c0 + C] = np.exp(-np.transpose(p, (1, 0, 2)))
    binned = dark + trans * (flat - dark)
    if flip:
        binned = binned[:, ::-1, :]
    frames = np.zeros((A, H, W), np.uint16)
    up = np.repeat(np.repeat(binned, sample, 1), sample, 2)
    frames[:, :hb * sample, :wb * sample] = np.clip(
        np.round(up + rng.normal(0, 20, up.shape)), 0, 65535).astype(np.uint16)
    corr = RawCorrection(dark, flat, (H, W), style=style, sample=sample,
                         crop=(r0, r0 + R, c0, c0 + C))
    host = O.raw_correct(frames, dark, flat, style, sample, flip, (r0, r0 + R, c0, c0 + C))
    a = recon.SliceReconstructor(None, g, verbose=False)
    b = recon.SliceReconstructor(None, g, verbose=False)
    order = np.random.default_rng(8).permutation(A)
    for i in range(0, A, 10):
        idx = order[i:i + 10]
        assert a.add_raw_projections(idx, frames[idx], corr) == min(i + 10, A)
        b.add_projections(idx, host[:, idx, :])
    qa, qb = a.filtered_views(0, A), b.filtered_views(0, A)
    assert rel_l2(qa, qb) < 2e-6, rel_l2(qa, qb)
    ia, _ = a.reconstruct_plane((0, 0, 0), (1.0, 0.5, 0.8), n=48)
    ib, _ = b.reconstruct_plane((0, 0, 0), (1.0, 0.5, 0.8), n=48)
    assert rel_l2(ia, ib) < 2e-6
    # re-sent views are ignored; shape errors raise
    v0 = a.version
    assert a.add_raw_projections(order[:3], frames[order[:3]], corr) == A and a.version == v0
    with pytest.raises(ValueError):
        a.add_raw_projections([0], frames[:1, :-1], corr)


# ---------------------------------------------------------------- volume threshold
def test_volume_percentiles_histogram_and_otsu_vs_numpy():
    """Device percentiles of the non-zero voxels, the clipped histogram and the Otsu
    threshold equal the reference's numpy computation (model.py:18-29) exactly, on
    volumes with zeros, negative values and ties, from host arrays and device memory."""
    from minirecast import _hip, model, recon
    ctx = _hip.Context(0)
    rng = np.random.default_rng(21)
    cases = [rng.normal(0.2, 0.5, (40, 50, 30)).astype(np.float32),
             np.round(rng.normal(0, 3, 120000)).astype(np.float32),          # ties, zeros
             np.concatenate([rng.uniform(0.1, 0.3, 30000), rng.uniform(0.8, 1.0, 9000),
                             np.zeros(20000), [-1e-30, 1e30]]).astype(np.float32)]
    for vol in cases:
        v = vol[vol != 0]
        q = np.true_divide([0.1, 99.95], np.float32(100))
        want = np.percentile(v, [0.1, 99.95])
        got, m = ctx.nonzero_percentiles(vol, q)
        assert m == v.size
        np.testing.assert_array_equal(got, want)
        lo, hi = want
        want_h, edges = np.histogram(np.clip(v, lo, hi), bins=512, range=(lo, hi))
        np.testing.assert_array_equal(ctx.clip_histogram(vol, lo, hi, edges), want_h)
        assert model.otsu_threshold(vol, ctx=ctx) == O.otsu_threshold(vol)
    # device-resident volume (a reconstruction left in HBM)
    vol = cases[0]
    d = ctx.alloc(vol.nbytes)
    try:
        ctx.h2d(d, vol)
        got, m = ctx.nonzero_percentiles(None, q, dptr=d, n=vol.size)
        np.testing.assert_array_equal(got, np.percentile(vol[vol != 0], [0.1, 99.95]))
    finally:
        ctx.free(d)
    with pytest.raises(ValueError):
        ctx.nonzero_percentiles(np.zeros(10, np.float32), q)
    # a constant volume: numpy's widened range
    c = np.full((5, 6, 7), 0.75, np.float32)
    assert model.otsu_threshold(c, ctx=ctx) == O.otsu_threshold(c)


# ---------------------------------------------------------------- serving
def test_slice_service_replies_and_coalescing(small_cone):
    """The server's requests answered on the GPU path equal direct calls; concurrent
    slice requests are served by fewer back-projection launches than requests."""
    import threading
    from multiprocessing import Pipe
    from minirecast import recon, serve
    w, proj, _, _ = small_cone
    rec = recon.SliceReconstructor(proj, w.geom, window=w.window, verbose=False)
    svc = serve.SliceService(window=w.window, coalesce_ms=20.0)
    h, arr = serve.geometry_message(w.geom)
    assert svc.handle(h, arr)[0] == {"ok": True}
    A = w.geom.n_angles
    for i in range(0, A, 60):
        rep, _ = svc.handle({"cmd": "projections", "indices": list(range(i, i + 60))},
                            proj[:, i:i + 60, :])
        assert rep == {"ok": True, "n": i + 60}
    planes = [dict(origin=(0.1 * k, 0.0, 0.2), normal=(1.0, 0.3 * k, 0.8), n=40 + k)
              for k in range(6)]
    direct = [rec.reconstruct_plane(p['origin'], p['normal'], n=p['n'])[0] for p in planes]
    got = [None] * len(planes)

    def worker(k):
        rep, img = svc.handle({"cmd": "slice", **planes[k]})
        got[k] = img

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(planes))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for k in range(len(planes)):
        assert got[k].dtype == np.float32
        np.testing.assert_allclose(got[k], direct[k], rtol=0, atol=1e-6 * np.abs(direct[k]).max())
    assert svc.slices_served == len(planes) and svc.launches < len(planes)
    # the other requests
    rep, flat = svc.handle({"cmd": "slices", "planes": planes[:2]})
    assert rep["sizes"] == [40, 41] and flat.size == 40 * 40 + 41 * 41
    # the same set again, with views arriving in between: served from running sums
    rec2 = recon.SliceReconstructor(None, w.geom, window=w.window, verbose=False)
    h, arr = serve.geometry_message(w.geom)
    svc.handle(h, arr)
    for i in range(0, A, 45):
        idx = list(range(i, min(A, i + 45)))
        svc.handle({"cmd": "projections", "indices": idx}, proj[:, idx, :])
        rec2.add_projections(idx, proj[:, idx, :])
        rep, flat = svc.handle({"cmd": "slices", "planes": planes[:3]})
        want = np.concatenate([im.ravel() for im in rec2.reconstruct_planes(planes[:3])])
        assert rep["version"] == rec2.version
        np.testing.assert_allclose(flat, want, rtol=0, atol=1e-6 * np.abs(want).max())
    assert svc.live_hits >= 2
    assert svc.handle({"cmd": "info"})[0]["n"] == A
    rep, vol = svc.handle({"cmd": "preview", "maxdim": 16})
    ref_vol, ref_origin, ref_h = rec.preview_volume(16)
    np.testing.assert_allclose(vol, ref_vol, atol=1e-6 * np.abs(ref_vol).max())
    assert rep["spacing"] == ref_h
    assert "best" in svc.handle({"cmd": "auto_cor"})[0]
    assert svc.handle({"cmd": "set_cor", "px": 0.5})[0] == {"ok": True}
    assert "error" in svc.handle({"cmd": "nope"})[0]
    # the request loop over a connection
    a, b = Pipe()
    t = threading.Thread(target=svc.serve, args=(b,), daemon=True)
    t.start()
    serve.send(a, {"cmd": "slice", **planes[0]})
    rep, img = serve.recv(a)
    assert img.shape == (40, 40) and rep["version"] == svc.recon.version
    serve.send(a, {"cmd": "projections", "indices": [0]}, proj[:, :2, :])   # count mismatch
    assert "error" in serve.recv(a)[0]
    a.close()
    t.join(timeout=10)


def test_served_replies_match_reference_fixtures_and_oracle(golden):
    """The server's replies (net.py:140-154) from the GPU path against the reference's
    own results: served slices (the three fixture planes without a size, one with zoom
    and centre), the slice after set_cor and the preview against the fixtures the
    reference's numpy backend wrote; auto_cor (estimate_cor's defaults: 25 shifts,
    n = 192) against the float64 oracle's search on the same projections."""
    from minirecast import serve
    from minirecast.geometry import ConeVecGeometry
    R, C = (int(x) for x in golden['cone_shape'])
    V = golden['cone_vectors']
    g = ConeVecGeometry(V, R, C)
    proj = golden['cone_proj']
    svc = serve.SliceService(window="shepp-logan", coalesce_ms=0.0)
    h, arr = serve.geometry_message(g)
    assert svc.handle(h, arr)[0] == {"ok": True}
    order = golden['cone_stream_order']
    for i in range(0, len(order), 7):
        idx = [int(k) for k in order[i:i + 7]]
        svc.handle({"cmd": "projections", "indices": idx}, proj[:, idx, :])
    for i in (1, 2, 3):                        # fixture planes whose size is the default
        a = golden[f'cone_plane{i}_args']
        assert a[7] == 0
        req = {"cmd": "slice", "origin": list(a[0:3]), "normal": list(a[3:6]), "n": int(a[6]),
               "zoom": float(a[8])}
        if not np.isnan(a[9]):
            req["center"] = list(a[9:12])
        rep, img = svc.handle(req)
        ref = golden[f'cone_plane{i}']
        assert img.dtype == np.float32 and img.shape == ref.shape
        np.testing.assert_array_equal(img == 0, ref == 0)
        assert rel_l2(img, ref) < TOL, (i, rel_l2(img, ref))
    assert svc.handle({"cmd": "set_cor", "px": 1.25})[0] == {"ok": True}
    rep, img = svc.handle({"cmd": "slice", "origin": [0, 0, 0], "normal": [1, 0.5, 0.8], "n": 32})
    assert rel_l2(img, golden['cone_plane_cor']) < TOL
    assert svc.handle({"cmd": "set_cor", "px": 0.0})[0] == {"ok": True}
    rep, vol = svc.handle({"cmd": "preview", "maxdim": 12})
    assert vol.shape == golden['cone_preview'].shape and rel_l2(vol, golden['cone_preview']) < TOL
    np.testing.assert_array_equal(rep["origin"], golden['cone_preview_origin'])
    assert rep["spacing"] == golden['cone_preview_h']
    rep, _ = svc.handle({"cmd": "auto_cor"})
    s_, d_, u_, v_ = np.split(V[0], 4)
    mag = np.linalg.norm(d_ - s_) / np.linalg.norm(s_)
    box = np.array([C * np.linalg.norm(u_) / mag / 2] * 2 + [R * np.linalg.norm(v_) / mag / 2])
    orc = O.OracleReconstructor("cone", R, C, box, vectors=V, store_f32=True)
    orc.add_projections(np.arange(len(V)), proj)
    shifts = np.arange(-3.0, 3.01, 0.25)
    want = np.asarray(orc.cor_scores(shifts, n=192))
    assert rep["best"] == float(shifts[int(np.argmax(want))])
    assert svc.recon.cor_shift == rep["best"]


def test_warmup_runs_the_device_path(golden):
    """warmup (recon.py:477-487) runs the 4-point back-projection on the device: a no-op
    before any view arrives, then the same values as reconstruct_points at the origin,
    which equal the float64 oracle's; there is no CPU backend to fall back to, so the
    backend stays "hip" (a failure would raise)."""
    from minirecast import recon
    from minirecast.geometry import ConeVecGeometry
    R, C = (int(x) for x in golden['cone_shape'])
    V = golden['cone_vectors']
    rec = recon.SliceReconstructor(None, ConeVecGeometry(V, R, C), verbose=False)
    assert rec.warmup() is None and rec.n == 0
    rec.add_projections(np.arange(len(V)), golden['cone_proj'])
    assert rec.warmup() is None and rec.backend == "hip"
    got = rec.reconstruct_points(np.zeros((4, 3)))
    orc = O.OracleReconstructor("cone", R, C, _box_cone(V, R, C), vectors=V, store_f32=True)
    orc.add_projections(np.arange(len(V)), golden['cone_proj'])
    want = orc.reconstruct_points(np.zeros((4, 3)))
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7 * np.abs(want).max())
