"""GPU tests of the batched forms: persistent launches over a device work queue
(plans), graph-replayed streaming updates, and the back-projection variants."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload():
    from minirecast import phantom, recon, workloads
    w = workloads.small_cone()
    proj = phantom.project_cone(w.phantom, w.geom)
    box = w.geom.volume_box()
    planes = [dict(origin=(0.0, 0.0, 0.0), normal=(1.0, 0.5, 0.8), n=96),
              dict(origin=(0.5, 0.0, -1.0), normal=(-0.3, 1.0, 0.4), n=64, zoom=1.7,
                   center=(1.0, 0.0, 0.0)),
              dict(origin=(0.0, 0.0, 2.0), normal=(0.0, 0.0, 1.0), n=47)]
    tasks, off = [], 0
    for p in planes:
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], p.get('size'),
                                      p.get('zoom', 1.0), p.get('center'), 0.0, off))
        off += p['n'] ** 2
    return w, proj, planes, np.array(tasks), off


def _ctx(w, storage="f32", variant="2", monkeypatch=None):
    from minirecast import _hip
    if monkeypatch is not None:
        monkeypatch.setenv("MINIRECAST_BP_VARIANT", variant)
    ctx = _hip.Context(0)
    ctx.set_cone(w.geom.vectors, w.geom.rows, w.geom.cols, w.window, storage)
    return ctx


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_plan_run_matches_tasks(workload, storage, monkeypatch):
    from minirecast import _hip
    w, proj, planes, tasks, out_len = workload
    ctx = _ctx(w, storage, "2", monkeypatch)
    A = w.geom.n_angles
    ctx.push(np.arange(A), proj)
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    ref = ctx.bp_tasks(T, A, out_len)
    plan = _hip.Plan(ctx, T, out_len)
    for _ in range(3):                                  # repeated runs reuse the queue
        got = plan.run()
        np.testing.assert_array_equal(got == 0, ref == 0)
        assert rel_l2(got, ref) < 1e-6
    half = plan.run(A // 2)
    assert rel_l2(half, ctx.bp_tasks(T, A // 2, out_len)) < 1e-6


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_plan_stream_graph_replays(workload, storage, monkeypatch):
    """Streaming updates replay one captured graph per batch size; every replay (with
    eager work in between) matches the eager ingest + back-projection."""
    from minirecast import _hip
    from minirecast.stream import arrival_order
    w, proj, planes, tasks, out_len = workload
    A = w.geom.n_angles
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    live = _ctx(w, storage, "2", monkeypatch)
    eager = _ctx(w, storage, "2", monkeypatch)
    plan = _hip.Plan(live, T, out_len)
    order = arrival_order(A, "interleaved")
    batches = [order[i:i + 10] for i in range(0, A, 10)]
    for bi, idx in enumerate(batches):
        blk = np.ascontiguousarray(proj[:, idx, :])
        got = plan.stream(idx, blk)
        eager.push(idx, blk)
        want = eager.bp_tasks(T, eager.num_views(), out_len)
        assert live.num_views() == eager.num_views()
        assert rel_l2(got, want) < 1e-6, (bi, rel_l2(got, want))
    # the final state equals the float64 oracle
    orc = O.OracleReconstructor("cone", w.geom.rows, w.geom.cols, w.geom.volume_box(),
                                window=w.window, vectors=w.geom.vectors)
    for idx in batches:
        orc.add_projections(idx, proj[:, idx, :])
    off = 0
    tol = 1e-5 if storage == "f32" else 1e-3
    for p in planes:
        img, _ = orc.reconstruct_plane(p['origin'], p['normal'], n=p['n'], zoom=p.get('zoom', 1.0),
                                       center=p.get('center'))
        assert rel_l2(got[off:off + p['n'] ** 2], img.ravel()) < tol
        off += p['n'] ** 2


@pytest.mark.parametrize("variant,ppt,layout,depth", [
    ("2", "4", "0", "2"), ("4", "4", "0", "2"), ("4", "2", "0", "2"), ("4", "2", "1", "3"),
    ("4", "1", "1", "2"), ("4", "1", "0", "3"), ("4", "4", "1", "3"), ("5", "4", "0", "2"),
    ("5", "2", "0", "2"), ("6", "2", "0", "2"), ("6", "1", "1", "3"), ("6", "4", "0", "2"),
    ("7", "2", "0", "2"), ("7", "2", "0", "3"), ("7", "1", "1", "3"), ("7", "4", "0", "2")])
def test_bp_variants_full_check(workload, variant, ppt, layout, depth, monkeypatch):
    """Every tile-kernel configuration (kernel, points per thread, lane layout, views in
    flight) against kernel 2 on the f32 ring on all outputs, for view counts that leave
    every pipeline tail length (n, n - 1, 2, 1)."""
    from minirecast import _hip
    monkeypatch.setenv("MINIRECAST_BP_PPT", ppt)
    monkeypatch.setenv("MINIRECAST_BP_LAYOUT", layout)
    monkeypatch.setenv("MINIRECAST_BP_DEPTH", depth)
    w, proj, planes, tasks, out_len = workload
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    A = w.geom.n_angles
    ref_ctx = _ctx(w, "f32", "2", monkeypatch)
    ref_ctx.push(np.arange(A), proj)
    refs = {nv: ref_ctx.bp_tasks(T, nv, out_len) for nv in (A, A - 1, 2, 1)}
    for storage, tol in (("f32", 1e-6), ("f16", 1e-3), ("bf16", 5e-3)):
        ctx = _ctx(w, storage, variant, monkeypatch)
        ctx.push(np.arange(A), proj)
        for nv, ref in refs.items():
            got = ctx.bp_tasks(T, nv, out_len)
            np.testing.assert_array_equal(got == 0, ref == 0)
            assert rel_l2(got, ref) < tol, (variant, ppt, layout, depth, storage, nv,
                                            rel_l2(got, ref))
    # parallel beam through the same variant
    from minirecast import phantom, recon
    from minirecast.geometry import ParallelGeometry
    g = ParallelGeometry(np.arange(90) * np.pi / 90, 24, 80, 1.0, 1.0)
    ph = phantom.random_ellipsoids(9, 8, g.volume_box(), (0.1, 1.0))
    pp = phantom.project_parallel(ph, g)
    rec = recon.SliceReconstructor(pp, g, window="hann", verbose=False)
    img, _ = rec.reconstruct_plane((1.0, 2.0, 3.0), (1.0, 0.5, 0.8), n=70, cor_shift=0.6)
    orc = O.OracleReconstructor("parallel", 24, 80, g.volume_box(), window="hann",
                                angles=g.angles, pixel_u=1.0, pixel_v=1.0)
    orc.add_projections(np.arange(90), pp)
    ref2, _ = orc.reconstruct_plane((1.0, 2.0, 3.0), (1.0, 0.5, 0.8), n=70, cor_shift=0.6)
    assert rel_l2(img, ref2) < 1e-5


def test_autotune_picks_a_configuration_and_keeps_parity(workload):
    """The autotuner times the configurations, keeps one, and the tuned context still
    matches kernel 2 on the ring; configurations are validated."""
    from minirecast import _hip, recon
    w, proj, planes, tasks, out_len = workload
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    A = w.geom.n_angles
    ref_ctx = _hip.Context(0)
    ref_ctx.set_cone(w.geom.vectors, w.geom.rows, w.geom.cols, w.window)
    ref_ctx.set_bp_config(kernel=2, ppt=4, layout=0, depth=2)
    ref_ctx.push(np.arange(A), proj)
    ref = ref_ctx.bp_tasks(T, A, out_len)
    rec = recon.SliceReconstructor(proj, w.geom, window=w.window, verbose=False)
    cfg, ms = rec.autotune(planes, max_tiles=64, reps=1)
    assert ms > 0 and cfg == rec._ctx.bp_config()
    assert cfg["kernel"] in (2, 4, 5, 6, 7) and cfg["ppt"] in (1, 2, 4)
    got = rec._ctx.bp_tasks(T, A, out_len)
    assert rel_l2(got, ref) < 1e-6, (cfg, rel_l2(got, ref))
    with pytest.raises(ValueError):
        rec._ctx.set_bp_config(kernel=8)
    with pytest.raises(ValueError):
        rec._ctx.set_bp_config(kernel=4, ppt=3)


@pytest.mark.parametrize("variant,vchunk", [("6", "1"), ("6", "7"), ("4", "64"), ("4", "5"), ("7", "3")])
def test_view_chunked_launches_are_exact(workload, variant, vchunk, monkeypatch):
    """Cone-beam updates split into launches of a few views each (the FP32 sums carried
    through the output between launches) equal the single launch bit for bit.  The
    split applies to grids of at least 8 workgroups per CU: the slices here are small,
    so MINIRECAST_BP_VCHUNK_ANY lifts that condition."""
    from minirecast import _hip
    w, proj, planes, tasks, out_len = workload
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    A = w.geom.n_angles
    monkeypatch.setenv("MINIRECAST_BP_PPT", "4")
    for storage in ("f32", "f16"):
        monkeypatch.setenv("MINIRECAST_BP_VCHUNK", "0")
        one = _ctx(w, storage, variant, monkeypatch)
        monkeypatch.setenv("MINIRECAST_BP_VCHUNK", vchunk)
        monkeypatch.setenv("MINIRECAST_BP_VCHUNK_ANY", "1")
        split = _ctx(w, storage, variant, monkeypatch)
        monkeypatch.delenv("MINIRECAST_BP_VCHUNK_ANY")
        for c in (one, split):
            c.push(np.arange(A), proj)
        for nv in (A, A - 1, 2, 1):
            np.testing.assert_array_equal(split.bp_tasks(T, nv, out_len), one.bp_tasks(T, nv, out_len))


@pytest.mark.parametrize("variant", ["6", "7"])
def test_pair_ring_follows_appends_and_resets(workload, variant, monkeypatch):
    """Kernels 6 and 7 read a row-pair or band copy of the ring built on first use: it must
    take in views appended after that use, and rebuild after the views are reset and
    replaced."""
    from minirecast import _hip
    w, proj, planes, tasks, out_len = workload
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    A = w.geom.n_angles
    ref_ctx = _ctx(w, "f16", "2", monkeypatch)
    ctx = _ctx(w, "f16", variant, monkeypatch)
    h = A // 2
    for c in (ref_ctx, ctx):
        c.push(np.arange(h), proj[:, :h])
    assert rel_l2(ctx.bp_tasks(T, h, out_len), ref_ctx.bp_tasks(T, h, out_len)) < 1e-6
    for c in (ref_ctx, ctx):
        c.push(np.arange(h, A), proj[:, h:])
    assert rel_l2(ctx.bp_tasks(T, A, out_len), ref_ctx.bp_tasks(T, A, out_len)) < 1e-6
    other = proj[:, ::-1].copy()
    for c in (ref_ctx, ctx):
        c.reset_views()
        c.push(np.arange(A), other)
    assert rel_l2(ctx.bp_tasks(T, A, out_len), ref_ctx.bp_tasks(T, A, out_len)) < 1e-6


@pytest.mark.parametrize("rows", [1, 2, 4, 37, 38, 39])
def test_band_ring_rows_and_borders(rows, monkeypatch):
    """Kernel 7 (band ring) against kernel 2 on the same 16-bit ring, for detector heights
    of every residue mod 3 (down to single-row detectors), cone and parallel beam: oblique
    slices, and a volume grid over 1.6x the box across and 4x along z whose outer points
    project off the detector (the zero border and clamps)."""
    from minirecast import _hip, phantom, recon
    from minirecast.geometry import ParallelGeometry, circular_cone_vectors
    C = 50
    cone = circular_cone_vectors(90, 400.0, 600.0, rows, C, 0.3)
    par = ParallelGeometry(np.arange(90) * np.pi / 90, rows, C, 1.0, 1.0)
    for g in (cone, par):
        box = g.volume_box()
        ph = phantom.random_ellipsoids(3, 6, box, (0.1, 1.0))
        if g is cone:
            proj = phantom.project_cone(ph, g)
        else:
            proj = phantom.project_parallel(ph, g)
        tasks, off = [], 0
        for normal, n in (((1.0, 0.5, 0.8), 61), ((0.2, -0.1, 1.0), 45)):
            tasks.append(recon.plane_task(box, (0.0, 0.0, 0.3), normal, n, out_offset=off))
            off += n * n
        # 1.6x the box across, 4x along z: the outer z-planes project off the detector
        dims, origin, h = recon.volume_grid(np.asarray(box) * (1.6, 1.6, 4.0), 24)
        vt = recon.volume_task(origin, h, dims, 0, int(dims[2]))
        vt['out_offset'] = off
        tasks.append(vt)
        out_len = off + int(np.prod(dims))
        T = np.array(tasks, dtype=_hip.TASK_DTYPE)
        for storage in ("f16", "bf16"):
            outs = []
            for variant in ("2", "7"):
                monkeypatch.setenv("MINIRECAST_BP_VARIANT", variant)
                monkeypatch.setenv("MINIRECAST_BP_PPT", "2")
                ctx = _hip.Context(0)
                if g is cone:
                    ctx.set_cone(g.vectors, rows, C, "ram-lak", storage)
                else:
                    ctx.set_parallel(g.angles, rows, C, 1.0, 1.0, "ram-lak", storage)
                assert ctx.bp_config()["kernel"] == int(variant)
                ctx.push(np.arange(g.n_angles), proj)
                outs.append(ctx.bp_tasks(T, g.n_angles, out_len))
            ref, got = outs
            np.testing.assert_array_equal(got == 0, ref == 0)
            assert rel_l2(got, ref) < 1e-6, (rows, type(g).__name__, storage, rel_l2(got, ref))
            assert np.count_nonzero(ref[off:]) < ref[off:].size   # border points present


def test_autotune_views_per_pass(workload):
    """With every tile timed and a grid that fills the device, the autotuner also picks
    the views per launch; the tuned context still matches kernel 2 on the ring, and the
    setting is validated and reported."""
    from minirecast import _hip, recon
    w, proj, planes, tasks, out_len = workload
    A = w.geom.n_angles
    box = w.geom.volume_box()
    big = [dict(origin=(0.0, 0.0, 0.0), normal=nrm, n=1024)
           for nrm in ((1.0, 0.5, 0.8), (-0.3, 1.0, 0.4), (0.2, -0.1, 1.0))]
    T, off = [], 0
    for p in big:
        T.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], out_offset=off))
        off += p['n'] ** 2
    T = np.array(T, dtype=_hip.TASK_DTYPE)
    ref_ctx = _hip.Context(0)
    ref_ctx.set_cone(w.geom.vectors, w.geom.rows, w.geom.cols, w.window, "f16")
    ref_ctx.set_bp_config(kernel=2, ppt=4, layout=0, depth=2)
    ref_ctx.push(np.arange(A), proj)
    ref = ref_ctx.bp_tasks(T, A, off)
    ctx = _hip.Context(0)
    ctx.set_cone(w.geom.vectors, w.geom.rows, w.geom.cols, w.window, "f16")
    ctx.push(np.arange(A), proj)
    assert ctx.views_per_pass() == 64                    # the 16-bit default
    cfg, ms = ctx.autotune(T, A, off, max_tiles=10 ** 6, reps=1)
    assert ms > 0 and ctx.views_per_pass() in (0, 16, 32, 64, 128)
    got = ctx.bp_tasks(T, A, off)
    np.testing.assert_array_equal(got == 0, ref == 0)
    assert rel_l2(got, ref) < 1e-6, (cfg, ctx.views_per_pass(), rel_l2(got, ref))
    ctx.set_views_per_pass(0)
    assert ctx.views_per_pass() == 0
    with pytest.raises(ValueError):
        ctx.set_views_per_pass(-1)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_plan_cor_scores_graph(workload, storage, monkeypatch):
    """The rotation-centre search as one graph (persistent back-projection of every trial
    slice + device scoring): every replay, with eager work in between, matches the
    one-launch search and picks the same shift; a plan whose tasks are not single images
    of contiguous rows (here a stack of z-planes) raises."""
    from minirecast import _hip, recon
    w, proj, planes, tasks, out_len = workload
    A = w.geom.n_angles
    box = w.geom.volume_box()
    shifts = np.arange(-2.0, 2.01, 0.5)
    n = 48
    T = np.array([recon.plane_task(box, (0.0, 0.0, 0.3), (0.0, 0.0, 1.0), n, cor_shift=s,
                                   out_offset=i * n * n) for i, s in enumerate(shifts)],
                 dtype=_hip.TASK_DTYPE)
    ctx = _ctx(w, storage, "2", monkeypatch)
    ctx.push(np.arange(A), proj)
    plan = _hip.Plan(ctx, T, len(shifts) * n * n)
    first = None
    for nv in (A, A // 2, A):
        want = ctx.cor_scores(T, nv)
        got = plan.cor_scores(nv)
        np.testing.assert_allclose(got, want, rtol=2e-3, atol=2e-3)
        assert np.argmax(got) == np.argmax(want)
        if nv == A:
            if first is None:
                first = got
            else:
                np.testing.assert_array_equal(got, first)   # replay after eager work
        imgs = plan.run(nv)                                 # the images stay readable
        assert rel_l2(imgs, ctx.bp_tasks(T, nv, len(shifts) * n * n)) < 1e-6
    vt = recon.volume_task(-box / 2, 0.1, (8, 8, 4), 0, 4)
    bad = _hip.Plan(ctx, np.array([vt], dtype=_hip.TASK_DTYPE), 8 * 8 * 4)
    with pytest.raises(ValueError):
        bad.cor_scores()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_cor_scores_of_uneven_images(workload, storage):
    """The searches of several slice sizes scored together, as one request and as one
    plan (one persistent launch over every trial image + scoring, one graph), equal the
    per-size calls exactly: each image's sums and histogram do not depend on the other
    images.  The plan holds the images in any order and with gaps between them."""
    from minirecast import _hip, recon
    w, proj, planes, tasks, out_len = workload
    A = w.geom.n_angles
    box = w.geom.volume_box()
    shifts = np.arange(-1.0, 1.01, 0.5)
    sizes = ((40, 0.3), (24, -0.2), (33, 0.0))
    groups = [np.array([recon.plane_task(box, (0.0, 0.0, z), (0.0, 0.0, 1.0), n, cor_shift=s)
                        for s in shifts], dtype=_hip.TASK_DTYPE) for n, z in sizes]
    ctx = _hip.Context(0)
    ctx.set_cone(w.geom.vectors, w.geom.rows, w.geom.cols, "ram-lak", storage)
    ctx.push(np.arange(A), proj)
    sep = [ctx.cor_scores(G, A) for G in groups]
    allT = np.concatenate(groups)
    np.testing.assert_array_equal(ctx.cor_scores(allT, A), np.concatenate(sep))
    # plans: per size (contiguous images) and one plan of all sizes, reversed, 7-pixel gaps
    per = []
    for G in groups:
        P = G.copy()
        npx = int(P[0]['np']) * int(P[0]['ns'])
        P['out_offset'] = np.arange(len(P)) * npx
        pl = _hip.Plan(ctx, P, len(P) * npx)
        per.append(pl.cor_scores(A))
        pl.close()
    P = allT[::-1].copy()
    off = 0
    for i in range(len(P)):
        P[i]['out_offset'] = off
        off += int(P[i]['np']) * int(P[i]['ns']) + 7
    pl = _hip.Plan(ctx, P, off)
    for _ in range(2):                                   # capture, then a replay
        got = pl.cor_scores(A)[::-1]
        np.testing.assert_array_equal(got, np.concatenate(per))
    pl.close()
    np.testing.assert_allclose(np.concatenate(per), np.concatenate(sep), rtol=2e-3, atol=2e-3)


@pytest.mark.parametrize("beam", ["cone", "parallel"])
def test_plan_stream_running_sums_are_exact(workload, beam):
    """A streaming plan keeps the unscaled FP32 sums and adds only each batch's new views:
    every output equals a full pass over all stored views bit for bit, also after views
    pushed by another ingest path (the sums catch up), after the views are reset and
    refilled (the sums restart), and for both storage forms.  A plan refuses to run once
    the context's geometry has been set again."""
    from minirecast import _hip, phantom, recon
    from minirecast.geometry import ParallelGeometry
    from minirecast.stream import arrival_order
    if beam == "cone":
        w, proj, planes, tasks, out_len = workload
        g = w.geom
        T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    else:
        g = ParallelGeometry(np.arange(60) * np.pi / 60, 24, 80, 1.0, 1.0)
        box = g.volume_box()
        ph = phantom.random_ellipsoids(5, 8, box, (0.1, 1.0))
        proj = phantom.project_parallel(ph, g)
        T, off = [], 0
        for nrm, n in (((1.0, 0.5, 0.8), 70), ((0.0, 0.0, 1.0), 40)):
            T.append(recon.plane_task(box, (1.0, 2.0, 3.0), nrm, n, cor_shift=0.6, out_offset=off))
            off += n * n
        T, out_len = np.array(T, dtype=_hip.TASK_DTYPE), off
    A = g.n_angles
    order = arrival_order(A, "interleaved")
    batches = [order[i:i + 7] for i in range(0, A, 7)]
    for storage in ("f32", "f16"):
        ctx = _hip.Context(0)
        if beam == "cone":
            ctx.set_cone(g.vectors, g.rows, g.cols, "ram-lak", storage)
        else:
            ctx.set_parallel(g.angles, g.rows, g.cols, 1.0, 1.0, "ram-lak", storage)
        plan = _hip.Plan(ctx, T, out_len)
        h = len(batches) // 2
        for bi, idx in enumerate(batches[:h]):
            got = plan.stream(idx, np.ascontiguousarray(proj[:, idx, :]))
            np.testing.assert_array_equal(got, plan.run(), err_msg=f"{storage} batch {bi}")
        # views from another ingest path, then streaming again
        idx = batches[h]
        ctx.push(idx, np.ascontiguousarray(proj[:, idx, :]))
        for idx in batches[h + 1:]:
            got = plan.stream(idx, np.ascontiguousarray(proj[:, idx, :]))
            np.testing.assert_array_equal(got, plan.run())
        assert ctx.num_views() == A
        # reset and refill with other data: the sums restart from zero
        ctx.reset_views()
        other = proj[:, ::-1].copy()
        for idx in batches[:3]:
            got = plan.stream(idx, np.ascontiguousarray(other[:, idx, :]))
            np.testing.assert_array_equal(got, plan.run())
        if beam == "cone":
            ctx.set_cone(g.vectors, g.rows, g.cols, "ram-lak", storage)
        else:
            ctx.set_parallel(g.angles, g.rows, g.cols, 1.0, 1.0, "ram-lak", storage)
        with pytest.raises(RuntimeError):
            plan.stream(batches[0], np.ascontiguousarray(proj[:, batches[0], :]))
        with pytest.raises(RuntimeError):
            plan.run()
        plan.close()


def test_live_planes_follow_add_projections(workload):
    """SliceReconstructor.live_planes: after every batch ingested by add_projections,
    refresh() equals a full pass of the same plan bit for bit and reconstruct_planes
    (another kernel) to FP32 rounding; a refresh with no new views returns the same
    images, and the views of a reset are summed from zero again."""
    from minirecast import recon
    from minirecast.stream import arrival_order
    w, proj, planes, tasks, out_len = workload
    A = w.geom.n_angles
    rec = recon.SliceReconstructor(None, w.geom, window=w.window, verbose=False)
    live = rec.live_planes(planes)
    order = arrival_order(A, "interleaved")
    for i in range(0, A, 9):
        idx = order[i:i + 9]
        rec.add_projections(idx, np.ascontiguousarray(proj[:, idx, :]))
        imgs = live.refresh()
        full = live._plan.run()
        np.testing.assert_array_equal(np.concatenate([im.ravel() for im in imgs]),
                                      full.astype(np.float64))
        ref = rec.reconstruct_planes(planes)
        for a, b in zip(imgs, ref):
            assert rel_l2(a, b) < 1e-6
    again = live.refresh()
    for a, b in zip(again, imgs):
        np.testing.assert_array_equal(a, b)
    live.close()


def test_repeated_requests_reuse_tiles_exactly(workload):
    """A context reuses the tile list of a request that repeats the previous one; any
    change in the tasks (one plane moved, a culled tile appearing) or in the operation
    rebuilds it: every result equals a fresh context's bit for bit."""
    from minirecast import _hip, recon
    w, proj, planes, tasks, out_len = workload
    A = w.geom.n_angles
    box = w.geom.volume_box()
    T1 = np.array(tasks, dtype=_hip.TASK_DTYPE)
    T2 = T1.copy()
    T2[1] = recon.plane_task(box, (0.5, 0.0, 9.0), (-0.3, 1.0, 0.4), 64, None, 1.7,
                             (1.0, 0.0, 0.0), 0.0, int(T1[1]['out_offset']))
    ctx = _ctx(w, "f16")
    ctx.push(np.arange(A), proj)
    seq = [ctx.bp_tasks(T, A, out_len) for T in (T1, T1, T2, T1, T2, T2)]
    fresh = {}
    for name, T in (("1", T1), ("2", T2)):
        c = _ctx(w, "f16")
        c.push(np.arange(A), proj)
        fresh[name] = c.bp_tasks(T, A, out_len)
    for got, name in zip(seq, "112122"):
        np.testing.assert_array_equal(got, fresh[name])
    assert not np.array_equal(fresh["1"], fresh["2"])
    # more distinct requests than the cache holds, cycled twice
    sets = []
    for k in range(6):
        Tk = T1.copy()
        Tk['cor_shift'] = 0.25 * k
        sets.append(Tk)
    ref = []
    for Tk in sets:
        c = _ctx(w, "f16")
        c.push(np.arange(A), proj)
        ref.append(c.bp_tasks(Tk, A, out_len))
    for _ in range(2):
        for Tk, r in zip(sets, ref):
            np.testing.assert_array_equal(ctx.bp_tasks(Tk, A, out_len), r)
    for Tk, r in zip(sets[::-1], ref[::-1]):
        np.testing.assert_array_equal(ctx.bp_tasks(Tk, A, out_len), r)


@pytest.mark.parametrize("beam", ["cone", "parallel"])
def test_plan_band_copy_matches_ring(workload, beam, monkeypatch):
    """Persistent plan launches read a copy of the ring: the band copy in a kernel-7
    (default 16-bit) context, the row-pair copy for f32 parallel beam.  Plan runs,
    streamed batches (whose graph converts each batch it stores) and updates after other
    ingest paths equal the same plans on the ring itself (kernel 2) to FP32 rounding, for
    f16, bf16 and f32."""
    from minirecast import _hip, phantom, recon
    from minirecast.geometry import ParallelGeometry
    from minirecast.stream import arrival_order
    if beam == "cone":
        w, proj, planes, tasks, out_len = workload
        g = w.geom
        T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    else:
        g = ParallelGeometry(np.arange(60) * np.pi / 60, 24, 80, 1.0, 1.0)
        box = g.volume_box()
        ph = phantom.random_ellipsoids(5, 8, box, (0.1, 1.0))
        proj = phantom.project_parallel(ph, g)
        T, off = [], 0
        for nrm, n in (((1.0, 0.5, 0.8), 70), ((0.0, 0.0, 1.0), 40)):
            T.append(recon.plane_task(box, (1.0, 2.0, 3.0), nrm, n, cor_shift=0.6, out_offset=off))
            off += n * n
        T, out_len = np.array(T, dtype=_hip.TASK_DTYPE), off
    A = g.n_angles
    order = arrival_order(A, "interleaved")
    batches = [order[i:i + 8] for i in range(0, A, 8)]
    for storage in ("f16", "bf16", "f32"):
        outs = {}
        for variant in ("7", "2"):
            monkeypatch.setenv("MINIRECAST_BP_VARIANT", variant)
            ctx = _hip.Context(0)
            if beam == "cone":
                ctx.set_cone(g.vectors, g.rows, g.cols, "ram-lak", storage)
            else:
                ctx.set_parallel(g.angles, g.rows, g.cols, 1.0, 1.0, "ram-lak", storage)
            plan = _hip.Plan(ctx, T, out_len)
            res = []
            h = len(batches) // 2
            for idx in batches[:h]:
                res.append(plan.stream(idx, np.ascontiguousarray(proj[:, idx, :])))
            for idx in batches[h:]:
                ctx.push(idx, np.ascontiguousarray(proj[:, idx, :]))
            res.append(plan.update())
            res.append(plan.run())
            res.append(plan.run(A // 3))
            outs[variant] = res
            plan.close()
        for i, (a, b) in enumerate(zip(outs["7"], outs["2"])):
            np.testing.assert_array_equal(a == 0, b == 0)
            assert rel_l2(a, b) < 1e-6, (storage, beam, i, rel_l2(a, b))
        np.testing.assert_array_equal(outs["7"][-3], outs["7"][-2])   # update == run, exactly
    monkeypatch.delenv("MINIRECAST_BP_VARIANT")


@pytest.mark.parametrize("beam", ["cone", "parallel"])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_oversized_slice_requests(workload, storage, beam, monkeypatch):
    """Slice requests whose records exceed the records buffer run the v2 loop in the
    default configuration (on the band copy for 16-bit rings, the row-pair copy for f32
    parallel beam), and the chunked records path in an explicit kernel-4/7 configuration;
    both equal kernel 2.  The buffer is shrunk to 64 KB so that these small slices need
    several chunks."""
    from minirecast import _hip, phantom, recon
    from minirecast.geometry import ParallelGeometry
    if beam == "cone":
        w, proj, planes, tasks, out_len = workload
        g = w.geom
        T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    else:
        g = ParallelGeometry(np.arange(120) * np.pi / 120, 24, 80, 1.0, 1.0)
        box = g.volume_box()
        ph = phantom.random_ellipsoids(5, 8, box, (0.1, 1.0))
        proj = phantom.project_parallel(ph, g)
        T, out_len = [], 0
        for k, sh in enumerate((-1.0, 0.0, 0.75)):
            T.append(recon.plane_task(box, (0.0, 0.0, 2.0), (0.0, 0.0, 1.0), 70, cor_shift=sh,
                                      out_offset=out_len))
            out_len += 70 * 70
        T = np.array(T, dtype=_hip.TASK_DTYPE)
    A = g.n_angles
    monkeypatch.setenv("MINIRECAST_BP_RECORD_KB", "64")
    outs, launches = {}, {}
    for name, variant in (("default", None), ("explicit", "7" if storage != "f32" else "4"),
                          ("ref", "2")):
        if variant is None:
            monkeypatch.delenv("MINIRECAST_BP_VARIANT", raising=False)
        else:
            monkeypatch.setenv("MINIRECAST_BP_VARIANT", variant)
        ctx = _hip.Context(0)
        if beam == "cone":
            ctx.set_cone(g.vectors, g.rows, g.cols, "ram-lak", storage)
        else:
            ctx.set_parallel(g.angles, g.rows, g.cols, 1.0, 1.0, "ram-lak", storage)
        ctx.push(np.arange(A), proj)
        outs[name] = ctx.bp_tasks(T, A, out_len)
        launches[name] = ctx.last_tile_launches()
    assert launches["default"] == 0               # the v2 loop (no kernel-4 launches)
    assert launches["explicit"] > 1                # records in several chunks
    for name in ("default", "explicit"):
        np.testing.assert_array_equal(outs[name] == 0, outs["ref"] == 0)
        assert rel_l2(outs[name], outs["ref"]) < 1e-6, (storage, name, rel_l2(outs[name], outs["ref"]))


def test_stream_graphs_survive_context_buffer_regrowth(workload):
    """A streaming plan's captured graphs keep their own f16 staging: an eager push of a
    larger batch and an adjoint with a host gradient regrow (free and reallocate) the
    context's staging buffers between two replays of the same batch-size graph, and the
    replay still equals a full pass bit for bit."""
    from minirecast import _hip
    w, proj, planes, tasks, out_len = workload
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    ctx = _hip.Context(0)
    ctx.set_cone(w.geom.vectors, w.geom.rows, w.geom.cols, w.window, "f16")
    plan = _hip.Plan(ctx, T, out_len)
    got = plan.stream(np.arange(0, 7), np.ascontiguousarray(proj[:, 0:7, :]))
    np.testing.assert_array_equal(got, plan.run())
    ctx.push(np.arange(7, 71), np.ascontiguousarray(proj[:, 7:71, :]))
    gq = ctx.bp_tasks_adjoint(T, ctx.num_views(), np.ones(out_len, np.float32),
                              w.geom.rows, w.geom.cols)
    assert gq.shape == (71, w.geom.rows, w.geom.cols) and np.isfinite(gq).all()
    for i0 in (71, 78):
        got = plan.stream(np.arange(i0, i0 + 7), np.ascontiguousarray(proj[:, i0:i0 + 7, :]))
        np.testing.assert_array_equal(got, plan.run())
    assert ctx.num_views() == 85
    plan.close()
