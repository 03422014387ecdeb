"""GPU parity at the full sizes of the streaming, rotation-centre and differentiable
workloads, each against the float64 oracle:

  stream_cone   600 views of a 384 x 486 detector (Poisson noise) in 60 interleaved
                batches of 10; three oblique 512^2 slices and the 64^3 preview after
                every batch, from one replayed graph (running sums) and eagerly.
  parallel_cor  1800 views of a 64 x 2048 detector, hann; 25 shifts (-3..+3 px by
                0.25) x axial slices 1024^2 (z = 0), 768^2 (z = +16), 512^2 (z = -16);
                all 75 trial images in one persistent launch + device entropy scores.
  diff_cone     360 views of a 256 x 256 detector; three oblique 256^2 slices
                (120,266 in-box points): adjoints and gradients.
"""
import numpy as np
import pytest

from conftest import cor_trial_images, entropy_score, rel_l2

pytestmark = pytest.mark.gpu


def _sample(rng, mask, k):
    idx = np.flatnonzero(np.asarray(mask).ravel())
    return np.sort(rng.choice(idx, size=min(k, len(idx)), replace=False))


# ---------------------------------------------------------------- stream_cone
@pytest.fixture(scope="module")
def stream():
    from minirecast import _hip, phantom, recon, workloads
    from minirecast.stream import arrival_order
    from oracle import oracle_c as OC
    w = workloads.stream_cone()
    g = w.geom
    A, R, C = g.n_angles, g.rows, g.cols
    print(f"[stream_cone] {A} views {R}x{C}, batches of {w.extra['batch']}", flush=True)
    proj = phantom.project_cone(w.phantom, g)
    proj = workloads.add_poisson_noise(proj, w.extra["photons"], w.seed + 100)
    box = g.volume_box()
    dims, origin, h = recon.volume_grid(box, w.extra["preview"])
    rng = np.random.default_rng(22)
    tasks, off, samp_idx, samp_pts = [], 0, [], []
    for p in w.planes:
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], out_offset=off))
        pts, mask = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        sel = _sample(rng, mask, 256)
        samp_idx.append(off + sel)
        samp_pts.append(pts.reshape(-1, 3)[sel])
        off += p['n'] ** 2
    vt = recon.volume_task(origin, h, dims, 0, int(dims[2]))
    vt['out_offset'] = off
    tasks.append(vt)
    nvox = int(np.prod(dims))
    vsel = np.sort(rng.choice(nvox, size=256, replace=False))
    ijk = np.stack(np.unravel_index(vsel, tuple(int(d) for d in dims)), axis=1)
    samp_idx.append(off + vsel)
    samp_pts.append(origin + h * ijk)
    out_len = off + nvox
    samp_idx, samp_pts = np.concatenate(samp_idx), np.concatenate(samp_pts)
    OC.set_threads(16)
    acc = np.zeros(len(samp_pts))
    for v in range(A):
        q64 = OC.fdk_filter_view(proj, v, g.vectors[v], w.window)
        OC.bp_cone_acc_view(q64, g.vectors[v], 0.0, samp_pts, acc)
    order = arrival_order(A, w.extra["order"])
    b = w.extra["batch"]
    blocks = [(order[i:i + b], np.ascontiguousarray(proj[:, order[i:i + b], :]))
              for i in range(0, A, b)]
    return dict(w=w, T=np.array(tasks, dtype=_hip.TASK_DTYPE), out_len=out_len,
                blocks=blocks, samp_idx=samp_idx, ref=acc * (np.pi / A), dims=dims)


def test_stream_cone_graph_equals_eager_every_batch_and_oracle(stream):
    """The replayed graph of the plan (upload, filter, back-projection of the batch's
    views into running sums) against the eager reference-style requests after each of
    the 60 batches (reconstruct_planes + preview_volume over all views so far,
    recon.py:426-475), and the final state against the float64 oracle at 768 sampled
    slice pixels and 256 preview voxels; the f16 ring streams the same batches."""
    from minirecast import _hip, recon
    w, T, out_len = stream["w"], stream["T"], stream["out_len"]
    g = w.geom
    rec = recon.SliceReconstructor(None, g, window=w.window, verbose=False)
    plans = {}
    for storage in ("f32", "f16"):
        ctx = _hip.Context(0)
        ctx.set_cone(g.vectors, g.rows, g.cols, w.window, storage)
        plans[storage] = (ctx, _hip.Plan(ctx, T, out_len))
    worst = 0.0
    for bi, (idx, blk) in enumerate(stream["blocks"]):
        got = plans["f32"][1].stream(idx, blk)
        plans["f16"][1].stream(idx, blk, copy=False)
        rec.add_projections(idx, blk)
        imgs = rec.reconstruct_planes(w.planes)
        vol, _, _ = rec.preview_volume(w.extra["preview"])
        want = np.concatenate([im.ravel() for im in imgs] + [vol.ravel()])
        np.testing.assert_array_equal(got == 0, want == 0, err_msg=f"batch {bi}")
        err = rel_l2(got, want)
        worst = max(worst, err)
        assert err < 1e-5, (bi, err)          # two FP32 kernels (v2 loop, kernel 4)
    assert rec.n == g.n_angles
    f16 = plans["f16"][1].run()
    ref = stream["ref"]
    e32 = rel_l2(got[stream["samp_idx"]], ref)
    e16 = rel_l2(f16[stream["samp_idx"]], ref)
    print(f"[stream_cone] graph vs eager worst rel-L2 {worst:.2e}; final vs float64: "
          f"f32 {e32:.2e}, f16 {e16:.2e}")
    assert e32 < 1e-5 and e16 < 1e-3
    for ctx, plan in plans.values():
        plan.close()


# ---------------------------------------------------------------- parallel_cor
@pytest.fixture(scope="module")
def cor():
    from minirecast import _hip, phantom, recon, workloads
    from oracle import oracle_c as OC
    w = workloads.parallel_cor()
    g = w.geom
    A = g.n_angles
    print(f"[parallel_cor] {A} views {g.rows}x{g.cols}", flush=True)
    proj = phantom.project_parallel(w.phantom, g)
    box = g.volume_box()
    shifts = w.extra["shifts"]
    # every trial image of the three slices: task order (slice, shift), contiguous outputs
    tasks, sizes, off = [], [], 0
    for p in w.planes:
        for s in shifts:
            tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], cor_shift=float(s),
                                          out_offset=off))
            sizes.append(p['n'])
            off += p['n'] ** 2
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    # float64 oracle at 256 sampled in-box pixels per slice for four of the shifts
    check = (-3.0, 0.0, 2.0, 3.0)
    rng = np.random.default_rng(44)
    q64 = OC.filter_rows(proj, g.pixel_u, w.window)                  # [R, A, C] float64
    views = np.transpose(q64, (1, 0, 2))
    cos_t, sin_t = np.cos(g.angles), np.sin(g.angles)
    OC.set_threads(16)
    samples = []
    for j, p in enumerate(w.planes):
        pts, mask = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        sel = _sample(rng, mask, 256)
        for s in check:
            acc = np.zeros(len(sel))
            OC.bp_parallel_acc(views, cos_t, sin_t, g.pixel_u, g.pixel_v, s,
                               pts.reshape(-1, 3)[sel], acc)
            k = j * len(shifts) + int(np.flatnonzero(np.isclose(shifts, s))[0])
            samples.append((k, sel, acc * (np.pi / A)))
    del q64, views
    return dict(w=w, proj=proj, T=T, sizes=sizes, out_len=off, samples=samples, shifts=shifts)


def _cor_images(rec, T, out_len):
    """(device scores, trial images) of one plan of all trial images: one persistent
    launch over the tiles of every image + device scoring (mr_plan_cor_scores); the
    images are the plan's output the scores were computed from."""
    from minirecast import _hip
    plan = _hip.Plan(rec._ctx, T, out_len)
    try:
        scores = plan.cor_scores(rec.n)
        imgs = rec._ctx.d2h(np.empty(out_len, np.float32), plan.output_ptr())
    finally:
        plan.close()
    return scores, imgs


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16", 1e-3)])
def test_parallel_cor_full_size(cor, storage, tol):
    """The 75 trial images at sampled pixels against the float64 oracle for shifts
    -3, 0, +2, +3 px (each slice size), the device entropy scores against numpy's
    histogram of the very images they were computed from (estimate_cor's formula,
    recon.py:509-511), the same best shift; and the same through estimate_cor's own
    request path (one scoring request per slice), whose images differ from the plan's
    only by FP32 rounding (another launch shape), so that its scores agree with the
    plan's to the histogram's count resolution."""
    from minirecast import recon
    w, T, sizes, shifts = cor["w"], cor["T"], cor["sizes"], cor["shifts"]
    rec = recon.SliceReconstructor(cor["proj"], w.geom, window=w.window, verbose=False,
                                   storage=storage)
    scores, imgs = _cor_images(rec, T, cor["out_len"])
    offs = np.concatenate([[0], np.cumsum(np.square(sizes))])
    for k, sel, ref in cor["samples"]:
        got = imgs[offs[k]:offs[k + 1]][sel]
        err = rel_l2(got, ref)
        assert err < tol, (storage, k, err)
    host = np.array([entropy_score(imgs[offs[k]:offs[k + 1]]) for k in range(len(sizes))])
    np.testing.assert_allclose(scores, host, rtol=1e-12, atol=0)
    ns = len(shifts)
    for j, p in enumerate(w.planes):
        dev = scores[j * ns:(j + 1) * ns]
        assert int(np.argmax(dev)) == int(np.argmax(host[j * ns:(j + 1) * ns]))
        best, sc = rec.estimate_cor(shifts, n=p['n'], z=p['origin'][2], verbose=False)
        mine = [entropy_score(im) for im in cor_trial_images(rec, shifts, p['n'], p['origin'][2])]
        np.testing.assert_allclose(sc, mine, rtol=1e-12, atol=0)
        assert best == shifts[int(np.argmax(mine))]
        # a few of ~10^6 pixels may change histogram bin between the two launch shapes
        np.testing.assert_allclose(sc, dev, rtol=2e-4, atol=0)
        print(f"[parallel_cor] {storage} {p['n']}^2 z={p['origin'][2]:+.0f}: best "
              f"{shifts[int(np.argmax(dev))]:+.2f} px (estimate_cor {best:+.2f}), score spread "
              f"{dev.max() - dev.min():.3f}")


def test_parallel_cor_shift_equivariance_and_true_shift(cor):
    """What the unshifted search picks, explained.  Data whose detector moved by +2 px
    (np.roll of the columns, the reference's own recipe, selftest.py:84-87) must score
    at shift s + 2 what the original data scores at s, wherever every tap of every
    point stays on the detector: slices of 512 mm centred on the object.  There the
    search recovers the true centre (0 for the original, +2 for the moved data) on the
    1024^2 and 768^2 slices.  The full-box slices of the workload (2048 mm wide, 2-4 mm
    pixels for 1 mm detector pixels) do not: their corners lie outside the field of
    view of many views, those artefacts set the histogram range, and the score varies
    by ~0.1-0.3 between neighbouring 0.25 px trials, so their picks say nothing
    about the centre."""
    from minirecast import _hip, recon
    w, shifts = cor["w"], cor["shifts"]
    g = w.geom
    box = g.volume_box()
    zoom = [dict(p, size=512.0) for p in w.planes]
    T, off = [], 0
    for p in zoom:
        for s in shifts:
            T.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], size=p['size'],
                                      cor_shift=float(s), out_offset=off))
            off += p['n'] ** 2
    T = np.array(T, dtype=_hip.TASK_DTYPE)
    ns = len(shifts)
    res = {}
    for roll in (0, 2):
        proj = np.roll(cor["proj"], roll, axis=2) if roll else cor["proj"]
        rec = recon.SliceReconstructor(proj, g, window=w.window, verbose=False)
        res[roll] = _cor_images(rec, T, off)[0].reshape(3, ns)
        del rec
    step = 8                                            # 2 px in 0.25 px steps
    np.testing.assert_allclose(res[2][:, step:], res[0][:, :-step], rtol=1e-4, atol=0)
    for j in range(2):
        assert shifts[int(np.argmax(res[0][j]))] == 0.0, res[0][j]
        assert shifts[int(np.argmax(res[2][j]))] == 2.0, res[2][j]
    # the workload's full-box slices: neighbouring trials differ by far more than the
    # zoomed curves' smooth trend near their optimum
    full = cor["T"].copy()
    rec = recon.SliceReconstructor(cor["proj"], g, window=w.window, verbose=False)
    sc = _cor_images(rec, full, cor["out_len"])[0].reshape(3, ns)
    assert np.all(np.abs(np.diff(sc, axis=1)).max(axis=1) > 0.1)


# ---------------------------------------------------------------- diff_cone
@pytest.fixture(scope="module")
def diff():
    from minirecast import _hip, phantom, recon, workloads
    w = workloads.diff_cone()
    g = w.geom
    print(f"[diff_cone] {g.n_angles} views {g.rows}x{g.cols}", flush=True)
    box = g.volume_box()
    tasks, off, pts, masks = [], 0, [], []
    for p in w.planes:
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], out_offset=off))
        pp, m = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        pts.append(pp.reshape(-1, 3))
        masks.append(m.ravel())
        off += p['n'] ** 2
    proj = phantom.project_cone(w.phantom, g)
    ctx = _hip.Context(0)
    ctx.set_cone(g.vectors, g.rows, g.cols, w.window)
    return dict(w=w, T=np.array(tasks, dtype=_hip.TASK_DTYPE), out_len=off, ctx=ctx, proj=proj,
                pts=np.concatenate(pts), mask=np.concatenate(masks))


def test_diff_cone_adjoint_dot_product_and_oracle(diff):
    """<BP q, w> = <q, BP^T w> for random q and w at full size, and BP^T w against the
    float64 oracle's scatter (recomputed bilinear and FDK weights)."""
    from oracle import oracle as O
    w, T, out_len, ctx = diff["w"], diff["T"], diff["out_len"], diff["ctx"]
    g = w.geom
    A, R, C = g.n_angles, g.rows, g.cols
    rng = np.random.default_rng(50)
    q = rng.normal(size=(A, R, C)).astype(np.float32)
    ctx.reset_views()
    ctx.store(np.arange(A), q)
    y = ctx.bp_tasks(T, A, out_len).astype(np.float64)
    wt = rng.normal(size=out_len).astype(np.float32)
    gq = ctx.bp_tasks_adjoint(T, A, wt, R, C)
    lhs = float(y @ wt.astype(np.float64))
    rhs = float(np.sum(q.astype(np.float64) * gq))
    assert abs(lhs - rhs) <= 1e-5 * np.linalg.norm(y) * np.linalg.norm(wt)
    K = O.cone_constants(g.vectors)
    m = diff["mask"]
    want = O.bp_cone_adjoint(wt.astype(np.float64)[m], K['s'], K['d'], K['uu'], K['vv'], K['n'],
                             K['sdd'], K['sod'], diff["pts"][m], np.pi / A, R, C)
    assert rel_l2(gq, want) < 1e-5


def test_diff_cone_gradients_vs_float64(diff):
    """Loss = sum_i w_i y_i(p, s) over the three slices.  dL/dp (autograd through the
    hand-written adjoints) against the float64 chain W F BP^T w; dy/ds per point against
    the float64 analytic derivative, which a float64 central difference at h = 1e-6
    confirms; and the float64 central difference of the loss at h = 1e-3 px against the
    GPU derivative averaged over the step (y is piecewise linear in s, so the quotient
    is that mean)."""
    import torch
    from minirecast import autograd
    from oracle import oracle as O
    from oracle import oracle_c as OC
    w, out_len = diff["w"], diff["out_len"]
    g = w.geom
    A, R, C = g.n_angles, g.rows, g.cols
    op = autograd.FDKSlices(g, w.planes, w.window)
    proj = torch.tensor(diff["proj"], device="cuda", requires_grad=True)
    s0 = 0.35
    shift = torch.tensor(s0, device="cuda", requires_grad=True)
    wts = torch.tensor(np.random.default_rng(51).normal(size=out_len), dtype=torch.float32,
                       device="cuda")
    loss = (autograd.fdk_slices(proj, shift, op) * wts).sum()
    loss.backward()
    print("[diff_cone] GPU forward + backward done", flush=True)
    m, pts = diff["mask"], diff["pts"]
    w64 = wts.cpu().numpy().astype(np.float64)
    K = O.cone_constants(g.vectors)
    gq = O.bp_cone_adjoint(w64[m], K['s'], K['d'] - s0 * K['u'], K['uu'], K['vv'], K['n'],
                           K['sdd'], K['sod'], pts[m], np.pi / A, R, C)
    gref = O.fdk_filter_adjoint(gq, K, w.window)
    print("[diff_cone] float64 W F BP^T w done", flush=True)
    errp = rel_l2(proj.grad.cpu().numpy(), gref)
    assert errp < 1e-5, errp
    # the shift: the float64 chain on the views the operator filtered
    op.load(proj.detach())
    q = op.ctx.read_views(0, A, R, C).astype(np.float64)
    dy64 = O.bp_cone_dshift(q, K['s'], K['d'], K['u'], K['uu'], K['vv'], K['n'], K['sdd'],
                            K['sod'], pts[m], np.pi / A, s0)
    an64 = float(dy64 @ w64[m])
    print("[diff_cone] float64 dy/ds done", flush=True)
    OC.set_threads(16)
    pm = np.ascontiguousarray(pts[m])

    def y64(sft):
        acc = np.zeros(len(pm))
        for v in range(A):
            OC.bp_cone_acc_view(q[v], g.vectors[v], sft, pm, acc)
        return acc * (np.pi / A)

    def loss64(sft):
        return float(y64(sft) @ w64[m])

    # per point: y is piecewise linear in s (col moves by s, row stays: u . vv = 0 on a
    # circular orbit), so the quotient at h = 1e-6 is dy/ds except at the points where
    # one of the 360 samples has a column within h of an integer (a kink inside the
    # step): a fraction of about 2 h A = 7e-4 of the points
    h6 = 1e-6
    fd_pt = (y64(s0 + h6) - y64(s0 - h6)) / (2 * h6)
    bad = np.abs(fd_pt - dy64) > 1e-6 * np.abs(dy64).max()
    print(f"[diff_cone] float64 central difference h = 1e-6 done: {bad.mean():.2e} of the "
          f"points have a kink in the step", flush=True)
    assert bad.mean() <= 3 * (2 * h6 * A), bad.mean()
    # the GPU derivative per point (FP32): a sample whose column lies within FP32
    # rounding (~1e-5 px) of an integer may take the other one-sided slope of the
    # bilinear kink; every other point agrees with float64
    d32 = op.dshift(proj.detach(), s0).cpu().numpy().astype(np.float64)[m]
    off = np.abs(d32 - dy64) > 1e-4 * np.abs(dy64).max()
    print(f"[diff_cone] GPU dy/ds: {off.mean():.2e} of the points at a kink within FP32 "
          f"rounding", flush=True)
    assert off.mean() <= 1e-3, off.mean()
    assert rel_l2(d32[~off], dy64[~off]) < 1e-4
    # autograd's dL/ds is the weighted sum of that derivative; the kink samples move it
    # by a few percent at this size (4.3e7 samples), the step mean below pins it
    gs = float(shift.grad)
    assert abs(gs - float(d32 @ w64[m])) <= 1e-5 * np.abs(d32) @ np.abs(w64[m])
    assert abs(gs - an64) <= 5e-2 * abs(an64), (gs, an64)
    h, M = 1e-3, 64
    mean_g = 0.0
    for k in range(M):
        ds = op.dshift(proj.detach(), s0 - h + (k + 0.5) * (2 * h / M))
        mean_g += float((ds * wts).sum()) / M
    print("[diff_cone] GPU dy/ds over the step done", flush=True)
    fd3 = (loss64(s0 + h) - loss64(s0 - h)) / (2 * h)
    print(f"[diff_cone] dL/dp rel-L2 {errp:.2e}; dL/ds {gs:.6f} (float64 {an64:.6f}); "
          f"fd h=1e-3 {fd3:.6f} vs step mean {mean_g:.6f}")
    assert abs(fd3 - mean_g) <= 1e-2 * abs(fd3), (fd3, mean_g)
