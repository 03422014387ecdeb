"""GPU tests of the differentiable path: the hand-written adjoints against the forward
maps (dot-product tests), against the float64 oracle's adjoints, and the shift
derivative against central finite differences; then autograd end to end."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    from minirecast import _hip, recon
    from minirecast.geometry import circular_cone_vectors
    g = circular_cone_vectors(24, 400.0, 600.0, 20, 32, 0.6, det_shift_u=0.3)
    box = g.volume_box()
    planes = [dict(origin=(0.0, 0.0, 0.0), normal=(1.0, 0.5, 0.8), n=40),
              dict(origin=(0.3, -0.2, 0.5), normal=(-0.3, 1.0, 0.4), n=33)]
    tasks, off = [], 0
    for p in planes:
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], out_offset=off))
        off += p['n'] ** 2
    tasks = np.array(tasks, dtype=_hip.TASK_DTYPE)
    ctx = _hip.Context(0)
    ctx.set_cone(g.vectors, g.rows, g.cols, "shepp-logan")
    return g, box, planes, tasks, off, ctx


def _plane_points(box, planes):
    from minirecast import recon
    pts, masks = [], []
    for p in planes:
        pp, m = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        pts.append(pp.reshape(-1, 3))
        masks.append(m.ravel())
    return np.concatenate(pts), np.concatenate(masks)


def test_bp_adjoint_dot_product_and_oracle(setup):
    g, box, planes, tasks, out_len, ctx = setup
    rng = np.random.default_rng(0)
    A, R, C = g.n_angles, g.rows, g.cols
    q = rng.normal(size=(A, R, C)).astype(np.float32)
    ctx.reset_views()
    ctx.store(np.arange(A), q)
    y = ctx.bp_tasks(tasks, A, out_len)
    w = rng.normal(size=out_len).astype(np.float32)
    gq = ctx.bp_tasks_adjoint(tasks, A, w, R, C)
    lhs = float(np.dot(y.astype(np.float64), w))
    rhs = float(np.sum(q.astype(np.float64) * gq))
    assert abs(lhs - rhs) <= 1e-5 * np.sqrt(np.sum(y.astype(float) ** 2) * np.sum(w.astype(float) ** 2))
    # float64 oracle adjoint on the masked points
    pts, mask = _plane_points(box, planes)
    K = O.cone_constants(g.vectors)
    want = O.bp_cone_adjoint(np.where(mask, w, 0.0).astype(np.float64)[mask], K['s'], K['d'],
                             K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts[mask], np.pi / A,
                             R, C)
    assert rel_l2(gq, want) < 1e-5


def test_filter_adjoint_dot_product_and_oracle(setup):
    g, box, planes, tasks, out_len, ctx = setup
    rng = np.random.default_rng(1)
    A, R, C = g.n_angles, g.rows, g.cols
    x = rng.normal(size=(R, A, C)).astype(np.float32)
    ctx.reset_views()
    ctx.push(np.arange(A), x)
    fx = ctx.read_views(0, A, R, C)                      # F(W x): [A, R, C]
    z = rng.normal(size=(A, R, C)).astype(np.float32)
    ftz = ctx.filter_adjoint(np.arange(A), z)            # W F(z): [R, A, C]
    lhs = float(np.sum(fx.astype(np.float64) * z))
    rhs = float(np.sum(x.astype(np.float64) * ftz))
    assert abs(lhs - rhs) <= 1e-5 * np.linalg.norm(fx) * np.linalg.norm(z)
    K = O.cone_constants(g.vectors)
    assert rel_l2(ftz, O.fdk_filter_adjoint(z, K, "shepp-logan")) < 1e-5


def test_dshift_vs_oracle_and_finite_difference(setup):
    """dy/ds against the float64 analytic derivative, and against central differences.

    y(s) is piecewise linear in s (bilinear interpolation; the shift moves every
    sample along the detector columns), so a central difference with step h equals
    the mean of dy/ds over [s - h, s + h], not dy/ds at s: at h = 1e-3 px the kinks
    move the quotient by several percent.  The checks are therefore (1) the float64
    analytic derivative against a float64 quotient at h = 1e-6 (no kink inside for
    almost every sample), and (2) the quotient at h = 1e-3 of the float64 and of the
    float32 forward against the GPU derivative averaged over 128 midpoints of the step.
    The midpoint rule leaves an error of a few 1e-3 (each kink inside the step is
    placed to within 1/128 of it); the quotient differs from dy/ds at s by ~8% here,
    so a 1% bound still tells the two apart.
    """
    from minirecast import phantom
    g, box, planes, tasks, out_len, ctx = setup
    A, R, C = g.n_angles, g.rows, g.cols
    ph = phantom.random_ellipsoids(3, 6, box, (0.2, 1.0))
    proj = phantom.project_cone(ph, g)
    ctx.reset_views()
    ctx.push(np.arange(A), proj)
    s0 = 0.37
    T = tasks.copy()
    T['cor_shift'] = s0
    d = ctx.bp_tasks_dshift(T, A, out_len)
    q = ctx.read_views(0, A, R, C).astype(np.float64)
    pts, mask = _plane_points(box, planes)
    K = O.cone_constants(g.vectors)
    want = np.zeros(out_len)
    want[mask] = O.bp_cone_dshift(q, K['s'], K['d'], K['u'], K['uu'], K['vv'], K['n'], K['sdd'],
                                  K['sod'], pts[mask], np.pi / A, s0)
    assert rel_l2(d, want) < 1e-4
    wts = np.random.default_rng(2).normal(size=out_len)

    def loss64(sft):
        y = np.zeros(out_len)
        y[mask] = O.bp_cone(q, K['s'], K['d'] - sft * K['u'], K['uu'], K['vv'], K['n'], K['sdd'],
                            K['sod'], pts[mask], np.pi / A)
        return float(y @ wts)
    an64 = float(want @ wts)
    fd_fine = (loss64(s0 + 1e-6) - loss64(s0 - 1e-6)) / 2e-6
    assert abs(fd_fine - an64) <= 1e-6 * abs(an64), (fd_fine, an64)

    h, M = 1e-3, 128
    mean_d = np.zeros(out_len)
    for k in range(M):
        Tk = tasks.copy()
        Tk['cor_shift'] = s0 - h + (k + 0.5) * (2 * h / M)
        mean_d += ctx.bp_tasks_dshift(Tk, A, out_len).astype(np.float64) / M
    an_mean = float(mean_d @ wts)
    fd64 = (loss64(s0 + h) - loss64(s0 - h)) / (2 * h)

    def loss32(sft):
        Tx = tasks.copy()
        Tx['cor_shift'] = sft
        return float(ctx.bp_tasks(Tx, A, out_len).astype(np.float64) @ wts)
    fd32 = (loss32(s0 + h) - loss32(s0 - h)) / (2 * h)
    print(f"dL/ds at s: {float(d.astype(np.float64) @ wts):.6f} (float64 {an64:.6f}, "
          f"fd64 h=1e-6 {fd_fine:.6f}); mean over step {an_mean:.6f}, fd64 h=1e-3 {fd64:.6f}, "
          f"fd32 h=1e-3 {fd32:.6f}")
    assert abs(fd64 - an_mean) <= 1e-2 * abs(fd64), (fd64, an_mean)
    assert abs(fd32 - an_mean) <= 1e-2 * abs(fd64), (fd32, an_mean)
    assert abs(fd64 - an64) > 3e-2 * abs(fd64)      # the kinks matter at this step


@pytest.mark.parametrize("storage,tol", [("f32", 1e-4), ("f16", 1e-3)])
def test_dshift_split_over_views_vs_oracle(setup, storage, tol):
    """A request of few tiles over at least 32 views splits its views into parts summed
    by separate workgroups (the part sums then added in part order): 96 views here, and
    31 views (one part), each against the float64 analytic derivative."""
    from minirecast import _hip, phantom, recon
    from minirecast.geometry import circular_cone_vectors
    _, _, planes, _, _, _ = setup
    g = circular_cone_vectors(96, 400.0, 600.0, 20, 32, 0.6, det_shift_u=0.3)
    box = g.volume_box()
    tasks, off = [], 0
    for p in planes:
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], cor_shift=0.37,
                                      out_offset=off))
        off += p['n'] ** 2
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    ctx = _hip.Context(0)
    ctx.set_cone(g.vectors, g.rows, g.cols, "shepp-logan", storage)
    ph = phantom.random_ellipsoids(6, 6, box, (0.2, 1.0))
    ctx.push(np.arange(g.n_angles), phantom.project_cone(ph, g))
    pts, mask = _plane_points(box, planes)
    K = O.cone_constants(g.vectors)
    for nv in (96, 31):
        d = ctx.bp_tasks_dshift(T, nv, off)
        q = ctx.read_views(0, nv, g.rows, g.cols).astype(np.float64)
        Kv = {k: v[:nv] for k, v in K.items()}
        want = np.zeros(off)
        want[mask] = O.bp_cone_dshift(q, Kv['s'], Kv['d'], Kv['u'], Kv['uu'], Kv['vv'], Kv['n'],
                                      Kv['sdd'], Kv['sod'], pts[mask], np.pi / nv, 0.37)
        err = rel_l2(d, want)
        print(f"{storage} {nv} views: dy/ds rel-L2 {err:.2e}")
        assert np.all(d[~mask] == 0) and err < tol, (nv, err)


def test_autograd_end_to_end(setup):
    import torch
    from minirecast import autograd, phantom
    g, box, planes, tasks, out_len, _ = setup
    op = autograd.FDKSlices(g, planes, "shepp-logan")
    ph = phantom.random_ellipsoids(4, 6, box, (0.2, 1.0))
    proj_np = phantom.project_cone(ph, g)
    proj = torch.tensor(proj_np, device="cuda", requires_grad=True)
    shift = torch.tensor(0.25, device="cuda", requires_grad=True)
    wts = torch.tensor(np.random.default_rng(3).normal(size=out_len), dtype=torch.float32,
                       device="cuda")
    y = autograd.fdk_slices(proj, shift, op)
    loss = (y * wts).sum()
    loss.backward()
    # float64 reference gradient with respect to the raw projections
    A, R, C = g.n_angles, g.rows, g.cols
    pts, mask = _plane_points(box, planes)
    K = O.cone_constants(g.vectors)
    D = K['d'] - 0.25 * K['u']
    gq = O.bp_cone_adjoint(wts.cpu().numpy().astype(np.float64)[mask], K['s'], D, K['uu'], K['vv'],
                           K['n'], K['sdd'], K['sod'], pts[mask], np.pi / A, R, C)
    gref = O.fdk_filter_adjoint(gq, K, "shepp-logan")
    assert rel_l2(proj.grad.cpu().numpy(), gref) < 1e-5
    # shift gradient: float64 central difference at h = 1e-6 on the views the
    # operator filtered (the analytic derivative at s), and at h = 1e-3 against the
    # operator's derivative averaged over the step (y is piecewise linear in s)
    op.load(proj.detach())
    q = op.ctx.read_views(0, A, R, C).astype(np.float64)
    w64 = wts.cpu().numpy().astype(np.float64)[mask]

    def loss64(sft):
        return float(O.bp_cone(q, K['s'], K['d'] - sft * K['u'], K['uu'], K['vv'], K['n'],
                               K['sdd'], K['sod'], pts[mask], np.pi / A) @ w64)
    fd = (loss64(0.25 + 1e-6) - loss64(0.25 - 1e-6)) / 2e-6
    print(f"shift grad {float(shift.grad):.6f}  fd64(1e-6) {fd:.6f}")
    assert abs(float(shift.grad) - fd) <= 1e-4 * abs(fd) + 1e-6
    h, M = 1e-3, 128
    mean_g = 0.0
    for k in range(M):
        ds = op.dshift(proj.detach(), 0.25 - h + (k + 0.5) * (2 * h / M))
        mean_g += float((ds * wts).sum()) / M
    fd3 = (loss64(0.25 + h) - loss64(0.25 - h)) / (2 * h)
    assert abs(fd3 - mean_g) <= 1e-2 * abs(fd3), (fd3, mean_g)


def test_new_projection_tensor_in_reused_memory_is_filtered(setup):
    """A projection tensor freed and a new one allocated in its memory (PyTorch's caching
    allocator hands the block back; the new tensor's version counter starts at 0 again)
    must be filtered anew: the operator is linear, so projections times 2 give slices
    times 2."""
    import torch
    from minirecast import autograd, phantom
    g, box, planes, tasks, out_len, _ = setup
    op = autograd.FDKSlices(g, planes, "shepp-logan")
    ph = phantom.random_ellipsoids(5, 6, box, (0.2, 1.0))
    base = torch.tensor(phantom.project_cone(ph, g), device="cuda")
    with torch.no_grad():
        p1 = base.clone()
        addr = p1.data_ptr()
        y1 = autograd.fdk_slices(p1, 0.0, op).clone()
        del p1
        p2 = base * 2.0                    # a new tensor, version 0
        reused = p2.data_ptr() == addr
        y2 = autograd.fdk_slices(p2, 0.0, op)
    print(f"second tensor in the first one's memory: {reused}")
    err = rel_l2(y2.cpu().numpy().astype(np.float64), 2.0 * y1.cpu().numpy().astype(np.float64))
    assert err < 1e-6, (reused, err)
    # the same tensor changed in place is filtered again as well
    with torch.no_grad():
        p2.mul_(0.5)
        y3 = autograd.fdk_slices(p2, 0.0, op)
    assert rel_l2(y3.cpu().numpy().astype(np.float64), y1.cpu().numpy().astype(np.float64)) < 1e-6


def test_adjoint_gather_and_scatter_paths_agree(setup):
    """Requests of up to 4096 point tiles run the adjoint as a gather (one workgroup per
    detector tile and view, sums in LDS), larger ones as a scatter with atomics.  A stack
    of 65 planes of 256^2 points (4160 tiles) takes the scatter; its two halves (33 and
    32 planes) take the gather; the adjoint is linear, so the halves sum to the whole
    (two float32 evaluations of the same sums), and the dot-product test holds for
    both."""
    from minirecast import _hip, recon
    g, box, planes, tasks, out_len, ctx = setup
    print("[adjoint] gather vs scatter, 65 planes of 256^2", flush=True)
    A, R, C = g.n_angles, g.rows, g.cols
    rng = np.random.default_rng(5)
    q = rng.normal(size=(A, R, C)).astype(np.float32)
    ctx.reset_views()
    ctx.store(np.arange(A), q)
    h = 1.6 * float(np.max(box)) / 256
    origin = -0.8 * np.asarray(box, float)
    dims = (256, 256, 65)
    full = recon.volume_task(origin, h, dims, 0, 65)
    lo = recon.volume_task(origin, h, dims, 0, 33)
    hi = recon.volume_task(origin, h, dims, 33, 65)
    n = int(np.prod(dims))
    w = rng.normal(size=n).astype(np.float32)
    T = lambda t: np.array([t], dtype=_hip.TASK_DTYPE)
    g_full = ctx.bp_tasks_adjoint(T(full), A, w, R, C)            # scatter
    g_lo = ctx.bp_tasks_adjoint(T(lo), A, w, R, C)                # gather
    g_hi = ctx.bp_tasks_adjoint(T(hi), A, w, R, C)                # gather
    assert rel_l2(g_lo + g_hi, g_full) < 1e-5
    for t, gq in ((full, g_full), (lo, g_lo)):
        y = ctx.bp_tasks(T(t), A, n)
        lhs = float(np.dot(y.astype(np.float64), w))
        rhs = float(np.sum(q.astype(np.float64) * gq))
        assert abs(lhs - rhs) <= 1e-5 * np.sqrt(np.sum(y.astype(float) ** 2) * np.sum(w.astype(float) ** 2))


def test_adjoint_non_finite_weights_propagate(setup):
    """A weight of +inf reaches every detector cell its point taps (as float sums do);
    the workgroups that meet no such point keep their fixed-point sums, equal to the
    float64 adjoint."""
    g, box, planes, tasks, out_len, ctx = setup
    A, R, C = g.n_angles, g.rows, g.cols
    rng = np.random.default_rng(9)
    ctx.reset_views()
    ctx.store(np.arange(A), rng.normal(size=(A, R, C)).astype(np.float32))
    pts, mask = _plane_points(box, planes)
    w = rng.normal(size=out_len).astype(np.float32)
    k = int(np.flatnonzero(mask)[len(np.flatnonzero(mask)) // 2])
    w[k] = np.inf
    gq = ctx.bp_tasks_adjoint(tasks, A, w, R, C)
    K = O.cone_constants(g.vectors)
    with np.errstate(invalid="ignore", over="ignore"):
        want = O.bp_cone_adjoint(w.astype(np.float64)[mask], K['s'], K['d'], K['uu'], K['vv'],
                                 K['n'], K['sdd'], K['sod'], pts[mask], np.pi / A, R, C)
    bad = ~np.isfinite(want)
    assert bad.any()
    np.testing.assert_array_equal(~np.isfinite(gq), bad)
    # views whose every cell is finite (the inf point projects off the detector in none of
    # them would be rare): compare those cells
    ok = np.isfinite(want)
    assert rel_l2(gq[ok], want[ok]) < 1e-5
