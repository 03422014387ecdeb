"""Randomised GPU parity: seeded random requests against the float64 oracle.

Each case draws a geometry (cone beam with the detector's v axis along z or tilted out of
it, or parallel beam), a storage form, a detector size, a view count, the order and
batches in which the views arrive, and a request: oblique and axial planes of sizes that
are not tile multiples with random zoom and centre, a z-stack of planes, and for parallel
beam the trial images of a rotation-centre search on a grid of 1, 1/2 or 1/4 px with gaps
(the shift-group kernel when at least 4 per group of 6 px on average, else the plane
tiles) or off any grid (the plane tiles).  Every task is checked at up
to 150 sampled points against the oracle's float64 back-projection of its own float64
filter (relative L2 per task: f32 <= 1e-5, f16 <= 1e-3, bf16 <= 3e-3), and points
outside the box of masked planes must be exactly 0.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-5, "f16": 1e-3, "bf16": 3e-3}


def _geometry(rng):
    from minirecast.geometry import ConeVecGeometry, ParallelGeometry, circular_cone_vectors
    kind = ["cone", "cone_tilt", "parallel"][int(rng.integers(3))]
    A = int(rng.integers(40, 121))
    R = int(rng.integers(8, 48))
    C = int(rng.integers(16, 160))
    if kind == "parallel":
        pu, pv = float(rng.uniform(0.5, 1.5)), float(rng.uniform(0.5, 1.5))
        return kind, ParallelGeometry(np.arange(A) * np.pi / A, R, C, pu, pv)
    pixel = float(rng.uniform(0.2, 0.8))
    g = circular_cone_vectors(A, 400.0, 600.0, R, C, pixel,
                              det_shift_u=float(rng.uniform(-2.0, 2.0)))
    if kind == "cone_tilt":
        v = g.vectors.copy()
        a = np.deg2rad(float(rng.uniform(-3.0, 3.0)))
        u = v[:, 6:9] / np.linalg.norm(v[:, 6:9], axis=1, keepdims=True)
        v[:, 9:12] = pixel * (np.cos(a) * np.array([0.0, 0.0, 1.0]) + np.sin(a) * u)
        g = ConeVecGeometry(v, R, C)
    return kind, g


def _tasks(rng, kind, box):
    """Requests: each a list of task records, per task (sample points, flat output indices,
    mask or None, shift), its output length, and whether its shifts lie on a grid."""
    from minirecast import _hip, recon
    reqs = []
    T, checks, off = [], [], 0
    shift = 0.0 if kind != "parallel" else float(rng.uniform(-2.0, 2.0))
    for _ in range(int(rng.integers(1, 4))):
        normal = rng.normal(size=3)
        if rng.random() < 0.3:
            normal = np.array([0.0, 0.0, 1.0])
        n = int(rng.integers(20, 71))
        origin = rng.uniform(-0.3, 0.3, size=3) * box
        zoom = float(rng.uniform(0.7, 1.6))
        center = tuple(rng.uniform(-0.2, 0.2, size=3) * box)
        t = recon.plane_task(box, origin, normal, n, zoom=zoom, center=center, cor_shift=shift,
                             out_offset=off)
        pts, mask = recon.slice_points(box, origin, normal, n, zoom=zoom, center=center)
        T.append(t)
        checks.append((pts.reshape(-1, 3), off + np.arange(n * n), mask.ravel(), shift))
        off += n * n
    if rng.random() < 0.5:
        dims = tuple(int(x) for x in rng.integers(10, 40, size=3))
        h = 2.0 * float(box[0]) / max(dims)
        origin = -np.array(dims) * h / 2.0 + h / 2.0
        t = recon.volume_task(origin, h, dims, 0, dims[2], shift)
        t['out_offset'] = off
        ijk = np.stack(np.meshgrid(*(np.arange(d) for d in dims), indexing="ij"), -1).reshape(-1, 3)
        T.append(t)
        checks.append((origin + h * ijk, off + np.arange(int(np.prod(dims))), None, shift))
        off += int(np.prod(dims))
    reqs.append((np.array(T, dtype=_hip.TASK_DTYPE), checks, off, False))
    if kind == "parallel" and rng.random() < 0.8:
        # a rotation-centre search of its own: one axial plane, shifts on a grid (with gaps)
        # or off it
        T, checks, off = [], [], 0
        q = [1, 2, 4][int(rng.integers(3))]
        m = np.sort(rng.choice(np.arange(6 * q * 2 + 1), size=int(rng.integers(5, 14)),
                               replace=False))
        shifts = -3.0 + m / q
        on_grid = rng.random() >= 0.25
        if not on_grid:
            shifts = shifts + rng.uniform(-0.1, 0.1, size=len(shifts))
        n = int(rng.integers(24, 60))
        z = float(rng.uniform(-0.8, 0.8) * box[2])
        pts, mask = recon.slice_points(box, (0.0, 0.0, z), (0.0, 0.0, 1.0), n)
        for s in shifts:
            T.append(recon.plane_task(box, (0.0, 0.0, z), (0.0, 0.0, 1.0), n, cor_shift=float(s),
                                      out_offset=off))
            checks.append((pts.reshape(-1, 3), off + np.arange(n * n), mask.ravel(), float(s)))
            off += n * n
        reqs.append((np.array(T, dtype=_hip.TASK_DTYPE), checks, off, _grouped(shifts)))
    return reqs


def _grouped(shifts):
    """Whether one plane's shifts run the shift-group kernel: on a grid of 1/q px (q = 1,
    2, 4, the smallest that fits), split greedily into groups spanning at most 6 px, at
    least 4 shifts per group on average."""
    s = np.sort(np.asarray(shifts, float))
    for q in (1, 2, 4):
        m = (s - s[0]) * q
        if np.all(np.abs(m - np.round(m)) <= 1e-9 * np.maximum(1.0, np.abs(m))) and \
                len(np.unique(np.round(m))) == len(m):
            m = np.round(m).astype(int)
            groups, start = 0, None
            for x in m:
                if start is None or x - start > 6 * q:
                    groups, start = groups + 1, x
            return len(m) >= 2 and len(m) >= 4 * groups
    return False


@pytest.mark.parametrize("seed", range(12))
def test_random_request_vs_oracle(seed):
    from minirecast import _hip, phantom
    from oracle import oracle_c as OC
    rng = np.random.default_rng(5000 + seed)
    kind, g = _geometry(rng)
    storage = ["f32", "f16", "bf16"][int(rng.integers(3))]
    window = ["ram-lak", "shepp-logan", "hann", "cosine"][int(rng.integers(4))]
    box = g.volume_box()
    ph = phantom.random_ellipsoids(seed, 6, box, (0.1, 1.0))
    A, R, C = g.n_angles, g.rows, g.cols
    proj = (phantom.project_parallel(ph, g) if kind == "parallel"
            else phantom.project_cone(ph, g)).astype(np.float32)
    ctx = _hip.Context(0)
    if kind == "parallel":
        ctx.set_parallel(g.angles, R, C, g.pixel_u, g.pixel_v, window, storage)
    else:
        ctx.set_cone(g.vectors, R, C, window, storage)
    # views arrive in a random order and batches
    order = rng.permutation(A)
    cuts = np.sort(rng.choice(np.arange(1, A), size=int(rng.integers(0, 3)), replace=False))
    for idx in np.split(order, cuts):
        ctx.push(idx, np.ascontiguousarray(proj[:, idx, :]))
    reqs = _tasks(rng, kind, box)
    print(f"[random {seed}] {kind} {storage} {window} A={A} R={R} C={C}: requests of "
          f"{[len(r[0]) for r in reqs]} tasks", flush=True)
    # the float64 oracle: filtered views, then every view's back-projection
    if kind == "parallel":
        q64 = np.ascontiguousarray(np.transpose(OC.filter_rows(proj, g.pixel_u, window), (1, 0, 2)))
    else:
        q64 = [OC.fdk_filter_view(proj, v, g.vectors[v], window) for v in range(A)]
    errs, kerns = [], []
    for T, checks, out_len, grouped in reqs:
        out = ctx.bp_tasks(T, A, out_len)
        kerns.append(ctx.last_bp_kernel())
        assert (kerns[-1] == 10) == grouped, kerns
        errs += _check(rng, kind, g, q64, out, checks, A)
    print(f"[random {seed}] kernels {kerns}, worst rel-L2 {max(errs):.2e} over {len(errs)} tasks")
    assert max(errs) < TOL[storage], (seed, kind, storage, errs)


def _check(rng, kind, g, q64, out, checks, A):
    from oracle import oracle_c as OC
    if kind == "parallel":
        cos_t, sin_t = np.cos(g.angles), np.sin(g.angles)
    errs = []
    for pts, idx, mask, shift in checks:
        sel = np.arange(len(pts))
        if mask is not None:
            assert not out[idx[~mask]].any(), "a point outside the box is not 0"
            sel = np.flatnonzero(mask)
        if len(sel) == 0:
            continue
        sel = np.sort(rng.choice(sel, size=min(150, len(sel)), replace=False))
        acc = np.zeros(len(sel))
        if kind == "parallel":
            OC.bp_parallel_acc(q64, cos_t, sin_t, g.pixel_u, g.pixel_v, shift, pts[sel], acc)
        else:
            for v in range(A):
                OC.bp_cone_acc_view(q64[v], g.vectors[v], shift, pts[sel], acc)
        ref = acc * (np.pi / A)
        if np.linalg.norm(ref) == 0.0:
            assert not out[idx[sel]].any()
            continue
        errs.append(rel_l2(out[idx[sel]].astype(np.float64), ref))
    return errs
