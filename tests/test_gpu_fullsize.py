"""GPU tests at the benchmark's full size: 1200 views of a 1536 x 1944 detector, three
oblique 1024^2 slices (1,985,962 in-box points), f32 and f16 rings in their default
back-projection configurations.

Checked: the float64 oracle on 3 x 256 sampled in-box pixels (f32 <= 1e-5, f16 <= 1e-3,
bf16 <= 1e-3 relative L2), on 512 sampled voxels of the 768 x 768 x 607 volume over
the box (reconstruct_volume(maxdim=768), recon.py:442-470; the z-column kernel, staged
through LDS for f16) and on 256 of the 384^3 grid (boxes too large to stage), stream == batch (interleaved arrival in batches of 40), exact linearity
of the f32 path (projections x 2 give outputs x 2 bit for bit: scaling by 2 is exact in
binary floating point), and the default f16 kernel (band ring) against kernel 2 (the
work-queue loop on the ring itself) on every output.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    import torch
    from minirecast import _hip, phantom, recon, workloads
    from minirecast.stream import arrival_order
    from oracle import oracle_c as OC
    w = workloads.full_cone()
    g = w.geom
    A, R, C = g.n_angles, g.rows, g.cols
    box = g.volume_box()
    names = ("f32", "f16", "bf16", "f32_stream", "f32_x2")
    ctx = {}
    for name in names:
        ctx[name] = _hip.Context(0)
        ctx[name].set_cone(g.vectors, R, C, w.window, name.split("_")[0])
    batch = 40
    raw = torch.empty(R * batch * C, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(2024)
    samp_idx, samp_pts = [], []
    tasks, off = [], 0
    for p in w.planes:
        pts, mask = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        sel = np.sort(rng.choice(np.flatnonzero(mask.ravel()), size=256, replace=False))
        samp_idx.append(off + sel)
        samp_pts.append(pts.reshape(-1, 3)[sel])
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], out_offset=off))
        off += p['n'] ** 2
    samp_idx = np.concatenate(samp_idx)
    n_slice = len(samp_idx)
    # the 768^3 volume: its grid over the box (volume_grid, recon.py:278-283)
    dims, vorigin, vh = recon.volume_grid(box, w.extra["volume"])
    nvox = int(np.prod(dims))
    vsel = np.sort(rng.choice(nvox, size=512, replace=False))
    ijk = np.stack(np.unravel_index(vsel, tuple(int(d) for d in dims)), axis=1)
    samp_pts.append(vorigin + vh * ijk)
    # a 384^3 grid: voxels of 3.4 detector pixels, whose z-column boxes do not fit the LDS
    # buffer of the staged kernel (its views are gathered from memory)
    dims2, vorigin2, vh2 = recon.volume_grid(box, 384)
    vsel2 = np.sort(rng.choice(int(np.prod(dims2)), size=256, replace=False))
    ijk2 = np.stack(np.unravel_index(vsel2, tuple(int(d) for d in dims2)), axis=1)
    samp_pts.append(vorigin2 + vh2 * ijk2)
    samp_pts = np.concatenate(samp_pts)
    acc64 = np.zeros(len(samp_pts))
    host = np.empty((R, batch, C), np.float32)
    OC.set_threads(16)

    def ingest(names_, order):
        for b0 in range(0, A, batch):
            idx = order[b0:b0 + batch]
            phantom.project_cone_device(w.phantom, type(g)(g.vectors[idx], R, C), ctx["f32"],
                                        raw.data_ptr())
            torch.cuda.synchronize()
            for name in names_:
                if name == "f32_x2":
                    raw.mul_(2.0)
                    torch.cuda.synchronize()
                ctx[name].push_device(idx, raw.data_ptr())
                if name == "f32_x2":
                    raw.mul_(0.5)
                    torch.cuda.synchronize()
            if "f32" in names_:
                ctx["f32"].d2h(host, raw.data_ptr())
                for j, v in enumerate(idx):
                    q64 = OC.fdk_filter_view(host, j, g.vectors[v], w.window)
                    OC.bp_cone_acc_view(q64, g.vectors[v], 0.0, samp_pts, acc64)

    ingest(("f32", "f16", "bf16", "f32_x2"), np.arange(A))
    ingest(("f32_stream",), arrival_order(A, "interleaved"))
    del raw
    torch.cuda.empty_cache()
    T = np.array(tasks, dtype=_hip.TASK_DTYPE)
    out = {name: ctx[name].bp_tasks(T, A, off) for name in names}
    ref = acc64 * (np.pi / A)
    n_vol = len(vsel)
    return dict(w=w, ctx=ctx, T=T, out_len=off, out=out, samp_idx=samp_idx,
                ref=ref[:n_slice], vol_ref=ref[n_slice:n_slice + n_vol], vsel=vsel, dims=dims,
                vorigin=vorigin, vh=vh, vol2_ref=ref[n_slice + n_vol:], vsel2=vsel2,
                dims2=dims2, vorigin2=vorigin2, vh2=vh2)


def test_full_size_vs_float64_oracle(full):
    for name, tol in (("f32", 1e-5), ("f16", 1e-3), ("bf16", 1e-3)):
        got = full["out"][name][full["samp_idx"]].astype(np.float64)
        err = rel_l2(got, full["ref"])
        print(f"{name}: rel-L2 {err:.2e} vs float64 at {len(got)} sampled pixels")
        assert err < tol, (name, err)


def test_full_size_stream_equals_batch(full):
    """The ring holds views in arrival order, so interleaved arrival only changes the
    FP32 summation order over 1200 views (~2e-6 of the maximum here); the reference's
    own stream == batch check allows 1e-4."""
    a, b = full["out"]["f32"], full["out"]["f32_stream"]
    np.testing.assert_array_equal(a == 0, b == 0)
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(a))
    assert rel_l2(b, a) < 1e-6


def test_full_size_linearity_is_exact(full):
    np.testing.assert_array_equal(full["out"]["f32_x2"], 2.0 * full["out"]["f32"])


def test_full_size_plane_alone_equals_plane_in_batch(full):
    """Cone updates on the tile kernels sum fixed parts of the views (64 for 16-bit rings,
    32 for f32) and add the parts in order, in one launch whatever the request's size, so
    each plane alone (one third of the grid) equals its values within the three-plane
    batch bit for bit."""
    from minirecast import _hip
    A = full["w"].geom.n_angles
    T, out_len = full["T"], full["out_len"]
    for name in ("f32", "f16"):
        ctx = full["ctx"][name]
        batch = full["out"][name]
        for t in T:
            n2 = int(t['np']) * int(t['ns'])
            lo = int(t['out_offset'])
            one = ctx.bp_tasks(np.array([t], dtype=_hip.TASK_DTYPE), A, out_len)
            np.testing.assert_array_equal(one[lo:lo + n2], batch[lo:lo + n2], err_msg=name)
        assert ctx.last_tile_launches() == 1


def test_full_size_default_f16_kernel_vs_plain_kernel(full):
    ctx = full["ctx"]["f16"]
    cfg = ctx.bp_config()
    assert cfg["kernel"] == 7                 # the default for 16-bit storage
    ctx.set_bp_config(kernel=2, ppt=4, layout=0, depth=2)
    try:
        plain = ctx.bp_tasks(full["T"], full["w"].geom.n_angles, full["out_len"])
    finally:
        ctx.reset_bp_config()
    assert ctx.bp_config() == cfg
    got = full["out"]["f16"]
    np.testing.assert_array_equal(got == 0, plain == 0)
    assert rel_l2(got, plain) < 1e-6


@pytest.mark.parametrize("name,tol", [("f32", 1e-5), ("f16", 1e-3)])
def test_full_size_volume_vs_float64_oracle(full, name, tol):
    """The 768 x 768 x 607 volume (4.3e11 updates) in one device-resident request, each
    z-plane a tile stack of the volume task (the default volume kernel of each storage),
    at 512 sampled voxels against the float64 oracle."""
    from minirecast import recon
    ctx = full["ctx"][name]
    dims = full["dims"]
    nvox = int(np.prod(dims))
    A = full["w"].geom.n_angles
    T = recon.volume_task(full["vorigin"], full["vh"], dims, 0, int(dims[2]))
    dptr = ctx.alloc(4 * nvox)
    try:
        ctx.bp_tasks_device(T[None], A, dptr, nvox)
        vol = ctx.d2h(np.empty(nvox, np.float32), dptr)
    finally:
        ctx.free(dptr)
    got = vol[full["vsel"]].astype(np.float64)
    err = rel_l2(got, full["vol_ref"])
    kern = ctx.last_bp_kernel()
    print(f"{name}: volume {tuple(int(d) for d in dims)} rel-L2 {err:.2e} at {len(got)} voxels "
          f"(kernel {kern})")
    assert np.isfinite(vol).all() and err < tol, (name, err)
    # z columns, staged through LDS
    assert kern == 9, kern


@pytest.mark.parametrize("name,tol", [("f32", 1e-5), ("f16", 1e-3)])
def test_full_size_coarse_volume_vs_float64_oracle(full, name, tol):
    """The 384^3 grid over the box (voxels of 3.4 detector pixels: the staged kernel's
    boxes exceed its LDS buffer and every view is gathered from memory) at 256 sampled
    voxels against the float64 oracle."""
    from minirecast import recon
    ctx = full["ctx"][name]
    dims = full["dims2"]
    nvox = int(np.prod(dims))
    A = full["w"].geom.n_angles
    T = recon.volume_task(full["vorigin2"], full["vh2"], dims, 0, int(dims[2]))
    dptr = ctx.alloc(4 * nvox)
    try:
        ctx.bp_tasks_device(T[None], A, dptr, nvox)
        vol = ctx.d2h(np.empty(nvox, np.float32), dptr)
    finally:
        ctx.free(dptr)
    err = rel_l2(vol[full["vsel2"]].astype(np.float64), full["vol2_ref"])
    kern = ctx.last_bp_kernel()
    print(f"{name}: volume {tuple(int(d) for d in dims)} rel-L2 {err:.2e} (kernel {kern})")
    assert np.isfinite(vol).all() and err < tol, (name, err)
    assert kern == 9, kern
