"""GPU tests of the z-column kernel (kernels 8 and 9) of volume requests.

A cone geometry whose detector column and depth do not depend on z (a circular orbit about
z with the detector's v axis along z) runs volume tasks (planes stacked along z) one voxel
column per thread: the reciprocal, column, column fraction and weight once per column and
view, the row per plane by one FMA.  By default each view's detector box is staged
through LDS (kernel 9: from the ring for f32, from the row-pair copy for 16-bit rings);
a 16-bit ring without room for the copy gathers from the ring (kernel 8).
Checked, for f32, f16 and bf16 rings:
- default volume requests of such a geometry run kernel 9 (mr_last_bp_kernel), and a
  16-bit ring whose copies are not allowed (copy budget 0) kernel 8, equal to it to
  float32 rounding;
- the volume equals the plane-tile kernels' (kernel 4 on the ring for f32, kernel 6 on
  the row-pair copy for 16-bit storage) to float32 rounding on every voxel;
- sampled voxels against the float64 oracle (f32 <= 1e-5, 16-bit <= 1e-3 relative L2),
  with and without a rotation-centre shift;
- grids that are not multiples of the tile (250 x 246 columns, 72 planes = 4.5 chunks),
  two slab tasks in one request, a view count below the stored views, voxels whose
  detector boxes straddle the LDS buffer size (staged and gathered views in one request);
- a detector tilted out of the vertical (no z separation) runs the plane tiles, and a
  volume too small to fill the device too.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

A, R, C, PIXEL = 96, 72, 96, 0.6
DIMS = (250, 246, 72)
# voxels of 1.25 mm (3 detector pixels), 3 chunks centred on the detector: some column
# tiles' boxes exceed the LDS buffer (about 120 of 17,300 tile-views, by a CPU emulation of
# the box rule), all inside the source orbit
COARSE, COARSE_H = (304, 304, 48), 1.25
WINDOW = "shepp-logan"


def _volume(ctx, T, n_views, nvox):
    dptr = ctx.alloc(4 * nvox)
    try:
        ctx.bp_tasks_device(T, n_views, dptr, nvox)
        vol = ctx.d2h(np.empty(nvox, np.float32), dptr)
    finally:
        ctx.free(dptr)
    return vol, ctx.last_bp_kernel()


@pytest.fixture(scope="module")
def zc():
    from minirecast import _hip, phantom, recon
    from minirecast.geometry import ConeVecGeometry, circular_cone_vectors
    from oracle import oracle_c as OC
    geom = circular_cone_vectors(A, 400.0, 600.0, R, C, PIXEL)
    box = geom.volume_box()
    ph = phantom.random_ellipsoids(11, 12, box, (0.1, 1.0))
    proj = phantom.project_cone(ph, geom)                  # [R, A, C]
    # the same orbit with the detector's v axis tilted by 2 degrees about the ray
    tv = geom.vectors.copy()
    a = np.deg2rad(2.0)
    u = tv[:, 6:9] / np.linalg.norm(tv[:, 6:9], axis=1, keepdims=True)
    tv[:, 9:12] = PIXEL * (np.cos(a) * np.array([0.0, 0.0, 1.0]) + np.sin(a) * u)
    tilted = ConeVecGeometry(tv, R, C)
    tproj = phantom.project_cone(ph, tilted)
    h = 2.0 * box[0] / DIMS[0]
    origin = -np.array([DIMS[0], DIMS[1], DIMS[2]]) * h / 2.0 + h / 2.0
    nvox = int(np.prod(DIMS))
    ctx = {}
    for st in ("f32", "f16", "bf16"):
        for mode in ("dflt", "tiles"):
            c = _hip.Context(0)
            c.set_cone(geom.vectors, R, C, WINDOW, st)
            c.push(np.arange(A), proj)
            if mode == "tiles":
                c.set_bp_config(4 if st == "f32" else 6, 4 if st != "f32" else 2, 0, 2)
            ctx[st, mode] = c
    ctx["tilted"] = _hip.Context(0)
    ctx["tilted"].set_cone(tilted.vectors, R, C, WINDOW, "f32")
    ctx["tilted"].push(np.arange(A), tproj)
    # float64 oracle at sampled voxels: unshifted and shifted, and the tilted geometry;
    # and at voxels of the coarse grid inside the box
    rng = np.random.default_rng(77)
    sel = np.sort(rng.choice(nvox, size=600, replace=False))
    ijk = np.stack(np.unravel_index(sel, DIMS), axis=1)
    pts = origin + h * ijk
    corigin = -np.array(COARSE) * COARSE_H / 2.0 + COARSE_H / 2.0
    cijk = np.stack(np.meshgrid(*(np.arange(n) for n in COARSE), indexing="ij"), -1).reshape(-1, 3)
    cpts = corigin + COARSE_H * cijk
    inside = np.flatnonzero(np.all(np.abs(cpts) <= box, axis=1))
    csel = np.sort(np.concatenate([rng.choice(inside, size=300, replace=False),
                                   rng.choice(len(cpts), size=300, replace=False)]))
    cin = np.all(np.abs(cpts[csel]) <= box, axis=1)
    ref = {0.0: np.zeros(len(pts)), 0.75: np.zeros(len(pts)), "tilted": np.zeros(len(pts)),
           "coarse": np.zeros(len(csel))}
    for v in range(A):
        q64 = OC.fdk_filter_view(proj, v, geom.vectors[v], WINDOW)
        OC.bp_cone_acc_view(q64, geom.vectors[v], 0.0, pts, ref[0.0])
        OC.bp_cone_acc_view(q64, geom.vectors[v], 0.75, pts, ref[0.75])
        OC.bp_cone_acc_view(q64, geom.vectors[v], 0.0, cpts[csel], ref["coarse"])
        tq = OC.fdk_filter_view(tproj, v, tilted.vectors[v], WINDOW)
        OC.bp_cone_acc_view(tq, tilted.vectors[v], 0.0, pts, ref["tilted"])
    for k in ref:
        ref[k] *= np.pi / A
    return dict(ctx=ctx, origin=origin, h=h, nvox=nvox, sel=sel, ref=ref, recon=recon,
                corigin=corigin, csel=csel, cin=cin, box=box)


@pytest.mark.parametrize("st,tol", [("f32", 1e-5), ("f16", 1e-3), ("bf16", 1e-3)])
@pytest.mark.parametrize("shift", [0.0, 0.75])
def test_zcol_volume_vs_tiles_and_oracle(zc, st, tol, shift):
    recon = zc["recon"]
    T = recon.volume_task(zc["origin"], zc["h"], DIMS, 0, DIMS[2], shift)[None]
    vol, kern = _volume(zc["ctx"][st, "dflt"], T, A, zc["nvox"])
    assert kern == 9, kern
    tiles, kern_t = _volume(zc["ctx"][st, "tiles"], T, A, zc["nvox"])
    assert kern_t in (4, 6), kern_t
    assert np.isfinite(vol).all()
    d = rel_l2(vol.astype(np.float64), tiles.astype(np.float64))
    dmax = float(np.max(np.abs(vol - tiles)) / np.max(np.abs(tiles)))
    err = rel_l2(vol[zc["sel"]].astype(np.float64), zc["ref"][shift])
    print(f"{st} shift {shift}: z columns vs plane tiles rel-L2 {d:.2e} (max {dmax:.2e}), "
          f"vs float64 {err:.2e}")
    # f16: the staged kernel lerps in packed half arithmetic (rounded to 2^-11 relative,
    # the order of the storage rounding), the plane tiles in float32
    dtol, mtol = (2e-4, 5e-4) if st == "f16" else (1e-6, 1e-5)
    assert d < dtol and dmax < mtol, (d, dmax)
    assert err < tol, err


def test_zcol_slabs_and_partial_views(zc):
    """Two slab tasks (planes 0-39 and 40-71) in one request equal the one-task volume;
    a request over the first 50 views equals the plane tiles over the same views."""
    recon = zc["recon"]
    ctx = zc["ctx"]["f32", "dflt"]
    whole, _ = _volume(ctx, recon.volume_task(zc["origin"], zc["h"], DIMS, 0, DIMS[2])[None], A,
                       zc["nvox"])
    T2 = np.concatenate([recon.volume_task(zc["origin"], zc["h"], DIMS, 0, 40)[None],
                         recon.volume_task(zc["origin"], zc["h"], DIMS, 40, DIMS[2])[None]])
    slabs, kern = _volume(ctx, T2, A, zc["nvox"])
    assert kern == 9
    assert rel_l2(slabs.astype(np.float64), whole.astype(np.float64)) < 1e-6
    T = recon.volume_task(zc["origin"], zc["h"], DIMS, 0, DIMS[2])[None]
    part, kern = _volume(ctx, T, 50, zc["nvox"])
    assert kern == 9
    part_t, _ = _volume(zc["ctx"]["f32", "tiles"], T, 50, zc["nvox"])
    assert rel_l2(part.astype(np.float64), part_t.astype(np.float64)) < 1e-6


def test_zcol_ring_form_equals_staged_form(zc):
    """Without room for the row-pair copy (a copy budget of 0) a 16-bit volume gathers
    its taps from the ring (kernel 8): the same taps as the staged form, whose f16 lerps
    are rounded to half precision (kernel 8 lerps in float32), so equal to that rounding."""
    recon = zc["recon"]
    T = recon.volume_task(zc["origin"], zc["h"], DIMS, 0, DIMS[2], 0.75)[None]
    ctx = zc["ctx"]["f16", "dflt"]
    staged, kern = _volume(ctx, T, A, zc["nvox"])
    assert kern == 9
    ctx.set_ring_copy_budget(0)
    try:
        ring, kern = _volume(ctx, T, A, zc["nvox"])
    finally:
        ctx.set_ring_copy_budget(-1)
    assert kern == 8, kern
    d = rel_l2(ring.astype(np.float64), staged.astype(np.float64))
    dmax = float(np.max(np.abs(ring - staged))) / float(np.max(np.abs(staged)))
    print(f"f16 gathered vs staged: rel-L2 {d:.2e}, max {dmax:.2e}")
    assert d < 2e-4 and dmax < 5e-4, (d, dmax)


@pytest.mark.parametrize("st", ["f32", "f16"])
def test_zcol_boxes_that_do_not_fit(zc, st):
    """The coarse grid (COARSE): the boxes of some column tiles exceed the 16 KB LDS
    buffer and are gathered from memory, the others staged, within one request.  Against
    the float64 oracle at 300 voxels inside the box and 300 anywhere in the grid (which
    reaches far outside the field of view), and against the plane tiles."""
    recon = zc["recon"]
    n = int(np.prod(COARSE))
    T = recon.volume_task(zc["corigin"], COARSE_H, COARSE, 0, COARSE[2])[None]
    vol, kern = _volume(zc["ctx"][st, "dflt"], T, A, n)
    assert kern == 9, kern
    tiles, _ = _volume(zc["ctx"][st, "tiles"], T, A, n)
    ref, cin = zc["ref"]["coarse"], zc["cin"]
    got, til = vol[zc["csel"]].astype(np.float64), tiles[zc["csel"]].astype(np.float64)
    err_in, err_all = rel_l2(got[cin], ref[cin]), rel_l2(got, ref)
    terr_in, terr_all = rel_l2(til[cin], ref[cin]), rel_l2(til, ref)
    d = rel_l2(vol.astype(np.float64), tiles.astype(np.float64))
    print(f"{st} coarse voxels: vs float64 inside {err_in:.2e} anywhere {err_all:.2e}; plane "
          f"tiles {terr_in:.2e} / {terr_all:.2e}; z columns vs tiles (every voxel) {d:.2e}")
    tol = 1e-5 if st == "f32" else 1e-3
    assert np.isfinite(vol).all() and err_in < tol and err_all < tol, (err_in, err_all)


def test_zcol_not_used_without_z_separation_or_enough_tiles(zc):
    recon = zc["recon"]
    T = recon.volume_task(zc["origin"], zc["h"], DIMS, 0, DIMS[2])[None]
    vol, kern = _volume(zc["ctx"]["tilted"], T, A, zc["nvox"])
    assert kern not in (8, 9), kern
    err = rel_l2(vol[zc["sel"]].astype(np.float64), zc["ref"]["tilted"])
    print(f"tilted detector: kernel {kern}, vs float64 {err:.2e}")
    assert err < 1e-5, err
    small = (64, 64, 64)
    n = int(np.prod(small))
    _, kern = _volume(zc["ctx"]["f32", "dflt"], recon.volume_task(zc["origin"], zc["h"], small, 0, 64)[None],
                      A, n)
    assert kern not in (8, 9), kern
