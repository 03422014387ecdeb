"""Writes the golden fixtures of the reference's end-to-end composition.

Run in a container where the reference package is readable (it is read, never
shipped):  python tests/golden/make_golden.py /path/to/reference

Only the composition of SliceReconstructor (recon.py:289-519) is taken from the
reference: plane centring and basis, per-tap detector edges, scale = pi / n,
arrival-order storage, the cone and parallel COR conventions, the volume grid and
the histogram-entropy COR scores.  Each piece on its own is pinned by closed forms
and brute force in tests/test_oracle.py.  Inputs are the reference's own analytic
modified Shepp-Logan phantom (scale 50, as its selftest uses) on small detectors,
so every fixture stays below a few hundred kB.
"""
import os
import sys

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_composition.npz")


def main(ref_root):
    sys.path.insert(0, ref_root)
    from minirecast import recon as R               # noqa: E402  (reference, numpy backend)
    from minirecast.geometry import ParallelGeometry, circular_cone_vectors
    from minirecast.phantom import EllipsoidPhantom, project_cone, project_parallel
    from minirecast.stream import arrival_order

    ph = EllipsoidPhantom(scale=50.0)
    fx = {}

    # ---------------------------------------------------------------- cone
    cg = circular_cone_vectors(36, 300.0, 600.0, 32, 48, 4.0, det_shift_u=1.5)
    cproj = project_cone(ph, cg)
    fx["cone_vectors"] = cg.vectors
    fx["cone_shape"] = np.array([cg.rows, cg.cols])
    fx["cone_proj"] = cproj
    rec = R.SliceReconstructor(cproj, cg, backend="numpy", verbose=False)
    planes = [
        dict(origin=(3.0, -2.0, 5.0), normal=(1.0, 0.5, 0.8), n=40, size=110.0),
        dict(origin=(0.0, 0.0, 0.0), normal=(0.0, 0.0, 1.0), n=32),
        dict(origin=(0.0, 0.0, 1.0), normal=(0.3, 0.0, 1.0), n=24, zoom=1.5, center=(2.0, 0.0, 0.0)),
        dict(origin=(0.0, 5.0, 0.0), normal=(0.0, 1.0, 0.05), n=33),
    ]
    for i, p in enumerate(planes):
        img, pts = rec.reconstruct_plane(p["origin"], p["normal"], n=p["n"], size=p.get("size"),
                                         zoom=p.get("zoom", 1.0), center=p.get("center"))
        fx[f"cone_plane{i}"] = img
        fx[f"cone_plane{i}_args"] = np.array(
            list(p["origin"]) + list(p["normal"]) + [p["n"], p.get("size") or 0.0,
                                                     p.get("zoom", 1.0)]
            + list(p.get("center") or (np.nan, np.nan, np.nan)))
    img, _ = rec.reconstruct_plane((0.0, 0.0, 0.0), (1.0, 0.5, 0.8), n=32, cor_shift=1.25)
    fx["cone_plane_cor"] = img
    vol, origin, h = rec.preview_volume(12)
    fx["cone_preview"], fx["cone_preview_origin"], fx["cone_preview_h"] = vol, origin, h
    shifts = np.arange(-2.0, 2.01, 0.5)
    best, scores = rec.estimate_cor(shifts, n=32, verbose=False)
    fx["cone_cor_shifts"], fx["cone_cor_scores"], fx["cone_cor_best"] = shifts, scores, best
    # streamed in interleaved groups of 7 (arrival-order storage)
    live = R.SliceReconstructor(None, cg, backend="numpy", verbose=False)
    order = arrival_order(36, "interleaved")
    for i in range(0, 36, 7):
        idx = order[i:i + 7]
        live.add_projections(idx, np.ascontiguousarray(cproj[:, idx, :]))
    img, _ = live.reconstruct_plane((0.0, 0.0, 0.0), (0.2, 0.1, 1.0), n=32)
    fx["cone_stream_plane"] = img
    fx["cone_stream_order"] = order

    # ---------------------------------------------------------------- parallel
    pg = ParallelGeometry(np.arange(45) * np.pi / 45, 24, 64, 1.7, 1.7)
    pproj = project_parallel(ph, pg)
    fx["par_angles"] = pg.angles
    fx["par_shape"] = np.array([pg.rows, pg.cols])
    fx["par_pixel"] = np.array([pg.pixel_u, pg.pixel_v])
    fx["par_proj"] = pproj
    prec = R.SliceReconstructor(pproj, pg, window="hann", backend="numpy", verbose=False)
    img, _ = prec.reconstruct_plane((0.0, 0.0, 2.0), (0.0, 0.0, 1.0), n=40)
    fx["par_plane_axial"] = img
    img, _ = prec.reconstruct_plane((1.0, 0.0, 0.0), (1.0, 0.5, 0.8), n=36, size=100.0)
    fx["par_plane_oblique"] = img
    img, _ = prec.reconstruct_plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), n=32, cor_shift=-0.75)
    fx["par_plane_cor"] = img
    best, scores = prec.estimate_cor(np.arange(-1.0, 1.01, 0.5), n=32, verbose=False)
    fx["par_cor_scores"], fx["par_cor_best"] = scores, best

    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1024:.0f} kB, {len(fx)} arrays)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
