"""Seeded synthetic workloads of the streaming FDK/FBP benchmarks.

Each builder returns a ``Workload``: geometry, random-ellipsoid phantom, the requested
slices (reconstruct_plane arguments) and per-workload extras.  Projections are exact
analytic line integrals (phantom.project_* on the GPU).

  small_cone     10 ellipsoids, 180 views / 360 deg, SOD 400 / SDD 600, 96 x 96
                 detector at 0.3 mm, shepp-logan, 3 orthogonal 128^2 slices.
  stream_cone    30 ellipsoids, Poisson noise (I0 = 1e4), 600 views arriving in
                 interleaved batches of 10, 384 x 486 detector at 0.3 mm,
                 3 oblique 512^2 slices + a 64^3 preview after every batch.
  full_cone      30 ellipsoids, 1200 views, 1536 x 1944 detector at 0.075 mm,
                 ram-lak, 3 oblique 1024^2 slices and a 768^3 volume.
  parallel_cor   parallel beam, 1800 views / 180 deg, 64 x 2048 detector at unit
                 pitch, hann, 25 centre shifts (-3..+3 by 0.25) x axial slices
                 1024^2 (z = 0), 768^2 (z = +16), 512^2 (z = -16).
  diff_cone      256^3-box phantom, 360 views, 256 x 256 detector at 0.3 mm,
                 3 oblique 256^2 slices (differentiable back-projection).
"""
from dataclasses import dataclass, field

import numpy as np

from .geometry import ParallelGeometry, circular_cone_vectors
from .phantom import random_ellipsoids

OBLIQUE_NORMALS = ((1.0, 0.5, 0.8), (-0.3, 1.0, 0.4), (0.2, -0.1, 1.0))
ORTHO_NORMALS = ((0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
SOD, SDD = 400.0, 600.0


@dataclass
class Workload:
    name: str
    geom: object
    phantom: object
    window: str
    planes: list
    seed: int
    extra: dict = field(default_factory=dict)


def _cone(name, seed, views, rows, cols, pixel, count, att, window, planes, **extra):
    geom = circular_cone_vectors(views, SOD, SDD, rows, cols, pixel)
    ph = random_ellipsoids(seed, count, geom.volume_box(), att)
    return Workload(name, geom, ph, window, planes, seed, extra)


def small_cone():
    planes = [dict(origin=(0.0, 0.0, 0.0), normal=n, n=128) for n in ORTHO_NORMALS]
    return _cone("small_cone", 1, 180, 96, 96, 0.3, 10, (0.1, 1.0), "shepp-logan", planes)


def stream_cone():
    planes = [dict(origin=(0.0, 0.0, 0.0), normal=n, n=512) for n in OBLIQUE_NORMALS]
    return _cone("stream_cone", 2, 600, 384, 486, 0.3, 30, (0.05, 1.0), "shepp-logan", planes,
                 batch=10, order="interleaved", photons=1e4, preview=64)


def full_cone():
    planes = [dict(origin=(0.0, 0.0, 0.0), normal=n, n=1024) for n in OBLIQUE_NORMALS]
    return _cone("full_cone", 3, 1200, 1536, 1944, 0.075, 30, (0.05, 1.0), "ram-lak", planes,
                 volume=768)


def parallel_cor():
    views = 1800
    geom = ParallelGeometry(np.arange(views) * np.pi / views, 64, 2048, 1.0, 1.0)
    ph = random_ellipsoids(4, 30, geom.volume_box(), (0.05, 1.0))
    planes = [dict(origin=(0.0, 0.0, 0.0), normal=(0.0, 0.0, 1.0), n=1024),
              dict(origin=(0.0, 0.0, 16.0), normal=(0.0, 0.0, 1.0), n=768),
              dict(origin=(0.0, 0.0, -16.0), normal=(0.0, 0.0, 1.0), n=512)]
    return Workload("parallel_cor", geom, ph, "hann", planes, 4,
                    dict(shifts=np.arange(-3.0, 3.01, 0.25)))


def diff_cone():
    planes = [dict(origin=(0.0, 0.0, 0.0), normal=n, n=256) for n in OBLIQUE_NORMALS]
    return _cone("diff_cone", 5, 360, 256, 256, 0.3, 30, (0.05, 1.0), "shepp-logan", planes)


WORKLOADS = dict(small_cone=small_cone, stream_cone=stream_cone, full_cone=full_cone,
                 parallel_cor=parallel_cor, diff_cone=diff_cone)


def add_poisson_noise(proj, photons, seed):
    """counts = Poisson(I0 exp(-p k)), p = -ln(max(counts, 1) / I0) / k, with k chosen so
    that max(p k) = 2 (the phantom recipe of the reference CLI, __main__.py:72-75)."""
    p = np.asarray(proj, np.float64)
    kappa = 2.0 / max(float(p.max()), 1e-12)
    rng = np.random.default_rng(seed)
    counts = rng.poisson(photons * np.exp(-p * kappa))
    return (-np.log(np.maximum(counts, 1) / photons) / kappa).astype(np.float32)
