"""Analytic ellipsoid phantoms (phantom.py of the reference) with projections computed
on the GPU: exact chord-length line integrals per detector ray in float64, stored
float32 in the [rows, angles, cols] layout.

``EllipsoidPhantom.table`` rows follow the reference (value, a, b, c, x0, y0, z0,
phi[deg], rotation about z).  ``random_ellipsoids`` builds seeded phantoms with full
3-D rotations for the benchmark workloads.
"""
import numpy as np

from . import _hip
from .geometry import ConeVecGeometry, ParallelGeometry

# Modified 3-D Shepp-Logan (phantom.py:11-22): value, a, b, c, x0, y0, z0, phi[deg]
MODIFIED_SHEPP_LOGAN = np.array([
    [1.0, .6900, .920, .810, 0.00, 0.0000, 0.00, 0],
    [-0.8, .6624, .874, .780, 0.00, -0.0184, 0.00, 0],
    [-0.2, .1100, .310, .220, 0.22, 0.0000, 0.00, -18],
    [-0.2, .1600, .410, .280, -0.22, 0.0000, 0.00, 18],
    [0.1, .2100, .250, .410, 0.00, 0.3500, -0.15, 0],
    [0.1, .0460, .046, .050, 0.00, 0.1000, 0.25, 0],
    [0.1, .0460, .046, .050, 0.00, -0.1000, 0.25, 0],
    [0.1, .0460, .023, .050, -0.08, -0.6050, 0.00, 0],
    [0.1, .0230, .023, .020, 0.00, -0.6060, 0.00, 0],
    [0.1, .0230, .046, .020, 0.06, -0.6050, 0.00, 0],
])


class EllipsoidPhantom:
    """Sum of constant-density ellipsoids in mm.  ``mats[k]`` maps x - centre[k] onto
    the unit sphere; ``vals[k]`` is the added density."""

    def __init__(self, scale=1.0, table=MODIFIED_SHEPP_LOGAN, vals=None, centers=None,
                 mats=None):
        self.scale = float(scale)
        if vals is not None:
            self.vals = np.asarray(vals, float)
            self.centers = np.asarray(centers, float).reshape(-1, 3)
            self.mats = np.asarray(mats, float).reshape(-1, 3, 3)
            return
        vals, cents, mats = [], [], []
        for v, a, b, c, x0, y0, z0, phi in np.asarray(table, float):
            p = np.deg2rad(phi)
            rot = np.array([[np.cos(p), -np.sin(p), 0.0], [np.sin(p), np.cos(p), 0.0],
                            [0.0, 0.0, 1.0]])
            inv_axes = np.diag([1.0 / (a * self.scale), 1.0 / (b * self.scale),
                                1.0 / (c * self.scale)])
            vals.append(v)
            cents.append(np.array([x0, y0, z0]) * self.scale)
            mats.append(inv_axes @ rot.T)
        self.vals, self.centers, self.mats = np.array(vals), np.array(cents), np.array(mats)

    def packed(self):
        """[K, 13] float64 records (value, centre, M row-major) for the GPU projector."""
        K = len(self.vals)
        E = np.zeros((K, 13))
        E[:, 0] = self.vals
        E[:, 1:4] = self.centers
        E[:, 4:13] = self.mats.reshape(K, 9)
        return E

    def evaluate(self, points):
        """Density at points (..., 3) (phantom.py:47-54)."""
        pts = np.asarray(points, dtype=np.float64)
        out = np.zeros(pts.shape[:-1])
        for v, c, M in zip(self.vals, self.centers, self.mats):
            y = (pts - c) @ M.T
            out += v * (np.einsum('...i,...i->...', y, y) <= 1.0)
        return out


def random_ellipsoids(seed, count, half, att=(0.1, 1.0)):
    """Seeded random phantom in a box of half-sizes ``half``: centres uniform in the
    inner 60% of the box, semi-axes 0.05-0.3 of the box, uniformly random 3-D
    rotations (QR of a Gaussian matrix, sign-fixed), additive densities U(att)."""
    rng = np.random.default_rng(seed)
    half = np.asarray(half, float)
    full = 2.0 * half
    vals, cents, mats = [], [], []
    for _ in range(int(count)):
        c = rng.uniform(-0.3, 0.3, 3) * full
        axes = rng.uniform(0.05, 0.3, 3) * full
        Q, Rm = np.linalg.qr(rng.normal(size=(3, 3)))
        Q = Q * np.sign(np.diag(Rm))
        vals.append(rng.uniform(att[0], att[1]))
        cents.append(c)
        mats.append(np.diag(1.0 / axes) @ Q.T)
    return EllipsoidPhantom(vals=vals, centers=cents, mats=mats)


def _ctx(ctx):
    return ctx if ctx is not None else _hip.Context(0)


def project_cone_device(phantom, geom, ctx, dptr):
    """Cone projections [rows, angles, cols] float32 into device memory ``dptr``."""
    ctx.phantom_cone(phantom.packed(), geom.vectors, geom.rows, geom.cols, dptr)


def project_parallel_device(phantom, geom, ctx, dptr, far=None):
    # the chord formula integrates the whole line, so the start distance only
    # mirrors the reference's 10 * scale (phantom.py:81)
    far = 10.0 * phantom.scale if far is None else far
    ctx.phantom_parallel(phantom.packed(), geom.angles, geom.rows, geom.cols, geom.pixel_u,
                         geom.pixel_v, far, dptr)


def _project(phantom, geom, ctx, fn):
    ctx = _ctx(ctx)
    shape = (geom.rows, geom.n_angles, geom.cols)
    out = np.empty(shape, dtype=np.float32)
    d = ctx.alloc(out.nbytes)
    try:
        fn(phantom, geom, ctx, d)
        ctx.d2h(out, d)
    finally:
        ctx.free(d)
    return out


def project_cone(phantom, geom: ConeVecGeometry, ctx=None):
    """Analytic cone-beam projections (phantom.py:86-98), computed on the GPU."""
    return _project(phantom, geom, ctx, project_cone_device)


def project_parallel(phantom, geom: ParallelGeometry, ctx=None):
    """Analytic parallel-beam projections (phantom.py:73-83), computed on the GPU."""
    return _project(phantom, geom, ctx, project_parallel_device)
