"""Scan geometries: the input contract of the reconstructor.

Projections are [rows, angles, cols] float32 arrays (geometry.py:1-6 of the reference).
Cone-beam geometries carry ASTRA ``cone_vec`` rows: source, detector centre, detector
u (column step) and v (row step), all in mm.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class ParallelGeometry:
    """Parallel beam around z (geometry.py:11-34): angles in radians, pixel pitch in mm.
    ``box`` = half-sizes (x, y, z) of the reconstructable volume; default is the
    detector's field of view."""
    angles: np.ndarray
    rows: int
    cols: int
    pixel_u: float = 1.0
    pixel_v: float = 1.0
    box: np.ndarray = None

    def detector_coords(self):
        u = (np.arange(self.cols) - (self.cols - 1) / 2.0) * self.pixel_u
        v = (np.arange(self.rows) - (self.rows - 1) / 2.0) * self.pixel_v
        return u, v

    @property
    def n_angles(self):
        return len(self.angles)

    def volume_box(self):
        if self.box is not None:
            return np.asarray(self.box, dtype=float)
        half_u = self.cols * self.pixel_u / 2.0
        return np.array([half_u, half_u, self.rows * self.pixel_v / 2.0])


@dataclass
class ConeVecGeometry:
    """Cone beam with one cone_vec row per view (geometry.py:37-57)."""
    vectors: np.ndarray
    rows: int
    cols: int
    angular_range: float = 2 * np.pi
    box: np.ndarray = None

    @property
    def n_angles(self):
        return self.vectors.shape[0]

    def volume_box(self):
        """Default box: the detector's field of view at the rotation axis."""
        if self.box is not None:
            return np.asarray(self.box, dtype=float)
        src, det, u, v = np.split(np.asarray(self.vectors[0], dtype=float), 4)
        mag = np.linalg.norm(det - src) / np.linalg.norm(src)
        half_u = self.cols * np.linalg.norm(u) / mag / 2.0
        half_v = self.rows * np.linalg.norm(v) / mag / 2.0
        return np.array([half_u, half_u, half_v])


def circular_cone_vectors(n_angles, sod, sdd, rows, cols, pixel, angular_range=2 * np.pi,
                          det_shift_u=0.0):
    """Ideal circular orbit around z (geometry.py:60-72): source at -sod along the ray,
    detector centre at (sdd - sod) along it, shifted det_shift_u pixels along u."""
    th = np.arange(n_angles) * angular_range / n_angles
    ray = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], axis=1)
    tan = np.stack([-np.sin(th), np.cos(th), np.zeros_like(th)], axis=1)
    vecs = np.zeros((n_angles, 12))
    vecs[:, 0:3] = -sod * ray
    vecs[:, 3:6] = (sdd - sod) * ray + det_shift_u * pixel * tan
    vecs[:, 6:9] = pixel * tan
    vecs[:, 11] = pixel
    return ConeVecGeometry(vecs, rows, cols, angular_range)
