"""Differentiable FDK/FBP back-projection of slices (PyTorch autograd wrapper).

``FDKSlices`` maps raw projections p [rows, views, cols] (float32, on the GPU) and a
rotation-centre shift s (detector columns) to the values of a list of slices:

    y = BP_s(F(W p))          W: FDK cosine weight, F: ramp filter, BP_s: back-projection

Forward and backward both run in libminirecast_hip.so:
    dL/dp = W F(BP_s^T g)      (mr_bp_tasks_adjoint: scatter with the forward's bilinear
                                weights, recomputed; mr_filter_adjoint: F is its own
                                transpose, W is diagonal)
    dL/ds = sum_i g_i dy_i/ds  (mr_bp_tasks_dshift)
PyTorch only provides the tensors and the autograd graph.
"""
import weakref

import numpy as np
import torch

from . import _hip, recon
from .geometry import ConeVecGeometry, ParallelGeometry


class FDKSlices:
    """Differentiable slice operator for one geometry and a fixed list of planes
    (dicts with reconstruct_plane's origin, normal, n, size, zoom, center)."""

    def __init__(self, geom, planes, window="shepp-logan", device=0):
        self.geom = geom
        self.box = np.asarray(geom.volume_box(), float)
        self.ctx = _hip.Context(device)
        if isinstance(geom, ConeVecGeometry):
            self.ctx.set_cone(geom.vectors, geom.rows, geom.cols, window)
        elif isinstance(geom, ParallelGeometry):
            self.ctx.set_parallel(geom.angles, geom.rows, geom.cols, geom.pixel_u, geom.pixel_v,
                                  window)
        else:
            raise TypeError(type(geom))
        self.planes = list(planes)
        self.sizes = [int(p.get("n", 256)) for p in self.planes]
        self.out_len = int(sum(n * n for n in self.sizes))
        self.device = torch.device("cuda", device)
        self._loaded = None             # (weak reference to the loaded tensor, its version)
        # the planes' task records (the shift enters only their cor_shift field): built
        # once, since plane_task takes ~0.2 ms of host time per plane
        out, off = [], 0
        for p, n in zip(self.planes, self.sizes):
            out.append(recon.plane_task(self.box, p["origin"], p["normal"], n, p.get("size"),
                                        p.get("zoom", 1.0), p.get("center"), 0.0, off))
            off += n * n
        self._tasks = np.array(out, dtype=_hip.TASK_DTYPE)

    def tasks(self, shift):
        T = self._tasks.copy()
        T['cor_shift'] = float(shift)
        return T

    def _check(self, proj):
        g = self.geom
        if proj.shape != (g.rows, g.n_angles, g.cols) or proj.dtype != torch.float32:
            raise ValueError(f"projections must be float32 [{g.rows}, {g.n_angles}, {g.cols}]")
        if proj.device.type != "cuda":
            raise ValueError("projections must be on the GPU")

    def load(self, proj):
        """Filter proj into the device ring, skipped when this same tensor object, unchanged
        since (same version counter), is the one loaded.  The address alone does not
        identify the data: a new tensor can take a freed tensor's memory with the same
        version number."""
        self._check(proj)
        if self._loaded is not None:
            ref, version = self._loaded
            if ref() is proj and proj._version == version:
                return
        self._loaded = None
        src = proj.contiguous()
        torch.cuda.synchronize(src.device)
        self.ctx.reset_views()
        self.ctx.push_device(np.arange(self.geom.n_angles), src.data_ptr())
        self._loaded = (weakref.ref(proj), proj._version)

    def forward(self, proj, shift):
        self.load(proj)
        out = torch.empty(self.out_len, dtype=torch.float32, device=proj.device)
        torch.cuda.synchronize(proj.device)
        self.ctx.bp_tasks_device(self.tasks(shift), self.geom.n_angles, out.data_ptr(),
                                 self.out_len)
        return out

    def grad_projections(self, g, shift=0.0):
        """dL/dp for output weights g (float32 [out_len] on the GPU) at shift s."""
        A, R, C = self.geom.n_angles, self.geom.rows, self.geom.cols
        g = g.contiguous()
        gq = torch.empty((A, R, C), dtype=torch.float32, device=g.device)
        graw = torch.empty((R, A, C), dtype=torch.float32, device=g.device)
        torch.cuda.synchronize(g.device)
        self.ctx.bp_tasks_adjoint_device(self.tasks(shift), A, g.data_ptr(), self.out_len,
                                         gq.data_ptr())
        self.ctx.filter_adjoint_device(np.arange(A), gq.data_ptr(), graw.data_ptr())
        return graw

    def dshift(self, proj, shift):
        """dy/ds for every output element at shift s."""
        self.load(proj)
        out = torch.empty(self.out_len, dtype=torch.float32, device=proj.device)
        torch.cuda.synchronize(proj.device)
        self.ctx.bp_tasks_dshift_device(self.tasks(shift), self.geom.n_angles, out.data_ptr(),
                                        self.out_len)
        return out

    def split(self, flat):
        """Output vector -> list of [n, n] slices."""
        res, off = [], 0
        for n in self.sizes:
            res.append(flat[off:off + n * n].view(n, n))
            off += n * n
        return res


class _FDKFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj, shift, op):
        out = op.forward(proj, float(shift))
        ctx.op = op
        ctx.save_for_backward(proj, shift)
        return out

    @staticmethod
    def backward(ctx, g):
        proj, shift = ctx.saved_tensors
        op = ctx.op
        gp = gs = None
        if ctx.needs_input_grad[0]:
            gp = op.grad_projections(g.float(), float(shift))
        if ctx.needs_input_grad[1]:
            ds = op.dshift(proj, float(shift))
            gs = (g.float() * ds).sum().to(shift.dtype).reshape(shift.shape)
        return gp, gs, None


def fdk_slices(proj, shift, op):
    """Differentiable slice values (flat float32 vector, op.split() for images)."""
    if not torch.is_tensor(shift):
        shift = torch.tensor(float(shift), dtype=torch.float32, device=proj.device)
    return _FDKFunction.apply(proj, shift, op)
