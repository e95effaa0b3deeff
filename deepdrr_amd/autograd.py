"""Differentiable DRR: density volume and 6-DoF volume pose -> intensity images.

deepdrr is not a differentiable renderer (README.md:47).  This module adds a
``torch.autograd.Function`` whose forward is the engine's single-volume
projector (drr_project_jobs: every view in one persistent launch, the
semantics of project_kernel.cu) and whose backward is the hand-written adjoint
kernel (drr_adjoint_jobs, likewise one launch): the per-ray march is
recomputed, density gradients are scattered atomically with the trilinear
weights, and pose gradients come from the sample-position derivatives.

Pose parameters ``theta = (t_x, t_y, t_z, w_x, w_y, w_z)`` perturb the volume
rigidly in world coordinates about ``center`` (default: the volume centre):
x -> R(w) (x - c) + c + t, R(w) = exp([w]_x).  The gradient with respect to t
is exact for any theta; with respect to w it is exact at w = 0 (it is the
left-perturbation derivative elsewhere), which is where a registration loop
evaluates it after folding the previous update into the pose.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _abi, geo
from ._abi_types import Job, Scene, make_view
from .material import attenuation_table
from .projector import world_from_index_rows
from .spectral_data import get_spectrum, spectrum_tables


def rodrigues(w: np.ndarray) -> np.ndarray:
    w = np.asarray(w, np.float64)
    th = float(np.linalg.norm(w))
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * (K @ K)


def perturbation(theta: np.ndarray, center: np.ndarray) -> np.ndarray:
    """4x4 world transform x -> R(w) (x - c) + c + t."""
    theta = np.asarray(theta, np.float64)
    R = rodrigues(theta[3:])
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = np.asarray(center, np.float64) - R @ np.asarray(center, np.float64) + theta[:3]
    return T


class PoseGeometry:
    """Views of one volume under a pose perturbation theta (host only)."""

    def __init__(self, shape, world_from_ijk, max_ray_length: float = -1.0, center=None):
        self.shape = tuple(int(s) for s in shape)
        self.world_from_ijk = np.asarray(world_from_ijk, np.float64)
        self.max_ray_length = float(max_ray_length)
        n = np.array(self.shape, np.float64)
        self.center = (np.asarray(center, np.float64) if center is not None
                       else (self.world_from_ijk @ np.append(n / 2, 1.0))[:3])

    def views(self, projections: Sequence[geo.CameraProjection], theta: np.ndarray):
        """[(View, source_world float32[3])] for the perturbed pose."""
        T = perturbation(theta, self.center)
        w_from_ijk = T @ self.world_from_ijk
        ijk_from_w = np.linalg.inv(w_from_ijk)
        out = []
        for proj in projections:
            src = np.array(proj.center_in_world)
            s_ijk = (ijk_from_w @ np.append(src, 1.0))[:3]
            w, h = proj.intrinsic.sensor_size
            vol = dict(ijk_from_world=ijk_from_w[:3].astype(np.float32).reshape(12),
                       source_ijk=s_ijk.astype(np.float32),
                       bounds_lo=np.full(3, -0.5, np.float32),
                       bounds_hi=np.array([s - 0.5 for s in self.shape], np.float32),
                       priority=0, enabled=1)
            out.append((make_view(world_from_index_rows(proj, 1.0), w, h, self.max_ray_length, [vol]),
                        src.astype(np.float32)))
        return out


class DifferentiableProjector(PoseGeometry):
    """Renders one volume whose density (and pose) are torch tensors.

    ``labels``: uint8 global material indices, shape [ni, nj, nk] like
    ``Volume.data``; ``materials``: names indexed by the labels;
    ``world_from_ijk``: 4x4 pose of the volume.
    """

    def __init__(self, labels: np.ndarray, materials: Sequence[str], world_from_ijk: np.ndarray,
                 step: float = 0.1, spectrum="90KV_AL40", device_id: int = 0,
                 max_ray_length: float = -1.0, center: Optional[np.ndarray] = None):
        import torch

        self.torch = torch
        self.device = torch.device("cuda", device_id)
        self.materials = list(materials)
        PoseGeometry.__init__(self, np.asarray(labels).shape, world_from_ijk, max_ray_length, center)
        self.labels_kji = torch.from_numpy(np.ascontiguousarray(
            np.moveaxis(np.asarray(labels, np.uint8), [0, 1, 2], [2, 1, 0]))).to(self.device)
        self.step = float(step)
        self.ctx = C.c_void_p()
        _abi.call("drr_ctx_create", device_id, C.byref(self.ctx))
        sc = Scene(step=self.step, n_materials=len(self.materials), attenuate_outside=0, air_index=0)
        _abi.call("drr_ctx_set_scene", self.ctx, C.byref(sc))
        self.E, self.pdf = spectrum_tables(get_spectrum(spectrum))
        self.mu = attenuation_table(self.materials, self.E)
        _abi.call("drr_ctx_set_spectrum", self.ctx, len(self.E), self.E.ctypes.data,
                  self.pdf.ctypes.data, self.mu.ctypes.data)

    def close(self):
        if self.ctx:
            _abi.lib().drr_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, density_kji):
        h = C.c_void_p()
        nz, ny, nx = density_kji.shape
        _abi.call("drr_volume_create", self.ctx, _abi.ptr(density_kji), _abi.ptr(self.labels_kji),
                  nx, ny, nz, 1, C.byref(h))
        return h

    def __call__(self, density, theta, projections: Sequence[geo.CameraProjection]):
        """density: torch float32 [ni, nj, nk] on the GPU; theta: torch float32 [6]
        (either may require grad).  Returns intensity [N, H, W] (equal sizes)."""
        return _DRRFunction.apply(density, theta, self, list(projections))


class _DRRFunction:
    pass


class _VolumeHandle:
    """Owns one device volume of the engine.  The handle is destroyed when
    ``release`` is called or, at the latest, when the owner is garbage
    collected (an autograd graph dropped without a backward pass frees it)."""

    def __init__(self, handle: C.c_void_p):
        import weakref

        self.handle = handle
        self._fin = weakref.finalize(self, _destroy_volume, handle.value)

    def release(self) -> None:
        self._fin()
        self.handle = None

    @property
    def alive(self) -> bool:
        return self._fin.alive


def _destroy_volume(value) -> None:
    if value:
        _abi.lib().drr_volume_destroy(C.c_void_p(value))


def _make_function():
    import torch

    class DRRFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, density, theta, dp: DifferentiableProjector, projections):
            dens_kji = density.detach().permute(2, 1, 0).contiguous().float()
            th = theta.detach().double().cpu().numpy()
            views = dp.views(projections, th)
            H, W = views[0][0].height, views[0][0].width
            if any(v.height != H or v.width != W for v, _ in views):
                raise ValueError("DifferentiableProjector: all projections need the same sensor size")
            jobs = (Job * len(views))()
            for i, (view, _) in enumerate(views):
                jobs[i].view = view
                jobs[i].volume_index = 0
                jobs[i].out_offset = i * H * W
            vol = _VolumeHandle(dp.upload(dens_kji))
            varr = (C.c_void_p * 1)(vol.handle.value)
            I = torch.empty((len(views), H, W), dtype=torch.float32, device=dp.device)
            need_grad = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
            # the area densities of the forward march, reused by the first backward pass
            AD = (torch.empty((len(views), len(dp.materials), H, W), dtype=torch.float32, device=dp.device)
                  if need_grad else None)
            stream = _abi.stream_ptr(torch.cuda.current_stream(dp.device))
            try:
                _abi.call("drr_project_jobs", dp.ctx, varr, 1, jobs, len(views), _abi.ptr(I), None,
                          _abi.ptr(AD), stream)
            except Exception:
                vol.release()
                raise
            if not need_grad:
                vol.release()
                vol = None
            else:
                ctx.save_for_backward(density)
            ctx.AD = AD
            ctx.dp = dp
            ctx.jobs = jobs
            ctx.sources = np.ascontiguousarray(np.stack([src for _, src in views]), np.float32)
            ctx.vol = vol
            ctx.shape = density.shape
            ctx.need_pose = theta.requires_grad
            ctx.need_density = density.requires_grad
            return I

        @staticmethod
        def backward(ctx, grad_out):
            dp = ctx.dp
            grad_out = grad_out.contiguous().float()
            ni, nj, nk = ctx.shape
            vol = ctx.vol
            if vol is None or not vol.alive:
                # a second backward through a retained graph: the first one
                # released the device volume; upload the saved density again
                (density,) = ctx.saved_tensors
                vol = _VolumeHandle(dp.upload(density.detach().permute(2, 1, 0).contiguous().float()))
            # pose-only backward (fixed density, e.g. registration) scatters nothing
            gD = (torch.zeros((nk, nj, ni), dtype=torch.float32, device=dp.device)
                  if ctx.need_density else None)
            gP = torch.zeros(6, dtype=torch.float32, device=dp.device) if ctx.need_pose else None
            stream = _abi.stream_ptr(torch.cuda.current_stream(dp.device))
            c = np.asarray(dp.center, np.float32)
            try:
                _abi.call("drr_adjoint_jobs", dp.ctx, vol.handle, ctx.jobs, len(ctx.jobs), _abi.ptr(grad_out),
                          None, ctx.sources.ctypes.data, c.ctypes.data, _abi.ptr(ctx.AD), _abi.ptr(gD),
                          _abi.ptr(gP), stream)
            finally:
                # the device volume (40 B per cell) and the saved area densities
                # are not kept between backward passes; a retained graph's next
                # backward re-uploads the density and re-marches the forward
                vol.release()
                ctx.vol = None
                ctx.AD = None
            grad_theta = gP if ctx.need_pose else None
            return (gD.permute(2, 1, 0) if gD is not None else None), grad_theta, None, None

    return DRRFunction


class _Lazy:
    fn = None


def _get():
    if _Lazy.fn is None:
        _Lazy.fn = _make_function()
    return _Lazy.fn


class _DRRFunctionProxy:
    @staticmethod
    def apply(*args):
        return _get().apply(*args)


_DRRFunction = _DRRFunctionProxy
