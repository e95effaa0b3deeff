"""Benchmark workloads and their algorithmic work counts.

Work unit of the forward march: one *in-volume sample*, a step whose alpha lies
inside an enabled volume's [minAlpha_vol, maxAlpha_vol]
(project_kernel.cu:470-474).  Its algorithmic footprint is 40 bytes: the 8
float32 density corners and 8 uint8 label corners a trilinear sample reads.
The count is evaluated per ray on the host in closed form from the float32 ray
setup: steps t with minA + t*step in [minA_v, maxA_v], t < N.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import phantoms
from ._abi_types import View

BYTES_PER_SAMPLE = 40          # 8 x f32 density + 8 x u8 label corners
ADJOINT_BYTES_PER_SAMPLE = 104  # 40 B gather + 8 x 4 B x 2 atomic read-modify-write
CLINICAL_CLASSES = ("air", "lung", "soft tissue", "bone", "titanium")
CLINICAL_UPPER_HU = (-900.0, -200.0, 300.0, 3000.0, 1.0e9)


def ray_intervals(view: View, stride: int = 1):
    """Per-ray [minAlpha_vol, maxAlpha_vol] and ray length of a single-volume view
    (FP64), for every ``stride``-th pixel along u and v."""
    F32 = np.float32
    W = np.array(view.world_from_index[:], F32)
    w, h = view.width, view.height
    vv, uu = np.meshgrid(np.arange(0, h, stride, dtype=F32), np.arange(0, w, stride, dtype=F32),
                         indexing="ij")
    u = uu.reshape(-1) + F32(0.5)
    v = vv.reshape(-1) + F32(0.5)
    rx = (u * W[0] + v * W[1]) + W[2]
    ry = (u * W[3] + v * W[4]) + W[5]
    rz = (u * W[6] + v * W[7]) + W[8]
    L = np.sqrt(rx * rx + ry * ry + rz * rz)
    rx, ry, rz = rx / L, ry / L, rz / L
    vol = view.vol[0]
    A = np.array(vol.ijk_from_world[:], F32).reshape(3, 4)
    s = np.array(vol.source_ijk[:], np.float64)
    lo = np.array(vol.bounds_lo[:], np.float64)
    hi = np.array(vol.bounds_hi[:], np.float64)
    r = np.stack([A[a, 0] * rx + A[a, 1] * ry + A[a, 2] * rz for a in range(3)], 1).astype(np.float64)
    mrl = float(view.max_ray_length)
    mn = np.zeros(len(u))
    mx = np.full(len(u), mrl if mrl > 0 else np.inf)
    for a in range(3):
        with np.errstate(divide="ignore", invalid="ignore"):
            a0 = (lo[a] - s[a]) / r[:, a]
            a1 = (hi[a] - s[a]) / r[:, a]
        mn = np.fmax(mn, np.fmin(a0, a1))
        mx = np.fmin(mx, np.fmax(a0, a1))
    return mn, mx, L.astype(np.float64)


def count_samples(view: View, step: float, stride: int = 1) -> int:
    """Closed-form count of in-volume samples of one single-volume view; with
    ``stride`` > 1 every stride-th pixel is counted and the sum scaled up."""
    mn, mx, L = ray_intervals(view, stride)
    minA = np.minimum(L, mn)
    maxA = np.maximum(0.0, mx)
    N = np.ceil((maxA - minA) / step)
    t0 = np.maximum(0.0, np.ceil((mn - minA) / step))
    t1 = np.minimum(N - 1, np.floor((mx - minA) / step))
    n = np.where(mx >= mn, np.maximum(0.0, t1 - t0 + 1), 0.0)
    scale = (view.width * view.height) / n.size
    return int(round(n.sum() * scale))


def clinical_label_map(materials):
    return np.array([materials.index(c) for c in CLINICAL_CLASSES], np.uint8)


def clinical_volume_device(ctx, materials, shape=(512, 512, 400), seed: int = 0):
    """Generate the clinical int16 HU phantom and ingest it on the GPU
    (drr_volume_create_from_hu); returns (handle, hu array)."""
    from . import _abi

    hu = phantoms.clinical_ct_hu(shape, seed)
    hu_kji = np.ascontiguousarray(np.moveaxis(hu, [0, 1, 2], [2, 1, 0]))
    upper = np.array(CLINICAL_UPPER_HU, np.float32)
    lab = clinical_label_map(materials)
    h = C.c_void_p()
    nz, ny, nx = hu_kji.shape
    _abi.call("drr_volume_create_from_hu", ctx, hu_kji.ctypes.data, nx, ny, nz, 0, len(upper),
              upper.ctypes.data, lab.ctypes.data, C.byref(h))
    return h, hu


def clinical_geometry_volume(shape=(512, 512, 400), spacing=(0.8, 0.8, 1.0)):
    """The clinical phantom's pose (no voxel data), for building views."""
    from . import geo
    from .volume import VolumePose

    origin = phantoms._centered_origin(shape, spacing)
    R = np.array([[spacing[0], 0, 0], [0, 0, spacing[2]], [0, -spacing[1], 0]], dtype=np.float64)
    return VolumePose(geo.FrameTransform.from_rt(R, origin), shape)
