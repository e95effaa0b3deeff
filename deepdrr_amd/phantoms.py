"""Seeded synthetic CT phantoms and camera poses for the DRR workloads.

No patient data ships with the engine (deepdrr's tests download
``CT-chest.nrrd``, deepdrr/utils/test_utils.py:11-53); these phantoms stand in
for it with the shapes of the benchmark workloads:

* ``small_ct``: 128^3 float32 HU, 1 mm voxels, air background, a soft-tissue
  ellipsoid (HU ~ N(40, 20^2)) and three bone ellipsoids (HU ~ U(700, 1500)).
* ``clinical_ct``: 512x512x400 int16 HU (0.8 x 0.8 x 1 mm) body cylinder with
  lungs, spine, ribs-like ellipsoids and titanium rods, plus Gaussian texture
  (sigma 30 HU); five materials air/lung/soft tissue/bone/titanium.
* ``batch_poses``: poses sampled as deepdrr's README loop does
  (README.md:103-125): point uniform in a 100 mm ball about the centre,
  direction uniform in a 30 degree cone about anterior, source fraction
  U(0.5, 0.9).
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from . import geo
from .device import SimpleDevice
from .volume import FIVE_MATERIAL_CLASSES, Volume


def _centered_origin(shape, spacing) -> np.ndarray:
    # LPS anatomical_from_IJK = [[sx,0,0],[0,0,sz],[0,-sy,0]] (volume.py:246-269);
    # choose the origin so the volume centre lands on the world origin.
    R = np.array([[spacing[0], 0, 0], [0, 0, spacing[2]], [0, -spacing[1], 0]], dtype=np.float64)
    return -R @ (np.array(shape, dtype=np.float64) / 2)


def _ellipsoid(shape, center, semi) -> np.ndarray:
    i, j, k = np.ogrid[: shape[0], : shape[1], : shape[2]]
    return (((i - center[0]) / semi[0]) ** 2 + ((j - center[1]) / semi[1]) ** 2
            + ((k - center[2]) / semi[2]) ** 2) <= 1.0


def small_ct_hu(n: int = 128, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    hu = np.full(shape, -1000.0, dtype=np.float32)
    c = np.array(shape) / 2
    s = n / 128.0
    soft = _ellipsoid(shape, c, (50 * s, 40 * s, 55 * s))
    hu[soft] = rng.normal(40.0, 20.0, size=int(soft.sum())).astype(np.float32)
    for off, semi in (((0, 0, 0), (10, 8, 30)), ((20, -10, 5), (6, 6, 12)), ((-22, 12, -8), (7, 5, 9))):
        b = _ellipsoid(shape, c + np.array(off) * s, np.array(semi) * s)
        hu[b] = rng.uniform(700.0, 1500.0, size=int(b.sum())).astype(np.float32)
    return hu


def small_ct(n: int = 128, seed: int = 0) -> Volume:
    hu = small_ct_hu(n, seed)
    spacing = (1.0, 1.0, 1.0)
    return Volume.from_hu(hu, origin=_centered_origin(hu.shape, spacing), spacing=spacing,
                          anatomical_coordinate_system="LPS")


def clinical_ct_hu(shape=(512, 512, 400), seed: int = 0) -> np.ndarray:
    """int16 HU phantom; memory ~2 bytes/voxel plus float32 temporaries per slab."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    hu = np.empty(shape, dtype=np.int16)
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    cx, cy = nx / 2, ny / 2
    body = ((i - cx) / (0.42 * nx)) ** 2 + ((j - cy) / (0.30 * ny)) ** 2 <= 1.0
    lungL = ((i - cx + 0.17 * nx) / (0.11 * nx)) ** 2 + ((j - cy + 0.02 * ny) / (0.16 * ny)) ** 2 <= 1.0
    lungR = ((i - cx - 0.17 * nx) / (0.11 * nx)) ** 2 + ((j - cy + 0.02 * ny) / (0.16 * ny)) ** 2 <= 1.0
    spine = ((i - cx) / (0.045 * nx)) ** 2 + ((j - cy - 0.18 * ny) / (0.045 * ny)) ** 2 <= 1.0
    rods = np.zeros_like(body)
    for dx, dy in ((-0.03, 0.12), (0.03, 0.12)):
        rods |= ((i - cx - dx * nx) ** 2 + (j - cy - dy * ny) ** 2) <= (0.006 * nx) ** 2
    for k in range(nz):
        z = (k - nz / 2) / nz
        sl = np.full((nx, ny), -1000.0, dtype=np.float32)
        sl[body] = 40.0
        lung_here = abs(z) < 0.35
        if lung_here:
            sl[lungL | lungR] = -800.0
        sl[spine] = 1100.0
        # rib-like ring segments on some slices
        if (k // 24) % 2 == 0:
            ring = np.abs(np.sqrt(((i - cx) / (0.40 * nx)) ** 2 + ((j - cy) / (0.28 * ny)) ** 2) - 0.93) < 0.02
            sl[ring & body] = 900.0
        if abs(z) < 0.2:
            sl[rods] = 5900.0
        noise = rng.normal(0.0, 30.0, size=(nx, ny)).astype(np.float32)
        sl = np.where(sl > -1000.0, sl + noise, sl)
        hu[:, :, k] = np.clip(np.rint(sl), -1024, 8000).astype(np.int16)
    return hu


def clinical_ct(shape=(512, 512, 400), spacing=(0.8, 0.8, 1.0), seed: int = 0) -> Volume:
    hu = clinical_ct_hu(shape, seed)
    return Volume.from_hu(hu, origin=_centered_origin(shape, spacing), spacing=spacing,
                          anatomical_coordinate_system="LPS", material_classes=FIVE_MATERIAL_CLASSES)


def uniform_volume(shape, spacing=(1.0, 1.0, 1.0), density: float = 1.0,
                   material: str = "soft tissue") -> Volume:
    data = np.full(shape, density, dtype=np.float32)
    return Volume.from_parameters(data, {material: np.ones(shape, bool)},
                                  origin=_centered_origin(shape, spacing), spacing=spacing,
                                  anatomical_coordinate_system="LPS")


def ap_device(width: int, height: int, pixel: float, sdd: float = 1020.0, center=(0, 0, 0),
              fraction: float = 0.5) -> SimpleDevice:
    """AP view (direction +y, image up +z) through ``center``."""
    d = SimpleDevice(sensor_height=height, sensor_width=width, pixel_size=pixel,
                     source_to_detector_distance=sdd)
    d.set_view(point=list(center), direction=[0, 1, 0], up=[0, 0, 1], source_to_point_fraction=fraction)
    return d


def _cone_direction(rng, axis, half_angle_deg) -> np.ndarray:
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    cmin = math.cos(math.radians(half_angle_deg))
    c = rng.uniform(cmin, 1.0)
    phi = rng.uniform(0, 2 * math.pi)
    s = math.sqrt(max(0.0, 1 - c * c))
    tmp = np.array([1.0, 0, 0]) if abs(axis[0]) < 0.9 else np.array([0, 0, 1.0])
    e1 = np.cross(axis, tmp)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    return c * axis + s * (math.cos(phi) * e1 + math.sin(phi) * e2)


def batch_poses(n: int, seed: int = 0, radius: float = 100.0, cone_deg: float = 30.0,
                axis=(0, 1, 0)) -> List[Tuple[np.ndarray, np.ndarray, float]]:
    """(point, direction, source fraction) triples."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        while True:
            p = rng.uniform(-radius, radius, 3)
            if np.linalg.norm(p) <= radius:
                break
        d = _cone_direction(rng, axis, cone_deg)
        f = rng.uniform(0.5, 0.9)
        out.append((p, d, f))
    return out


def pose_projection(width: int, height: int, pixel: float, point, direction, fraction,
                    sdd: float = 1020.0, center=(0, 0, 0)) -> geo.CameraProjection:
    d = SimpleDevice(sensor_height=height, sensor_width=width, pixel_size=pixel,
                     source_to_detector_distance=sdd)
    up = np.array([0.0, 0.0, 1.0])
    d.set_view(point=list(np.asarray(center) + np.asarray(point)), direction=list(direction),
               up=list(up), source_to_point_fraction=fraction)
    return d.get_camera_projection()
