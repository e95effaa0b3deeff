"""Monte-Carlo scatter with Woodcock tracking (MC-GPU method) on the GPU.

deepdrr's scatter path is removed (``projector.py:527-531``); its MC-GPU
inputs (``conv_to_mcgpu.py``, ``mcgpu_*.py``) describe the method this module
runs: a point source, the spectrum, photoelectric / Compton / Rayleigh
interactions and an image split into primary, single Compton, single
Rayleigh and multiple scatter.  Cross sections here are smooth synthetic
power laws per material, mu_x = rho * a_x * (E / 60 keV)^(-k_x), with a_x
chosen near the NIST values at 60 keV.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _abi, geo
from ._abi_types import MCParams

# (photoelectric, Compton, Rayleigh) mass attenuation at 60 keV, cm^2/g
SYNTHETIC_MU60: Dict[str, Tuple[float, float, float]] = {
    "air": (0.011, 0.170, 0.008),
    "lung": (0.018, 0.178, 0.011),
    "soft tissue": (0.020, 0.178, 0.012),
    "tissue_soft": (0.020, 0.178, 0.012),
    "bone": (0.090, 0.170, 0.025),
    "titanium": (0.780, 0.150, 0.060),
}
K_PE, K_COMPTON, K_RAYLEIGH = 3.0, 0.3, 1.8
CLASSES = ("primary", "compton", "rayleigh", "multiple")


def detector_geometry(proj: geo.CameraProjection, sdd: float):
    """World origin of pixel (0, 0)'s corner, per-pixel steps along u and v, unit normal."""
    R = proj.world_from_camera3d.data[:3, :3]
    kinv = np.linalg.inv(proj.intrinsic.data)
    src = np.array(proj.center_in_world)
    origin = src + sdd * (R @ kinv @ np.array([0.0, 0.0, 1.0]))
    du = sdd * (R @ kinv @ np.array([1.0, 0.0, 0.0]))
    dv = sdd * (R @ kinv @ np.array([0.0, 1.0, 0.0]))
    normal = R[:, 2] / np.linalg.norm(R[:, 2])
    return src, origin, du, dv, normal


def make_params(materials: Sequence[str], spectrum_arr: np.ndarray, proj: geo.CameraProjection,
                sdd: float, world_from_ijk: np.ndarray, max_density: float = 0.0,
                max_interactions: int = 64, mu60: Optional[Dict[str, Tuple[float, float, float]]] = None,
                exponents=(K_PE, K_COMPTON, K_RAYLEIGH)) -> MCParams:
    mu60 = mu60 or SYNTHETIC_MU60
    p = MCParams()
    p.n_materials = len(materials)
    E = np.asarray(spectrum_arr[:, 0], np.float64) / 1000.0
    w = np.clip(np.asarray(spectrum_arr[:, 1], np.float64), 0.0, None)
    cdf = np.cumsum(w) / w.sum()
    cdf[-1] = 1.0
    p.n_bins = len(E)
    for i in range(len(E)):
        p.energies_keV[i] = float(E[i])
        p.cdf[i] = float(cdf[i])
    for m, name in enumerate(materials):
        if name not in mu60:
            raise KeyError(f"no synthetic cross sections for material {name!r}")
        a = mu60[name]
        p.a_pe[m], p.a_compton[m], p.a_rayleigh[m] = (0.1 * a[0], 0.1 * a[1], 0.1 * a[2])  # 1/mm per g/cm^3
    p.k_pe, p.k_compton, p.k_rayleigh = exponents
    p.max_density = float(max_density)
    W = np.asarray(world_from_ijk, np.float64)
    Wi = np.linalg.inv(W)
    p.world_from_ijk[:] = [float(x) for x in W[:3].reshape(12)]
    p.ijk_from_world[:] = [float(x) for x in Wi[:3].reshape(12)]
    p.voxel_size[:] = [float(x) for x in np.linalg.norm(W[:3, :3], axis=0)]
    src, origin, du, dv, normal = detector_geometry(proj, sdd)
    wdt, hgt = proj.intrinsic.sensor_size
    p.source_world[:] = [float(x) for x in src]
    p.detector_origin[:] = [float(x) for x in origin]
    p.detector_u[:] = [float(x) for x in du]
    p.detector_v[:] = [float(x) for x in dv]
    p.detector_normal[:] = [float(x) for x in normal]
    p.det_width, p.det_height = int(wdt), int(hgt)
    center = origin + du * wdt / 2 + dv * hgt / 2
    axis = (center - src) / np.linalg.norm(center - src)
    corners = [origin, origin + du * wdt, origin + dv * hgt, origin + du * wdt + dv * hgt]
    cmin = min(float(np.dot(axis, (c - src) / np.linalg.norm(c - src))) for c in corners)
    p.cone_axis[:] = [float(x) for x in axis]
    p.cone_cos = cmin
    p.max_interactions = int(max_interactions)
    return p


def cone_solid_angle(p: MCParams) -> float:
    return 2.0 * np.pi * (1.0 - p.cone_cos)


def pixel_solid_angles(p: MCParams) -> np.ndarray:
    """Solid angle of each pixel seen from the source (small-pixel approximation)."""
    src = np.array(p.source_world[:])
    o = np.array(p.detector_origin[:])
    du = np.array(p.detector_u[:])
    dv = np.array(p.detector_v[:])
    n = np.array(p.detector_normal[:])
    vv, uu = np.meshgrid(np.arange(p.det_height) + 0.5, np.arange(p.det_width) + 0.5, indexing="ij")
    c = o[None, None] + uu[..., None] * du + vv[..., None] * dv - src
    r = np.linalg.norm(c, axis=-1)
    area = np.linalg.norm(np.cross(du, dv))
    return area * np.abs(c @ n) / r ** 3


class MonteCarlo:
    """Runs histories on the GPU for one volume handle of a drr_ctx."""

    def __init__(self, ctx, vol_handle, params: MCParams, device=None):
        import torch

        self.torch = torch
        self.ctx = ctx
        self.vol = vol_handle
        self.params = params
        self.device = device or torch.device("cuda", 0)
        H, W = params.det_height, params.det_width
        self.img = torch.zeros((4, H, W), dtype=torch.float32, device=self.device)
        self.img2 = torch.zeros((4, H, W), dtype=torch.float32, device=self.device)
        self.steps = torch.zeros(2, dtype=torch.int64, device=self.device)  # flights, brick steps
        self.n = 0

    def run(self, n_histories: int, seed: int = 12345, first: int = 0, chunk: int = 1 << 24,
            stream=None) -> None:
        s = _abi.stream_ptr(stream)
        done = 0
        while done < n_histories:
            k = min(chunk, n_histories - done)
            _abi.call("drr_mc_scatter", self.ctx, self.vol, C.byref(self.params), first + done, k,
                      C.c_uint64(seed), _abi.ptr(self.img), _abi.ptr(self.img2), _abi.ptr(self.steps),
                      _abi.ptr(self.steps) + 8, s)
            done += k
        self.n += n_histories

    def results(self):
        """Energy images [4, H, W] and sums of squares (float64) and the number
        of Woodcock flights."""
        self.torch.cuda.synchronize(self.device)
        return (self.img.cpu().numpy().astype(np.float64), self.img2.cpu().numpy().astype(np.float64),
                int(self.steps[0].item()))

    def brick_steps(self) -> int:
        """Brick majorants crossed by the flights so far."""
        self.torch.cuda.synchronize(self.device)
        return int(self.steps[1].item())

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        _abi.call("drr_ctx_last_kernel_ms", self.ctx, C.byref(ms))
        return ms.value


def scatter_for_projector(projector, projections: Sequence[geo.CameraProjection]) -> np.ndarray:
    """Scatter intensity to add to ``Projector`` images ([N, H, W], keV per photon
    aimed at the pixel), from ``projector.mc_scatter_histories`` histories per view.

    The MC image is the scattered energy reaching each pixel per emitted photon;
    dividing by the pixel's share of the emission cone (Omega_pixel / Omega_cone)
    puts it in the units of the primary intensity (expected energy of the photon
    travelling towards that pixel)."""
    if len(projector.volumes) != 1:
        raise NotImplementedError("scatter simulation supports single-volume projectors")
    vol = projector.volumes[0]
    n_hist = int(projector.mc_scatter_histories)
    out = []
    for proj in projections:
        p = make_params(projector.all_materials, projector.spectrum_arr, proj,
                        projector.source_to_detector_distance, vol.world_from_IJK.data,
                        max_density=float(np.max(vol.data)))
        mc = MonteCarlo(projector._ctx, projector._vols[0], p, projector._torch_device)
        mc.run(n_hist)
        img, _, _ = mc.results()
        sc = (img[1] + img[2] + img[3]) / n_hist
        frac = pixel_solid_angles(p) / cone_solid_angle(p)
        out.append((sc / frac).astype(np.float32))
    return np.stack(out)
