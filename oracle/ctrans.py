"""ctypes wrapper of liboracle.so (the FP32 C transcription).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from deepdrr_amd._abi_types import Scene, View

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        l.ora_project.restype = C.c_int
        l.ora_project.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(View), C.POINTER(Scene),
                                  C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_int]
        l.ora_solid_angle.restype = None
        l.ora_solid_angle.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        _lib = l
    return _lib


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:
        return os.cpu_count() or 1


def project(volumes, view: View, scene: Scene, energies, pdf, mu, n_threads: int | None = None):
    """volumes: list of (density [nz,ny,nx] f32, labels [nz,ny,nx] u8).
    Returns intensity [H,W], photon_prob [H,W], area_density [M,H,W] (float32)."""
    n_threads = n_threads or default_threads()
    V = view.n_volumes
    dens = [np.ascontiguousarray(d, dtype=np.float32) for d, _ in volumes[:V]]
    labs = [np.ascontiguousarray(l, dtype=np.uint8) for _, l in volumes[:V]]
    dptr = (C.c_void_p * V)(*[d.ctypes.data for d in dens])
    lptr = (C.c_void_p * V)(*[l.ctypes.data for l in labs])
    dims = np.array([s for d in dens for s in d.shape[::-1]], dtype=np.int32)
    H, W, M = view.height, view.width, scene.n_materials
    I = np.zeros((H, W), np.float32)
    P = np.zeros((H, W), np.float32)
    AD = np.zeros((M, H, W), np.float32)
    e = np.ascontiguousarray(energies, np.float32)
    p = np.ascontiguousarray(pdf, np.float32)
    m = np.ascontiguousarray(mu, np.float32)
    rc = lib().ora_project(dptr, lptr, dims.ctypes.data, C.byref(view), C.byref(scene), len(e),
                           e.ctypes.data, p.ctypes.data, m.ctypes.data, I.ctypes.data, P.ctypes.data,
                           AD.ctypes.data, int(n_threads))
    if rc != 0:
        raise ValueError("ora_project rejected its arguments")
    return I, P, AD


def solid_angle(world_from_index, width: int, height: int) -> np.ndarray:
    w = np.ascontiguousarray(np.asarray(world_from_index, np.float32).reshape(9))
    out = np.zeros((height, width), np.float32)
    lib().ora_solid_angle(w.ctypes.data, width, height, out.ctypes.data)
    return out
