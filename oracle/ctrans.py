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
        l.ora_project_rows.restype = C.c_int
        l.ora_project_rows.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(View), C.POINTER(Scene),
                                       C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        from deepdrr_amd._abi_types import MeshPrim

        l.ora_project_mesh.restype = C.c_int
        l.ora_project_mesh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(View), C.POINTER(Scene),
                                       C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int, C.POINTER(MeshPrim), C.c_int, C.c_int, C.c_int,
                                       C.c_void_p, C.c_float, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int]
        l.ora_mesh_rays.restype = C.c_int
        l.ora_mesh_rays.argtypes = [C.POINTER(View), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                    C.POINTER(MeshPrim), C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                    C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
        l.ora_ray_triangle.restype = C.c_int
        l.ora_ray_triangle.argtypes = [C.c_void_p] * 5 + [C.POINTER(C.c_float), C.POINTER(C.c_int)]
        l.ora_solid_angle.restype = None
        l.ora_solid_angle.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        _lib = l
    return _lib


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:
        return os.cpu_count() or 1


def project(volumes, view: View, scene: Scene, energies, pdf, mu, n_threads: int | None = None, rows=None):
    """volumes: list of (density [nz,ny,nx] f32, labels [nz,ny,nx] u8).
    Returns intensity [H,W], photon_prob [H,W], area_density [M,H,W] (float32);
    with ``rows`` (detector row indices) only those rows, as [len(rows), W]
    images, each ray computed exactly as in the full view."""
    n_threads = n_threads or default_threads()
    V = view.n_volumes
    dens = [np.ascontiguousarray(d, dtype=np.float32) for d, _ in volumes[:V]]
    labs = [np.ascontiguousarray(l, dtype=np.uint8) for _, l in volumes[:V]]
    dptr = (C.c_void_p * V)(*[d.ctypes.data for d in dens])
    lptr = (C.c_void_p * V)(*[l.ctypes.data for l in labs])
    dims = np.array([s for d in dens for s in d.shape[::-1]], dtype=np.int32)
    W, M = view.width, scene.n_materials
    rsel = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    H = view.height if rsel is None else len(rsel)
    I = np.zeros((H, W), np.float32)
    P = np.zeros((H, W), np.float32)
    AD = np.zeros((M, H, W), np.float32)
    e = np.ascontiguousarray(energies, np.float32)
    p = np.ascontiguousarray(pdf, np.float32)
    m = np.ascontiguousarray(mu, np.float32)
    if rsel is None:
        rc = lib().ora_project(dptr, lptr, dims.ctypes.data, C.byref(view), C.byref(scene), len(e),
                               e.ctypes.data, p.ctypes.data, m.ctypes.data, I.ctypes.data, P.ctypes.data,
                               AD.ctypes.data, int(n_threads))
    else:
        rc = lib().ora_project_rows(dptr, lptr, dims.ctypes.data, C.byref(view), C.byref(scene), len(e),
                                    e.ctypes.data, p.ctypes.data, m.ctypes.data, rsel.ctypes.data, len(rsel),
                                    I.ctypes.data, P.ctypes.data, AD.ctypes.data, int(n_threads))
    if rc != 0:
        raise ValueError("ora_project rejected its arguments")
    return I, P, AD


def project_mesh(volumes, view: View, scene: Scene, energies, pdf, mu, mesh, n_layers: int, max_hits: int,
                 source_world, sdd: float, rows=None, n_threads: int | None = None):
    """The transcription with meshes.  ``mesh`` = (vertices float32 [n,3],
    triangles int32 [m,3], triangle primitive int32 [m], MeshPrim table,
    n_prims) as ``deepdrr_amd.mesh.mesh_arrays`` returns; ``volumes`` may be
    empty (view.n_volumes == 0)."""
    n_threads = n_threads or default_threads()
    V = view.n_volumes
    dens = [np.ascontiguousarray(d, dtype=np.float32) for d, _ in volumes[:V]]
    labs = [np.ascontiguousarray(l, dtype=np.uint8) for _, l in volumes[:V]]
    dptr = (C.c_void_p * max(1, V))(*[d.ctypes.data for d in dens])
    lptr = (C.c_void_p * max(1, V))(*[l.ctypes.data for l in labs])
    dims = np.array([s for d in dens for s in d.shape[::-1]] or [0], dtype=np.int32)
    W, M = view.width, scene.n_materials
    rsel = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    H = view.height if rsel is None else len(rsel)
    I = np.zeros((H, W), np.float32)
    P = np.zeros((H, W), np.float32)
    AD = np.zeros((M, H, W), np.float32)
    e = np.ascontiguousarray(energies, np.float32)
    p = np.ascontiguousarray(pdf, np.float32)
    m = np.ascontiguousarray(mu, np.float32)
    verts, tris, tprim, prims, n_prims = mesh
    src = np.ascontiguousarray(source_world, np.float32)
    rc = lib().ora_project_mesh(dptr, lptr, dims.ctypes.data, C.byref(view), C.byref(scene), len(e),
                                e.ctypes.data, p.ctypes.data, m.ctypes.data, verts.ctypes.data, tris.ctypes.data,
                                tprim.ctypes.data, len(tris), prims, n_prims, n_layers, max_hits, src.ctypes.data,
                                float(sdd), None if rsel is None else rsel.ctypes.data,
                                0 if rsel is None else len(rsel), I.ctypes.data, P.ctypes.data, AD.ctypes.data,
                                int(n_threads))
    if rc != 0:
        raise ValueError("ora_project_mesh rejected its arguments")
    return I, P, AD


def mesh_rays(view: View, mesh, n_layers: int, max_hits: int, n_materials: int, source_world, sdd: float):
    """Per-ray mesh terms: intervals [H,W,K,2], counts [H,W], additive
    [n_layers * n_unique, H, W] and n_unique (K = n_layers * max_hits / 2)."""
    verts, tris, tprim, prims, n_prims = mesh
    H, W = view.height, view.width
    K = n_layers * max_hits // 2
    iv = np.zeros((H, W, K, 2), np.float32)
    cnt = np.zeros((H, W), np.int32)
    add = np.zeros((n_layers * 16, H, W), np.float32)
    src = np.ascontiguousarray(source_world, np.float32)
    nu = lib().ora_mesh_rays(C.byref(view), verts.ctypes.data, tris.ctypes.data, tprim.ctypes.data, len(tris),
                             prims, n_prims, n_layers, max_hits, n_materials, src.ctypes.data, float(sdd),
                             iv.ctypes.data, cnt.ctypes.data, add.ctypes.data)
    return iv, cnt, add[: n_layers * nu], nu


def ray_triangle(o, d, v0, v1, v2):
    """(hit, t, entry) of the oracle's FP32 Moller-Trumbore test."""
    arrs = [np.ascontiguousarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    t = C.c_float()
    ent = C.c_int()
    hit = lib().ora_ray_triangle(*[a.ctypes.data for a in arrs], C.byref(t), C.byref(ent))
    return bool(hit), t.value, bool(ent.value)


def solid_angle(world_from_index, width: int, height: int) -> np.ndarray:
    w = np.ascontiguousarray(np.asarray(world_from_index, np.float32).reshape(9))
    out = np.zeros((height, width), np.float32)
    lib().ora_solid_angle(w.ctypes.data, width, height, out.ctypes.data)
    return out


def _mc_lib():
    l = lib()
    if not getattr(l, "_mc_bound", False):
        from deepdrr_amd._abi_types import MCParams

        l.ora_mc.restype = C.c_long
        l.ora_mc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(MCParams),
                             C.c_int64, C.c_int64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int,
                             C.POINTER(C.c_long)]
        l.ora_sample_compton.restype = None
        l.ora_sample_compton.argtypes = [C.c_float, C.c_int, C.c_uint64, C.c_void_p, C.c_void_p]
        l.ora_sample_rayleigh.restype = None
        l.ora_sample_rayleigh.argtypes = [C.c_int, C.c_uint64, C.c_void_p]
        l.ora_mc_tally.restype = C.c_long
        l.ora_mc_tally.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(MCParams),
                                   C.c_int64, C.c_int64, C.c_uint64, C.c_void_p, C.c_void_p]
        l.ora_rotate.restype = None
        l.ora_rotate.argtypes = [C.c_float] * 5 + [C.c_void_p]
        l.ora_ranecu.restype = None
        l.ora_ranecu.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        l.ora_ranecu_seed.restype = None
        l.ora_ranecu_seed.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        l._mc_bound = True
    return l


def mc(density_kji, labels_kji, params, n_histories, seed=12345, first=0, n_threads=None, brick_steps=None):
    """CPU Monte Carlo: returns img [4,H,W], img2 [4,H,W] (float64) and the
    number of Woodcock flights; ``brick_steps`` (a list) receives the number of
    brick majorants crossed."""
    d = np.ascontiguousarray(density_kji, np.float32)
    l = np.ascontiguousarray(labels_kji, np.uint8)
    nz, ny, nx = d.shape
    H, W = params.det_height, params.det_width
    img = np.zeros((4, H, W), np.float64)
    img2 = np.zeros((4, H, W), np.float64)
    bs = C.c_long(0)
    steps = _mc_lib().ora_mc(d.ctypes.data, l.ctypes.data, nx, ny, nz, C.byref(params), first,
                             n_histories, C.c_uint64(seed), img.ctypes.data, img2.ctypes.data,
                             int(n_threads or default_threads()), C.byref(bs))
    if brick_steps is not None:
        brick_steps.append(int(bs.value))
    return img, img2, int(steps)


def sample_compton(E, n, seed=1):
    c = np.zeros(n, np.float32)
    e = np.zeros(n, np.float32)
    _mc_lib().ora_sample_compton(float(E), n, C.c_uint64(seed), c.ctypes.data, e.ctypes.data)
    return c, e


def sample_rayleigh(n, seed=1):
    c = np.zeros(n, np.float32)
    _mc_lib().ora_sample_rayleigh(n, C.c_uint64(seed), c.ctypes.data)
    return c


TALLY_KEYS = ("emitted", "scored", "photoelectric", "compton_recoil", "cutoff", "escaped")


def mc_tally(density_kji, labels_kji, params, n_histories, seed=12345, first=0):
    """Energy balance of CPU histories: dict of keV totals (TALLY_KEYS), the
    scored histories per class and the Woodcock step count."""
    d = np.ascontiguousarray(density_kji, np.float32)
    l = np.ascontiguousarray(labels_kji, np.uint8)
    nz, ny, nx = d.shape
    tally = np.zeros(len(TALLY_KEYS), np.float64)
    counts = np.zeros(4, np.int64)
    steps = _mc_lib().ora_mc_tally(d.ctypes.data, l.ctypes.data, nx, ny, nz, C.byref(params), first,
                                   n_histories, C.c_uint64(seed), tally.ctypes.data, counts.ctypes.data)
    return dict(zip(TALLY_KEYS, tally.tolist())), counts, int(steps)


def rotate(u, v, w, ct, phi):
    out = np.zeros(3, np.float32)
    _mc_lib().ora_rotate(float(u), float(v), float(w), float(ct), float(phi), out.ctypes.data)
    return out


def ranecu(seed, h, n):
    s1 = np.zeros(n, np.int32)
    s2 = np.zeros(n, np.int32)
    u = np.zeros(n, np.float32)
    _mc_lib().ora_ranecu(C.c_uint64(seed), C.c_uint64(h), n, s1.ctypes.data, s2.ctypes.data, u.ctypes.data)
    a, b = C.c_int(), C.c_int()
    _mc_lib().ora_ranecu_seed(C.c_uint64(seed), C.c_uint64(h), C.byref(a), C.byref(b))
    return (a.value, b.value), s1, s2, u
