"""ctypes mirrors of the structs in include/deepdrr_amd.h (no library loading)."""
from __future__ import annotations

import ctypes as C

import numpy as np

DRR_MAX_VOLUMES = 4
DRR_MAX_MATERIALS = 16
DRR_MAX_BINS = 256
DRR_MESH_MAX_LAYERS = 4
DRR_MESH_MAX_HITS = 128
DRR_MESH_RAY_HITS = 256


class ViewVolume(C.Structure):
    _fields_ = [
        ("ijk_from_world", C.c_float * 12),
        ("source_ijk", C.c_float * 3),
        ("bounds_lo", C.c_float * 3),
        ("bounds_hi", C.c_float * 3),
        ("priority", C.c_int32),
        ("enabled", C.c_int32),
    ]


class View(C.Structure):
    _fields_ = [
        ("world_from_index", C.c_float * 9),
        ("max_ray_length", C.c_float),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("n_volumes", C.c_int32),
        ("vol", ViewVolume * DRR_MAX_VOLUMES),
    ]


MARCH_EXACT = 0  # DRR_MARCH_EXACT
MARCH_FAST = 1  # DRR_MARCH_FAST


class Scene(C.Structure):
    _fields_ = [
        ("step", C.c_float),
        ("n_materials", C.c_int32),
        ("attenuate_outside", C.c_int32),
        ("air_index", C.c_int32),
        ("march", C.c_int32),  # MARCH_EXACT or MARCH_FAST
    ]


class Job(C.Structure):
    _fields_ = [
        ("view", View),
        ("volume_index", C.c_int32),
        ("out_offset", C.c_int64),
    ]


class MeshPrim(C.Structure):
    _fields_ = [
        ("material", C.c_int32),
        ("density", C.c_float),
        ("additive", C.c_int32),
        ("subtractive", C.c_int32),
        ("layer", C.c_int32),
    ]


class BatchTuning(C.Structure):
    _fields_ = [
        ("tile_w", C.c_int32),
        ("tile_h", C.c_int32),
        ("waves_per_block", C.c_int32),
        ("blocks_per_cu", C.c_int32),
    ]


class MCParams(C.Structure):
    _fields_ = [
        ("n_materials", C.c_int32),
        ("n_bins", C.c_int32),
        ("energies_keV", C.c_float * DRR_MAX_BINS),
        ("cdf", C.c_float * DRR_MAX_BINS),
        ("a_pe", C.c_float * DRR_MAX_MATERIALS),
        ("a_compton", C.c_float * DRR_MAX_MATERIALS),
        ("a_rayleigh", C.c_float * DRR_MAX_MATERIALS),
        ("k_pe", C.c_float),
        ("k_compton", C.c_float),
        ("k_rayleigh", C.c_float),
        ("max_density", C.c_float),
        ("source_world", C.c_float * 3),
        ("world_from_ijk", C.c_float * 12),
        ("ijk_from_world", C.c_float * 12),
        ("voxel_size", C.c_float * 3),
        ("detector_origin", C.c_float * 3),
        ("detector_u", C.c_float * 3),
        ("detector_v", C.c_float * 3),
        ("detector_normal", C.c_float * 3),
        ("det_width", C.c_int32),
        ("det_height", C.c_int32),
        ("cone_axis", C.c_float * 3),
        ("cone_cos", C.c_float),
        ("max_interactions", C.c_int32),
    ]


def make_view(world_from_index, width, height, max_ray_length, vols, allow_empty=False) -> View:
    """vols: list of dicts with ijk_from_world(12), source_ijk(3), bounds_lo(3),
    bounds_hi(3), priority, enabled (none only with ``allow_empty``, a scene of
    meshes alone)."""
    v = View()
    v.world_from_index[:] = [float(x) for x in np.asarray(world_from_index, np.float32).reshape(9)]
    v.max_ray_length = float(np.float32(max_ray_length))
    v.width = int(width)
    v.height = int(height)
    if not (0 if allow_empty else 1) <= len(vols) <= DRR_MAX_VOLUMES:
        raise ValueError(f"1..{DRR_MAX_VOLUMES} volumes per view are supported, got {len(vols)}")
    v.n_volumes = len(vols)
    for i, d in enumerate(vols):
        vv = v.vol[i]
        vv.ijk_from_world[:] = [float(x) for x in np.asarray(d["ijk_from_world"], np.float32).reshape(12)]
        vv.source_ijk[:] = [float(x) for x in np.asarray(d["source_ijk"], np.float32).reshape(3)]
        vv.bounds_lo[:] = [float(x) for x in np.asarray(d["bounds_lo"], np.float32).reshape(3)]
        vv.bounds_hi[:] = [float(x) for x in np.asarray(d["bounds_hi"], np.float32).reshape(3)]
        vv.priority = int(d["priority"])
        vv.enabled = int(d["enabled"])
    return v


def view_to_dict(v: View) -> dict:
    return dict(
        world_from_index=np.array(v.world_from_index[:], np.float32),
        max_ray_length=np.float32(v.max_ray_length),
        width=v.width,
        height=v.height,
        vols=[dict(ijk_from_world=np.array(v.vol[i].ijk_from_world[:], np.float32),
                   source_ijk=np.array(v.vol[i].source_ijk[:], np.float32),
                   bounds_lo=np.array(v.vol[i].bounds_lo[:], np.float32),
                   bounds_hi=np.array(v.vol[i].bounds_hi[:], np.float32),
                   priority=v.vol[i].priority, enabled=v.vol[i].enabled)
              for i in range(v.n_volumes)],
    )
