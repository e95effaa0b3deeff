"""Triangle meshes for the projector, without pyrender or OpenGL.

Mirrors deepdrr's mesh renderables: ``DRRMaterial``
(deepdrr/pyrenderdrr/material.py:5-184: density, additive, subtractive,
layer, drrMatName, tag), ``Mesh`` (deepdrr/vol/mesh.py:40-183: a pose,
``enabled``, ``from_stl``) holding primitives with ``positions``,
``indices`` and ``material`` as ``pyrender.Mesh.primitives`` do.  The
projector ray-casts them on the GPU (``drr_project_meshes``) in place of the
reference's depth peeling.

Semantics (projector.py:1055-1330, project_kernel.cu:363-375, 498-521,
569-579):
* an *additive* primitive adds ``density x chord length`` to its material's
  area density, per layer;
* a *subtractive* primitive hides the volumes inside it, and the additive
  primitives of lower layers where they overlap it;
* ``layer`` orders them (higher layers win); primitives on layers at or
  above the projector's ``mesh_layers`` are not rendered.
"""
from __future__ import annotations

import struct
from pathlib import Path
from typing import List, Optional, Sequence, Union

import numpy as np

from . import geo
from ._abi_types import MeshPrim
from .material import Material


class DRRMaterial:
    """pyrenderdrr's DRRMaterial, without the pyrender base class."""

    # the reference's default densities (g/cm^3) of from_name
    _default_densities = {
        "bone": 1.92, "soft tissue": 1.0, "tissue_soft": 1.0, "blood": 1.06, "muscle": 1.06,
        "air": 0.0012, "iron": 7.87, "lead": 11.34, "copper": 8.96, "lung": 0.26,
        "titanium": 4.51, "teflon": 2.2, "polyethylene": 0.94, "concrete": 2.3,
    }

    def __init__(self, name=None, density=1.0, additive=True, subtractive=False, layer=0,
                 drrMatName=None, tag=None, **_render_options):
        self.name = name
        self.density = density
        self.additive = additive
        self.subtractive = subtractive
        self.layer = layer
        self.drrMatName = drrMatName
        self.tag = tag

    @property
    def density(self) -> float:
        return self._density

    @density.setter
    def density(self, value):
        self._density = 1.0 if value is None else float(value)

    @property
    def additive(self) -> bool:
        return self._additive

    @additive.setter
    def additive(self, value):
        self._additive = bool(value)

    @property
    def subtractive(self) -> bool:
        return self._subtractive

    @subtractive.setter
    def subtractive(self, value):
        self._subtractive = bool(value)

    @property
    def layer(self) -> int:
        return self._layer

    @layer.setter
    def layer(self, value):
        self._layer = 0 if value is None else int(value)

    @property
    def drrMatName(self) -> str:
        return self._drrMatName

    @drrMatName.setter
    def drrMatName(self, value):
        if value is None:
            if self.name is None:
                raise ValueError("Please specify drrMatName")
            try:
                Material.from_string(self.name)
            except Exception:
                raise ValueError(f"Attempted to use material name {self.name} for drrMatName, but it was not "
                                 "found in material coefficients. Please specify drrMatName.")
            value = self.name
        self._drrMatName = value

    @classmethod
    def from_name(cls, name: str, **kwargs) -> "DRRMaterial":
        try:
            Material.from_string(name)
        except Exception:
            raise ValueError(f"Attempted to use material name {name}, but it was not found in material "
                             "coefficients. Please specify drrMatName")
        return cls(name=name, density=cls._default_densities[name], **kwargs)

    @classmethod
    def Iron(cls, **kw):
        return cls.from_name("iron", **kw)

    @classmethod
    def Bone(cls, **kw):
        return cls.from_name("bone", **kw)

    @classmethod
    def SoftTissue(cls, **kw):
        return cls.from_name("soft tissue", **kw)

    @classmethod
    def Titanium(cls, **kw):
        return cls.from_name("titanium", **kw)

    @classmethod
    def Lung(cls, **kw):
        return cls.from_name("lung", **kw)

    @classmethod
    def Air(cls, **kw):
        return cls.from_name("air", **kw)


class Primitive:
    """One triangle set with one material (a ``pyrender.Primitive``'s
    ``positions``, ``indices`` and ``material``)."""

    def __init__(self, positions, indices, material: DRRMaterial):
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        self.indices = np.asarray(indices, dtype=np.int64).reshape(-1, 3)
        if self.indices.size and (self.indices.min() < 0 or self.indices.max() >= len(self.positions)):
            raise ValueError("triangle indices out of range")
        if not isinstance(material, DRRMaterial):
            raise ValueError(f"unrecognized material type: {type(material)}.")
        self.material = material


class TriangleMesh:
    """Primitives plus visibility (the part of ``pyrender.Mesh`` the projector reads)."""

    def __init__(self, primitives: Sequence[Primitive], is_visible: bool = True):
        self.primitives = list(primitives)
        self.is_visible = is_visible

    @classmethod
    def from_arrays(cls, vertices, faces, material: Union[str, DRRMaterial] = "iron") -> "TriangleMesh":
        mat = DRRMaterial.from_name(material) if isinstance(material, str) else material
        return cls([Primitive(vertices, faces, mat)])


class Mesh:
    """A mesh renderable (deepdrr/vol/mesh.py:40-183).  Vertex positions are in
    the mesh's own (anatomical / IJK) frame; ``world_from_ijk`` places them."""

    def __init__(self, anatomical_from_IJK=None, world_from_anatomical=None, anatomical_from_ijk=None,
                 mesh: Optional[TriangleMesh] = None, **kwargs):
        if mesh is None:
            raise ValueError("mesh must be specified")
        self.mesh = mesh
        if anatomical_from_ijk is not None:
            anatomical_from_IJK = anatomical_from_ijk
        self.anatomical_from_IJK = (geo.FrameTransform.identity(3) if anatomical_from_IJK is None
                                    else geo.frame_transform(anatomical_from_IJK))
        self.world_from_anatomical = (geo.FrameTransform.identity(3) if world_from_anatomical is None
                                      else geo.frame_transform(world_from_anatomical))

    def set_enabled(self, enabled: bool) -> None:
        self.mesh.is_visible = bool(enabled)

    @property
    def enabled(self) -> bool:
        return bool(self.mesh.is_visible)

    @enabled.setter
    def enabled(self, enabled: bool) -> None:
        self.set_enabled(enabled)

    @property
    def world_from_ijk(self) -> geo.FrameTransform:
        return self.world_from_anatomical @ self.anatomical_from_IJK

    world_from_IJK = world_from_ijk

    def get_points(self) -> np.ndarray:
        return np.array(self.mesh.primitives[0].positions)

    def get_bounding_AABB(self):
        pts = np.concatenate([p.positions for p in self.mesh.primitives])
        return pts.min(0), pts.max(0)

    def get_center(self) -> geo.Point3D:
        lo, hi = self.get_bounding_AABB()
        return geo.point((lo + hi) / 2)

    @classmethod
    def from_stl(cls, path, material: Union[str, DRRMaterial, list] = "iron", convert_to_RAS: bool = False,
                 tag: Optional[str] = None, **kwargs) -> "Mesh":
        """STL file(s), one primitive each, with the default density of
        ``material`` (vol/mesh.py:86-134).  Faces are oriented outward (the
        reference runs trimesh's fix_normals); ``convert_to_RAS`` flips x
        and y (LPS <-> RAS)."""
        paths = path if isinstance(path, list) else [path]
        mats = material if isinstance(material, list) else [material] * len(paths)
        prims = []
        for p, m in zip(paths, mats):
            p = Path(p)
            if not p.exists():
                raise FileNotFoundError(f"Could not find file {p}")
            v, f = read_stl(p)
            if convert_to_RAS:
                v = v * np.array([-1.0, -1.0, 1.0])
                f = f[:, ::-1]
            f = orient_outward(v, f)
            mat = DRRMaterial.from_name(m, tag=tag) if isinstance(m, str) else m
            prims.append(Primitive(v, f, mat))
        return cls(mesh=TriangleMesh(prims), **kwargs)


def read_stl(path) -> (np.ndarray, np.ndarray):
    """Vertices [n, 3] (unique) and faces [m, 3] of a binary or ASCII STL file."""
    data = Path(path).read_bytes()
    tri = None
    if len(data) >= 84:
        n = struct.unpack("<I", data[80:84])[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            tri = rec["v"].astype(np.float64)
    if tri is None:
        text = data.decode("ascii", errors="replace").split()
        coords = []
        for i, w in enumerate(text):
            if w == "vertex":
                coords.append([float(text[i + 1]), float(text[i + 2]), float(text[i + 3])])
        if not coords or len(coords) % 3:
            raise ValueError(f"{path} is not an STL file")
        tri = np.asarray(coords, np.float64).reshape(-1, 3, 3)
    verts, inv = np.unique(tri.reshape(-1, 3), axis=0, return_inverse=True)
    return verts, inv.reshape(-1, 3).astype(np.int64)


def signed_volume(vertices, faces) -> float:
    v = np.asarray(vertices, np.float64)[np.asarray(faces)]
    return float(np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2])).sum() / 6.0)


def orient_outward(vertices, faces) -> np.ndarray:
    """Flip every face if the closed surface's signed volume is negative."""
    faces = np.asarray(faces)
    return faces[:, ::-1].copy() if signed_volume(vertices, faces) < 0 else faces


def mesh_arrays(meshes: Sequence[Mesh], all_materials: List[str], n_layers: int):
    """Flatten the enabled meshes into world-space arrays for drr_mesh_create:
    vertices float32 [n, 3], triangles int32 [m, 3], triangle primitive int32
    [m] and the ``MeshPrim`` table.  Primitives on layers >= n_layers are
    left out (the reference renders layers 0 .. mesh_layers - 1)."""
    verts, tris, tprim, prims = [], [], [], []
    base = 0
    for mesh in meshes:
        if not mesh.enabled:
            continue
        T = np.asarray(mesh.world_from_ijk.data, np.float64)
        for prim in mesh.mesh.primitives:
            mat = prim.material
            if mat.layer < 0 or mat.layer >= n_layers or len(prim.indices) == 0:
                continue
            pw = prim.positions @ T[:3, :3].T + T[:3, 3]
            if np.linalg.det(T[:3, :3]) < 0:  # a reflection turns the winding inside out
                idx = prim.indices[:, ::-1]
            else:
                idx = prim.indices
            verts.append(pw.astype(np.float32))
            tris.append((idx + base).astype(np.int32))
            tprim.append(np.full(len(idx), len(prims), np.int32))
            p = MeshPrim()
            p.material = all_materials.index(mat.drrMatName)
            p.density = float(mat.density)
            p.additive = 1 if mat.additive else 0
            p.subtractive = 1 if mat.subtractive else 0
            p.layer = int(mat.layer)
            prims.append(p)
            base += len(pw)
    V = np.concatenate(verts) if verts else np.zeros((0, 3), np.float32)
    F = np.concatenate(tris) if tris else np.zeros((0, 3), np.int32)
    P = np.concatenate(tprim) if tprim else np.zeros(0, np.int32)
    table = (MeshPrim * max(1, len(prims)))(*prims)
    return np.ascontiguousarray(V), np.ascontiguousarray(F), np.ascontiguousarray(P), table, len(prims)
