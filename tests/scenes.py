"""Seeded scenes shared by the parity tests, plus a thin runner that calls the
HIP library through its C ABI."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List

import numpy as np

from deepdrr_amd import phantoms, geo
from deepdrr_amd import projector as P
from deepdrr_amd._abi_types import Scene, View, make_view, view_to_dict
from deepdrr_amd.material import attenuation_table
from deepdrr_amd.spectral_data import spectrum_tables, synthetic_120kv
from deepdrr_amd.volume import Volume


@dataclass
class Case:
    name: str
    volumes: list            # [(density [nz,ny,nx] f32, labels u8)]
    view: View
    scene: Scene
    E: np.ndarray
    pdf: np.ndarray
    mu: np.ndarray
    materials: List[str] = field(default_factory=list)

    def scene_dict(self):
        return dict(step=self.scene.step, n_materials=self.scene.n_materials,
                    attenuate_outside=self.scene.attenuate_outside, air_index=self.scene.air_index)

    def view_dict(self):
        return view_to_dict(self.view)


def make_case(name, vols: List[Volume], proj: geo.CameraProjection, step=0.1, atten=False,
              priorities=None, max_ray_length=4080.0, spectrum=None) -> Case:
    mats = set()
    for v in vols:
        mats |= set(v.materials[0].keys())
    if atten:
        mats.add("air")
    mats = sorted(mats)
    E, pdf = spectrum_tables(synthetic_120kv() if spectrum is None else spectrum)
    mu = attenuation_table(mats, E)
    prios = priorities if priorities is not None else [len(vols) - 1 - i for i in range(len(vols))]
    view = P.build_view(vols, prios, proj, max_ray_length)
    arrays = [P.volume_arrays(v, mats) for v in vols]
    sc = Scene(step=step, n_materials=len(mats), attenuate_outside=1 if atten else 0,
               air_index=mats.index("air") if atten else 0)
    return Case(name, arrays, view, sc, E, pdf, mu, mats)


def small_case(n=32, det=(24, 20), step=0.1, atten=False, seed=0) -> Case:
    vol = phantoms.small_ct(n, seed)
    pix = 0.8 * 128 / n * 2
    dev = phantoms.ap_device(det[0], det[1], pix)
    return make_case(f"small{n}_{det[0]}x{det[1]}_atten{int(atten)}", [vol], dev.get_camera_projection(),
                     step=step, atten=atten)


def oblique_case(n=24, det=(16, 16), seed=1, step=0.1) -> Case:
    vol = phantoms.small_ct(n, seed)
    rng = np.random.default_rng(seed)
    (p, d, f), = phantoms.batch_poses(1, seed=seed, radius=0.2 * n)
    proj = phantoms.pose_projection(det[0], det[1], 0.8 * 128 / n * 2, p, d, f)
    return make_case(f"oblique{n}", [vol], proj, step=step)


def multi_case(step=0.1, same_grid=False, atten=False, priorities=None) -> Case:
    a = phantoms.small_ct(20, 0)
    if same_grid:
        # identical grid, different segmentation: exercises the shared label cache
        hu = phantoms.small_ct_hu(20, 3)
        b = Volume.from_hu(hu, origin=np.array(a.anatomical_from_IJK.t), spacing=(1, 1, 1),
                           anatomical_coordinate_system="LPS")
    else:
        b = phantoms.uniform_volume((10, 12, 8), spacing=(1.5, 1.0, 2.0), density=1.9, material="bone")
        b.place_center([3.0, 0.0, -2.0])
    dev = phantoms.ap_device(18, 14, 3.0)
    return make_case(f"multi_same{int(same_grid)}_atten{int(atten)}", [a, b], dev.get_camera_projection(),
                     step=step, atten=atten, priorities=priorities)


def many_case(n_vols, seed=0, atten=False, priorities=None) -> Case:
    """A CT and n_vols - 1 uniform blocks of other grids, spacings and materials,
    shifted so that they overlap it and each other (3 and 4 volumes run the 3-
    and 4-volume marches)."""
    rng = np.random.default_rng(seed)
    vols = [phantoms.small_ct(20, seed)]
    blocks = [((10, 12, 8), (1.5, 1.0, 2.0), 1.9, "bone"),
              ((14, 9, 11), (1.2, 1.7, 1.0), 0.3, "lung"),
              ((7, 16, 9), (2.0, 0.8, 1.6), 4.5, "titanium")]
    for shape, spacing, dens, mat in blocks[:n_vols - 1]:
        b = phantoms.uniform_volume(shape, spacing=spacing, density=dens, material=mat)
        b.place_center(rng.uniform(-6.0, 6.0, 3).tolist())
        vols.append(b)
    dev = phantoms.ap_device(18, 14, 3.0)
    return make_case(f"many{n_vols}_seed{seed}_atten{int(atten)}", vols, dev.get_camera_projection(),
                     atten=atten, priorities=priorities)


@dataclass
class MeshCase:
    """A scene of volumes (possibly none) and meshes."""
    case: Case
    mesh: tuple              # deepdrr_amd.mesh.mesh_arrays(...)
    n_layers: int
    max_hits: int
    source: np.ndarray       # world, float32
    sdd: float

    def oracle(self, rows=None):
        from oracle import ctrans

        c = self.case
        return ctrans.project_mesh(c.volumes, c.view, c.scene, c.E, c.pdf, c.mu, self.mesh, self.n_layers,
                                   self.max_hits, self.source, self.sdd, rows=rows)

    def oracle_rays(self):
        from oracle import ctrans

        return ctrans.mesh_rays(self.case.view, self.mesh, self.n_layers, self.max_hits,
                                self.case.scene.n_materials, self.source, self.sdd)


def make_mesh_case(name, vols: List[Volume], meshes, proj: geo.CameraProjection, sdd=1020.0, n_layers=2,
                   max_hits=32, step=0.1, atten=False, max_ray_length=4080.0) -> MeshCase:
    from deepdrr_amd.mesh import mesh_arrays

    mats = set()
    for v in vols:
        mats |= set(v.materials[0].keys())
    for m in meshes:
        for prim in m.mesh.primitives:
            mats.add(prim.material.drrMatName)
    if atten:
        mats.add("air")
    mats = sorted(mats)
    E, pdf = spectrum_tables(synthetic_120kv())
    mu = attenuation_table(mats, E)
    prios = [len(vols) - 1 - i for i in range(len(vols))]
    view = P.build_view(vols, prios, proj, max_ray_length, allow_empty=True)
    arrays = [P.volume_arrays(v, mats) for v in vols]
    sc = Scene(step=step, n_materials=len(mats), attenuate_outside=1 if atten else 0,
               air_index=mats.index("air") if atten else 0)
    case = Case(name, arrays, view, sc, E, pdf, mu, mats)
    return MeshCase(case, mesh_arrays(meshes, mats, n_layers), n_layers, max_hits,
                    np.asarray(proj.center_in_world, np.float32), float(sdd))


class GpuRunner:
    """Owns a drr_ctx and volume handles for one Case."""

    def __init__(self, case: Case):
        import torch
        from deepdrr_amd import _abi

        self.torch = torch
        self.abi = _abi
        self.case = case
        self.ctx = C.c_void_p()
        _abi.call("drr_ctx_create", 0, C.byref(self.ctx))
        _abi.call("drr_ctx_set_scene", self.ctx, C.byref(case.scene))
        _abi.call("drr_ctx_set_spectrum", self.ctx, len(case.E), case.E.ctypes.data,
                  case.pdf.ctypes.data, case.mu.ctypes.data)
        self.vols = []
        for d, l in case.volumes:
            h = C.c_void_p()
            nz, ny, nx = d.shape
            _abi.call("drr_volume_create", self.ctx, d.ctypes.data, l.ctypes.data, nx, ny, nz, 0, C.byref(h))
            self.vols.append(h)
        self.varr = (C.c_void_p * len(self.vols))(*[v.value for v in self.vols])

    def project(self, view: View = None, solid_angle=False):
        torch = self.torch
        view = view or self.case.view
        H, W, M = view.height, view.width, self.case.scene.n_materials
        dev = torch.device("cuda", 0)
        I = torch.full((H, W), float("nan"), device=dev)
        Pp = torch.full((H, W), float("nan"), device=dev)
        AD = torch.full((M, H, W), float("nan"), device=dev)
        SA = torch.full((H, W), float("nan"), device=dev) if solid_angle else None
        self.abi.call("drr_project", self.ctx, self.varr, C.byref(view), self.abi.ptr(I), self.abi.ptr(Pp),
                      self.abi.ptr(AD), self.abi.ptr(SA), self.abi.stream_ptr())
        torch.cuda.synchronize()
        out = (I.cpu().numpy(), Pp.cpu().numpy(), AD.cpu().numpy())
        return out + ((SA.cpu().numpy(),) if solid_angle else ())

    def project_mesh(self, mc: MeshCase, want_rays=False):
        """drr_project_meshes of the mesh case (this runner's volumes); with
        ``want_rays`` also drr_mesh_rays' intervals, counts and additive terms."""
        torch = self.torch
        view = mc.case.view
        H, W, M = view.height, view.width, self.case.scene.n_materials
        dev = torch.device("cuda", 0)
        verts, tris, tprim, prims, n_prims = mc.mesh
        h = C.c_void_p()
        self.abi.call("drr_mesh_create", self.ctx, verts.ctypes.data, len(verts), tris.ctypes.data,
                      tprim.ctypes.data, len(tris), prims, n_prims, mc.n_layers, mc.max_hits, C.byref(h))
        try:
            I = torch.full((H, W), float("nan"), device=dev)
            Pp = torch.full((H, W), float("nan"), device=dev)
            AD = torch.full((M, H, W), float("nan"), device=dev)
            src = np.ascontiguousarray(mc.source, np.float32)
            self.abi.call("drr_project_meshes", self.ctx, self.varr, C.byref(view), h, src.ctypes.data,
                          C.c_float(mc.sdd), self.abi.ptr(I), self.abi.ptr(Pp), self.abi.ptr(AD), None,
                          self.abi.stream_ptr())
            torch.cuda.synchronize()
            out = (I.cpu().numpy(), Pp.cpu().numpy(), AD.cpu().numpy())
            if want_rays:
                K = mc.n_layers * mc.max_hits // 2
                iv = torch.full((H, W, K, 2), float("nan"), device=dev)
                cnt = torch.full((H, W), -1, dtype=torch.int32, device=dev)
                add = torch.full((mc.n_layers * 16, H, W), float("nan"), device=dev)
                nu = C.c_int()
                self.abi.call("drr_mesh_rays", self.ctx, h, C.byref(view), src.ctypes.data, C.c_float(mc.sdd),
                              self.abi.ptr(iv), self.abi.ptr(cnt), self.abi.ptr(add), C.byref(nu),
                              self.abi.stream_ptr())
                torch.cuda.synchronize()
                out = out + (iv.cpu().numpy(), cnt.cpu().numpy(), add.cpu().numpy()[: mc.n_layers * nu.value],
                             nu.value)
            return out
        finally:
            self.abi.lib().drr_mesh_destroy(h)

    def close(self):
        for v in self.vols:
            self.abi.lib().drr_volume_destroy(v)
        self.abi.lib().drr_ctx_destroy(self.ctx)
