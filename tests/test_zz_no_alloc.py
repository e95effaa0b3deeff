"""drr_project_jobs_noalloc (job tables in the library's fixed device scratch)
against drr_project_jobs (job tables in fresh device allocations).

Bar: bit-identical intensity, photon_prob and area densities on two calls in a
row, the second with a smaller batch (the scratch keeps the first call's
tables), with the exact and the fast march, and a batch larger than the scratch
is refused on the host.
"""
import ctypes as C

import numpy as np
import pytest

from deepdrr_amd import _abi, phantoms
from deepdrr_amd import projector as P
from deepdrr_amd._abi_types import MARCH_EXACT, MARCH_FAST, Job, Scene
from deepdrr_amd.material import attenuation_table
from deepdrr_amd.spectral_data import spectrum_tables, synthetic_120kv

pytestmark = pytest.mark.gpu


def _scene(march):
    vols = [phantoms.small_ct(24, 0), phantoms.small_ct(16, 1), phantoms.small_ct(20, 2)]
    vols[1].place_center([5.0, 3.0, -4.0])
    mats = sorted({m for v in vols for m in v.materials[0]})
    E, pdf = spectrum_tables(synthetic_120kv())
    mu = attenuation_table(mats, E)
    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    sc = Scene(step=0.1, n_materials=len(mats), attenuate_outside=1, air_index=mats.index("air"), march=march)
    _abi.call("drr_ctx_set_scene", ctx, C.byref(sc))
    _abi.call("drr_ctx_set_spectrum", ctx, len(E), E.ctypes.data, pdf.ctypes.data, mu.ctypes.data)
    handles = []
    for v in vols:
        d, l = P.volume_arrays(v, mats)
        h = C.c_void_p()
        nz, ny, nx = d.shape
        _abi.call("drr_volume_create", ctx, d.ctypes.data, l.ctypes.data, nx, ny, nz, 0, C.byref(h))
        handles.append(h)
    return ctx, handles, vols, len(mats)


def _jobs(vols, n, seed, sizes, multi=False):
    views, vidx = [], []
    for i, (p, d, f) in enumerate(phantoms.batch_poses(n, seed=seed, radius=6.0)):
        w, h = sizes[i % len(sizes)]
        proj = phantoms.pose_projection(w, h, 48.0 / w, p, d, f, center=np.array(vols[i % 3].center_in_world))
        if multi and i % 2:
            # a view over volumes 1 and 2 (the multi-volume job table)
            views.append(P.build_view(vols[1:3], [1, 0], proj, 4080.0))
            vidx.append(1)
            continue
        views.append(P.build_view([vols[i % 3]], [0], proj, 4080.0))
        vidx.append(i % 3)
    jobs = (Job * n)()
    off = 0
    for i, (v, vi) in enumerate(zip(views, vidx)):
        jobs[i].view = v
        jobs[i].volume_index = vi
        jobs[i].out_offset = off
        off += v.width * v.height
    return jobs, off


def _run(name, ctx, varr, n_vols, jobs, n, total, M):
    import torch

    dev = torch.device("cuda", 0)
    I = torch.full((total,), float("nan"), device=dev)
    Pp = torch.full((total,), float("nan"), device=dev)
    AD = torch.full((M * total,), float("nan"), device=dev)
    _abi.call(name, ctx, varr, n_vols, jobs, n, _abi.ptr(I), _abi.ptr(Pp), _abi.ptr(AD), _abi.stream_ptr())
    torch.cuda.synchronize()
    return [t.cpu().numpy().view(np.uint32) for t in (I, Pp, AD)]


@pytest.mark.parametrize("march", [MARCH_EXACT, MARCH_FAST], ids=["exact", "fast"])
def test_noalloc_matches_allocating_entry_point(gpu, march):
    ctx, handles, vols, M = _scene(march)
    varr = (C.c_void_p * 3)(*[h.value for h in handles])
    try:
        # a larger batch, then a smaller one over the scratch the first left behind
        for n, seed, sizes, multi in [(12, 3, [(40, 33), (24, 20), (17, 9)], True), (4, 8, [(9, 7), (16, 13)], False)]:
            jobs, total = _jobs(vols, n, seed, sizes, multi)
            print(f"noalloc: march {march}, {n} jobs, {total} pixels", flush=True)
            ref = _run("drr_project_jobs", ctx, varr, 3, jobs, n, total, M)
            got = _run("drr_project_jobs_noalloc", ctx, varr, 3, jobs, n, total, M)
            for a, b, what in zip(ref, got, ("intensity", "photon_prob", "area_density")):
                assert np.array_equal(a, b), f"{n} jobs: {what} differs"
            assert not np.isnan(got[0].view(np.float32)).any()
        # a batch whose ray tiles exceed the scratch is refused before any launch
        big, _ = _jobs(vols, 1, 5, [(65536, 65536)])
        fn = _abi.lib().drr_project_jobs_noalloc
        import torch

        dummy = torch.zeros(16, device=gpu)
        rc = fn(ctx, varr, 3, big, 1, _abi.ptr(dummy), None, None, _abi.stream_ptr())
        assert rc == -1 and "scratch" in _abi.last_error()
    finally:
        for h in handles:
            _abi.lib().drr_volume_destroy(h)
        _abi.lib().drr_ctx_destroy(ctx)
