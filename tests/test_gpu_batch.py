"""Batched persistent forward (hipGraph) against the one-view path and the oracle.

Bar: each batched image's intensity is bit-identical to drr_project of the same
view (same march code), within 1e-5 relative of the transcription; the device
neglog matches the host NumPy neglog to 1e-5 absolute (values in [0, 1]; logf
differs by an ulp); graph replays give identical results.
"""
import ctypes as C

import numpy as np
import pytest

from deepdrr_amd import _abi, phantoms, utils
from deepdrr_amd import projector as P
from deepdrr_amd._abi_types import Scene
from deepdrr_amd.batch import BatchPlan
from deepdrr_amd.material import attenuation_table
from deepdrr_amd.spectral_data import spectrum_tables, synthetic_120kv

pytestmark = pytest.mark.gpu


def _setup(n_views=12, seed=0):
    vols = [phantoms.small_ct(24, 0), phantoms.small_ct(16, 1), phantoms.small_ct(20, 2)]
    vols[1].place_center([5.0, 3.0, -4.0])
    mats = sorted({m for v in vols for m in v.materials[0]})
    E, pdf = spectrum_tables(synthetic_120kv())
    mu = attenuation_table(mats, E)
    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    sc = Scene(step=0.1, n_materials=len(mats), attenuate_outside=0, air_index=0)
    _abi.call("drr_ctx_set_scene", ctx, C.byref(sc))
    _abi.call("drr_ctx_set_spectrum", ctx, len(E), E.ctypes.data, pdf.ctypes.data, mu.ctypes.data)
    handles, arrays = [], []
    for v in vols:
        d, l = P.volume_arrays(v, mats)
        h = C.c_void_p()
        nz, ny, nx = d.shape
        _abi.call("drr_volume_create", ctx, d.ctypes.data, l.ctypes.data, nx, ny, nz, 0, C.byref(h))
        handles.append(h)
        arrays.append((d, l))
    sizes = [(12, 10), (24, 20), (17, 9)]
    views, vidx = [], []
    for i, (p, d, f) in enumerate(phantoms.batch_poses(n_views, seed=seed, radius=6.0)):
        w, h = sizes[(i // 3) % 3]
        pix = 4.0 * 12 / w
        proj = phantoms.pose_projection(w, h, pix, p, d, f, center=np.array(vols[i % 3].center_in_world))
        views.append(P.build_view([vols[i % 3]], [0], proj, 4080.0))
        vidx.append(i % 3)
    return ctx, handles, arrays, views, vidx, sc, E, pdf, mu


def _teardown(ctx, handles):
    for h in handles:
        _abi.lib().drr_volume_destroy(h)
    _abi.lib().drr_ctx_destroy(ctx)


@pytest.mark.parametrize("tuning", [(8, 8, 4, 8), (16, 4, 1, 4), (8, 8, 2, 2)])
def test_batch_matches_single_view_and_oracle(gpu, tuning):
    import torch
    from oracle import ctrans

    ctx, handles, arrays, views, vidx, sc, E, pdf, mu = _setup()
    try:
        plan = BatchPlan(ctx, handles, views, vidx, neglog=True, tuning=tuning)
        print(f"batch tuning {tuning}: {len(views)} views", flush=True)
        plan.launch()
        torch.cuda.synchronize()
        I_b = plan.fetch("intensity")
        P_b = plan.fetch("photon_prob")
        img_b = plan.fetch("images")
        for i, (v, vi) in enumerate(zip(views, vidx)):
            I = torch.empty((v.height, v.width), device=gpu)
            Pp = torch.empty_like(I)
            varr = (C.c_void_p * 1)(handles[vi].value)
            _abi.call("drr_project", ctx, varr, C.byref(v), _abi.ptr(I), _abi.ptr(Pp), None, None,
                      _abi.stream_ptr())
            torch.cuda.synchronize()
            assert np.array_equal(I.cpu().numpy().view(np.uint32), I_b[i].view(np.uint32)), f"job {i}"
            assert np.array_equal(Pp.cpu().numpy().view(np.uint32), P_b[i].view(np.uint32)), f"job {i}"
            Io, Po, _ = ctrans.project([arrays[vi]], v, sc, E, pdf, mu)
            np.testing.assert_allclose(I_b[i], Io, rtol=1e-5)
            ref = utils.neglog(I_b[i][None].copy())[0]
            np.testing.assert_allclose(img_b[i], ref, atol=1e-5)
        plan.close()
    finally:
        _teardown(ctx, handles)


def test_graph_replays_are_identical(gpu):
    import torch

    ctx, handles, arrays, views, vidx, *_ = _setup(n_views=9, seed=3)
    try:
        plan = BatchPlan(ctx, handles, views, vidx, neglog=True)
        first = None
        for rep in range(4):
            plan.launch()
            torch.cuda.synchronize()
            cur = np.concatenate([a.ravel() for a in plan.fetch("images")])
            # eager work between replays
            junk = torch.randn(1 << 20, device=gpu).sum().item()
            if first is None:
                first = cur
            else:
                assert np.array_equal(first.view(np.uint32), cur.view(np.uint32)), f"replay {rep}"
        assert plan.march_ms() > 0
        plan.close()
    finally:
        _teardown(ctx, handles)


def test_fetch_async_keeps_each_launch(gpu):
    """fetch_async: each pending copy holds the images of the launch it follows,
    bit for bit those fetch returns, even when the device buffer is overwritten
    right after (as the next launch does) and several copies are in flight."""
    ctx, handles, arrays, views, vidx, *_ = _setup()
    try:
        plan = BatchPlan(ctx, handles, views, vidx, neglog=True, tuning=(8, 8, 4, 8))
        plan.launch()
        ref = plan.fetch("images")
        pend = []
        for i in range(3):
            plan.launch()
            pend.append(plan.fetch_async("images"))
            plan.images.fill_(float(i + 7))
            pend.append(plan.fetch_async("images"))
        for i in range(3):
            for x, y in zip(pend[2 * i].result(), ref):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"launch {i}"
            for x in pend[2 * i + 1].result():
                assert np.all(x == float(i + 7)), f"overwrite {i}"
        plan.close()
    finally:
        _teardown(ctx, handles)


def test_device_neglog_matches_host(gpu):
    import torch

    rng = np.random.default_rng(0)
    x = rng.uniform(1.0, 40.0, size=(3, 33, 47)).astype(np.float32)
    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    try:
        t = torch.from_numpy(x).to(gpu)
        out = torch.empty_like(t)
        _abi.call("drr_neglog", ctx, _abi.ptr(t), _abi.ptr(out), 3, 33 * 47, C.c_float(0.01), _abi.stream_ptr())
        np.testing.assert_allclose(out.cpu().numpy(), utils.neglog(x), atol=1e-5)
        # a constant image zeroes every image, as the reference does
        x[1] = 5.0
        t = torch.from_numpy(x).to(gpu)
        _abi.call("drr_neglog", ctx, _abi.ptr(t), _abi.ptr(out), 3, 33 * 47, C.c_float(0.01), _abi.stream_ptr())
        assert np.all(out.cpu().numpy() == 0)
        assert np.all(utils.neglog(x) == 0)
    finally:
        _abi.lib().drr_ctx_destroy(ctx)


def test_projector_batch_dropin(gpu):
    """Projector.project_batch returns what project(*views) returns."""
    from deepdrr_amd import Projector

    vol = phantoms.small_ct(24)
    dev = phantoms.ap_device(20, 16, 2.0)
    projs = [phantoms.pose_projection(20, 16, 2.0, p, d, f)
             for p, d, f in phantoms.batch_poses(5, seed=7, radius=5.0)]
    with Projector(vol, device=dev, spectrum=synthetic_120kv()) as proj:
        ref = proj(*projs)
        got = proj.project_batch(*projs)
    np.testing.assert_allclose(got, ref, atol=1e-5)


def test_autotuned_plan_gives_identical_images(gpu):
    """tuning='auto' picks one of the candidates and only changes the schedule."""
    import torch

    from deepdrr_amd import autotune

    ctx, handles, arrays, views, vidx, sc, E, pdf, mu = _setup(n_views=9, seed=2)
    try:
        key = autotune.workload_key(views, vidx, [a[0].shape[::-1] for a in arrays], sc.n_materials, "test")
        auto = BatchPlan(ctx, handles, views, vidx, neglog=False, tuning="auto", tune_key=key)
        print(f"autotuned schedule {auto.tuning}", flush=True)
        assert auto.tuning in autotune.CANDIDATES
        assert autotune.cached(key) == auto.tuning
        ref = BatchPlan(ctx, handles, views, vidx, neglog=False, tuning=(8, 8, 4, 8))
        auto.launch()
        ref.launch()
        torch.cuda.synchronize()
        for a, b in zip(auto.fetch("intensity"), ref.fetch("intensity")):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        auto.close()
        ref.close()
    finally:
        _teardown(ctx, handles)


def test_projector_batch_device_noise(gpu):
    """With add_noise, project_batch runs noise and neglog on the device: seeded
    calls repeat, the images stay in [0, 1] and differ from the noise-free ones,
    and the noisy intensity keeps the clean mean (zero-mean shot noise)."""
    from deepdrr_amd import Projector

    vol = phantoms.small_ct(24)
    dev = phantoms.ap_device(48, 40, 1.0)
    projs = [phantoms.pose_projection(48, 40, 1.0, p, d, f)
             for p, d, f in phantoms.batch_poses(4, seed=3, radius=3.0)]
    with Projector(vol, device=dev, spectrum=synthetic_120kv(), add_noise=True, photon_count=2000) as proj:
        proj.noise_seed = 11
        a = proj.project_batch(*projs)
        b = proj.project_batch(*projs)
        proj.neglog = False
        raw = proj.project_batch(*projs)
        proj.add_noise = False
        clean_raw = proj.project_batch(*projs)
        proj.neglog = True
        clean = proj.project_batch(*projs)
    assert a.shape == (4, 40, 48) and np.array_equal(a, b)
    assert np.all(a >= 0) and np.all(a <= 1) and not np.array_equal(a, clean)
    inner = (slice(None), slice(2, -2), slice(2, -2))
    rel = abs(raw[inner].mean() - clean_raw[inner].mean()) / clean_raw[inner].mean()
    assert rel < 0.01, rel


def test_multi_volume_jobs(gpu):
    """Views over 2 and 3 volumes (with priorities and air outside) mixed with
    single-volume views in one batch: each image bit-identical to drr_project
    of the same view, under both tile shapes; drr_project_jobs' area densities
    too."""
    import torch

    from deepdrr_amd._abi_types import Job

    ctx, handles, arrays, _, _, sc, E, pdf, mu = _setup(n_views=3)
    vols = [phantoms.small_ct(24, 0), phantoms.small_ct(16, 1), phantoms.small_ct(20, 2)]
    vols[1].place_center([5.0, 3.0, -4.0])
    try:
        views, vidx = [], []
        for i, (p, d, f) in enumerate(phantoms.batch_poses(9, seed=4, radius=6.0)):
            w, h = [(12, 10), (24, 20), (17, 9)][i % 3]
            proj = phantoms.pose_projection(w, h, 4.0 * 12 / w, p, d, f)
            first, n = [(0, 2), (1, 2), (0, 3), (2, 1), (0, 1)][i % 5]
            prios = [[1, 0], [0, 0], [2, 0, 1], [0], [0]][i % 5]
            views.append(P.build_view(vols[first:first + n], prios, proj, 4080.0))
            vidx.append(first)
        print(f"multi-volume batch: {[v.n_volumes for v in views]} volumes per view", flush=True)
        for tuning in ((8, 8, 4, 8), (16, 4, 2, 4)):
            plan = BatchPlan(ctx, handles, views, vidx, neglog=False, tuning=tuning)
            plan.launch()
            torch.cuda.synchronize()
            I_b = plan.fetch("intensity")
            plan.close()
            for i, (v, vi) in enumerate(zip(views, vidx)):
                I = torch.empty((v.height, v.width), device=gpu)
                Pp = torch.empty_like(I)
                varr = (C.c_void_p * v.n_volumes)(*[handles[vi + k].value for k in range(v.n_volumes)])
                _abi.call("drr_project", ctx, varr, C.byref(v), _abi.ptr(I), _abi.ptr(Pp), None, None,
                          _abi.stream_ptr())
                torch.cuda.synchronize()
                assert np.array_equal(I.cpu().numpy().view(np.uint32), I_b[i].view(np.uint32)), (tuning, i)
        # drr_project_jobs with area densities
        M = sc.n_materials
        jobs = (Job * len(views))()
        off = 0
        offs = []
        for i, (v, vi) in enumerate(zip(views, vidx)):
            jobs[i].view = v
            jobs[i].volume_index = vi
            jobs[i].out_offset = off
            offs.append(off)
            off += v.width * v.height
        I_j = torch.empty(off, device=gpu)
        AD_j = torch.empty(M * off, device=gpu)
        _abi.call("drr_project_jobs", ctx, (C.c_void_p * 3)(*[h.value for h in handles]), 3, jobs, len(views),
                  _abi.ptr(I_j), None, _abi.ptr(AD_j), _abi.stream_ptr())
        torch.cuda.synchronize()
        AD_all = AD_j.cpu().numpy()
        for i, (v, vi) in enumerate(zip(views, vidx)):
            I = torch.empty((v.height, v.width), device=gpu)
            Pp = torch.empty_like(I)
            AD = torch.empty((M, v.height, v.width), device=gpu)
            varr = (C.c_void_p * v.n_volumes)(*[handles[vi + k].value for k in range(v.n_volumes)])
            _abi.call("drr_project", ctx, varr, C.byref(v), _abi.ptr(I), _abi.ptr(Pp), _abi.ptr(AD), None,
                      _abi.stream_ptr())
            torch.cuda.synchronize()
            npx = v.width * v.height
            got = AD_all[M * offs[i]: M * offs[i] + M * npx].reshape(M, v.height, v.width)
            assert np.array_equal(AD.cpu().numpy().view(np.uint32), got.view(np.uint32)), i
        # a view whose volumes run past the batch's list is refused
        jobs[0].volume_index = 2
        rc = _abi.lib().drr_project_jobs(ctx, (C.c_void_p * 3)(*[h.value for h in handles]), 3, jobs, 1,
                                         _abi.ptr(I_j), None, None, _abi.stream_ptr())
        assert rc != 0 and "volume_index" in _abi.last_error()
    finally:
        _teardown(ctx, handles)


def test_projector_batch_two_volumes(gpu):
    """Projector.project_batch over a two-volume projector renders every view
    over both volumes, as project() does (the reference's multi-volume scene,
    tests/test_multivolume.py:19-32)."""
    from deepdrr_amd import Projector

    a = phantoms.small_ct(24, 0)
    b = phantoms.small_ct(16, 1)
    b.place_center([4.0, 2.0, -3.0])
    dev = phantoms.ap_device(20, 16, 2.0)
    projs = [phantoms.pose_projection(20, 16, 2.0, p, d, f)
             for p, d, f in phantoms.batch_poses(4, seed=9, radius=4.0)]
    with Projector([a, b], device=dev, spectrum=synthetic_120kv(), priorities=[1, 0]) as proj:
        ref = proj(*projs)
        got = proj.project_batch(*projs)
        single = proj.project_batch(*projs, volume_indices=[0, 1, 0, 1])
    np.testing.assert_allclose(got, ref, atol=1e-5)
    assert not np.allclose(single, ref, atol=1e-3)
