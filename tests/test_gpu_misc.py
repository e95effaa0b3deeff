"""Device noise, GPU HU ingestion and volume round trips."""
import ctypes as C

import numpy as np
import pytest
import torch

from deepdrr_amd import _abi, phantoms, utils, workload
from deepdrr_amd import projector as P
from deepdrr_amd._abi_types import Scene
from deepdrr_amd.volume import FIVE_MATERIAL_CLASSES, Volume

pytestmark = pytest.mark.gpu

K = np.array([[0.03, 0.06, 0.02], [0.11, 0.98, 0.11], [0.02, 0.06, 0.03]])


def _ctx(n_materials=5):
    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    sc = Scene(step=0.1, n_materials=n_materials, attenuate_outside=0, air_index=0)
    _abi.call("drr_ctx_set_scene", ctx, C.byref(sc))
    return ctx


@pytest.mark.parametrize("lam", [5.0, 5000.0])
def test_device_noise_statistics(gpu, lam):
    """Interior pixels of a constant image: mean I, variance I^2/lambda * sum(K^2)
    (Poisson shot noise, zero-mean, correlated by the 3x3 kernel)."""
    ctx = _ctx()
    try:
        H = W = 256
        I0, count = 30.0, 10000.0
        I = torch.full((2, H, W), I0, device=gpu)
        Pp = torch.full((2, H, W), lam / count, device=gpu)
        out = torch.empty_like(I)
        _abi.call("drr_add_noise", ctx, _abi.ptr(I), _abi.ptr(Pp), _abi.ptr(out), 2, H, W,
                  C.c_float(count), C.c_uint64(7), _abi.stream_ptr())
        x = out.cpu().numpy()[:, 1:-1, 1:-1].astype(np.float64)
        var_expect = I0 ** 2 / lam * float((K ** 2).sum())
        assert abs(x.mean() - I0) < 5 * np.sqrt(var_expect / x.size) * 3  # correlated samples
        assert x.var() == pytest.approx(var_expect, rel=0.05)
        # the host path (reference algorithm, NumPy generator) has the same statistics
        h = utils.add_noise(np.full((1, H, W), I0), np.full((1, H, W), lam / count), count,
                            rng=np.random.default_rng(1))[0, 1:-1, 1:-1]
        assert h.var() == pytest.approx(var_expect, rel=0.05)
        # same seed -> same noise; different seed -> different noise
        out2 = torch.empty_like(I)
        _abi.call("drr_add_noise", ctx, _abi.ptr(I), _abi.ptr(Pp), _abi.ptr(out2), 2, H, W,
                  C.c_float(count), C.c_uint64(7), _abi.stream_ptr())
        assert torch.equal(out, out2)
    finally:
        _abi.lib().drr_ctx_destroy(ctx)


def test_gpu_hu_ingestion_matches_host(gpu):
    """drr_volume_create_from_hu gives exactly Volume.from_hu's float32 density and labels."""
    hu = phantoms.clinical_ct_hu((40, 36, 28), seed=4)
    vol = Volume.from_hu(hu, origin=[0, 0, 0], anatomical_coordinate_system="LPS",
                         material_classes=FIVE_MATERIAL_CLASSES)
    mats = sorted(workload.CLINICAL_CLASSES)
    dens, labs = P.volume_arrays(vol, mats)
    ctx = _ctx(len(mats))
    try:
        hu_kji = np.ascontiguousarray(np.moveaxis(hu, [0, 1, 2], [2, 1, 0]))
        upper = np.array(workload.CLINICAL_UPPER_HU, np.float32)
        lab = workload.clinical_label_map(mats)
        h = C.c_void_p()
        nz, ny, nx = hu_kji.shape
        _abi.call("drr_volume_create_from_hu", ctx, hu_kji.ctypes.data, nx, ny, nz, 0, len(upper),
                  upper.ctypes.data, lab.ctypes.data, C.byref(h))
        d2 = np.empty_like(dens)
        l2 = np.empty_like(labs)
        _abi.call("drr_volume_download", h, d2.ctypes.data, l2.ctypes.data)
        assert np.array_equal(d2.view(np.uint32), dens.view(np.uint32))
        assert np.array_equal(l2, labs)
        _abi.lib().drr_volume_destroy(h)
    finally:
        _abi.lib().drr_ctx_destroy(ctx)


def test_volume_errors(gpu):
    ctx = _ctx(3)
    try:
        d = np.ones((4, 4, 4), np.float32)
        l = np.full((4, 4, 4), 3, np.uint8)  # label >= n_materials
        h = C.c_void_p()
        rc = _abi.lib().drr_volume_create(ctx, d.ctypes.data, l.ctypes.data, 4, 4, 4, 0, C.byref(h))
        assert rc != 0 and "n_materials" in _abi.last_error()
        rc = _abi.lib().drr_volume_create(ctx, d.ctypes.data, l.ctypes.data, 0, 4, 4, 0, C.byref(h))
        assert rc != 0
        # device-pointer labels are checked on the device
        dl = torch.from_numpy(l).to(gpu)
        dd = torch.from_numpy(d).to(gpu)
        rc = _abi.lib().drr_volume_create(ctx, _abi.ptr(dd), _abi.ptr(dl), 4, 4, 4, 1, C.byref(h))
        assert rc != 0
    finally:
        _abi.lib().drr_ctx_destroy(ctx)


def test_non_finite_view_rejected(gpu):
    """A view with a NaN or infinite geometry value is refused on the host."""
    from tests import scenes

    case = scenes.small_case(8, (6, 5))
    r = scenes.GpuRunner(case)
    try:
        for field, val in [("world_from_index", float("nan")), ("max_ray_length", float("inf"))]:
            v = type(case.view).from_buffer_copy(case.view)
            if field == "world_from_index":
                v.world_from_index[4] = val
            else:
                v.max_ray_length = val
            I = torch.empty((v.height, v.width), device=gpu)
            rc = _abi.lib().drr_project(r.ctx, r.varr, C.byref(v), _abi.ptr(I), _abi.ptr(I), None, None,
                                        _abi.stream_ptr())
            assert rc != 0 and "finite" in _abi.last_error()
    finally:
        r.close()


def test_job_view_and_plan_errors(gpu):
    """Invalid jobs, views and batch plans are refused on the host, before any
    launch, with a message (the reference's Projector raises for its own
    invalid inputs; the C ABI returns a status and drr_last_error)."""
    from deepdrr_amd._abi_types import Job
    from deepdrr_amd.batch import BatchPlan
    from tests import scenes

    case = scenes.small_case(8, (6, 5))
    r = scenes.GpuRunner(case)
    try:
        out = torch.empty(256, device=gpu)

        def project_jobs(jobs, n):
            return _abi.lib().drr_project_jobs(r.ctx, r.varr, 1, jobs, n, _abi.ptr(out), _abi.ptr(out), None,
                                               _abi.stream_ptr())

        J = (Job * 1)()
        J[0].view = case.view
        J[0].volume_index = 0
        J[0].out_offset = 0
        assert project_jobs(J, 0) != 0 and "jobs" in _abi.last_error()
        J[0].volume_index = 1
        assert project_jobs(J, 1) != 0 and "volume_index" in _abi.last_error()
        J[0].volume_index = 0
        J[0].out_offset = -1
        assert project_jobs(J, 1) != 0 and "offset" in _abi.last_error()
        J[0].out_offset = 0
        J[0].view.n_volumes = 2  # two volumes from a list of one
        assert project_jobs(J, 1) != 0 and "volume_index" in _abi.last_error()
        J[0].view.n_volumes = 5
        assert project_jobs(J, 1) != 0 and "n_volumes" in _abi.last_error()
        J[0].view.n_volumes = 1
        for field, val in (("n_volumes", 0), ("n_volumes", 5), ("width", 0), ("height", 65537)):
            v = type(case.view).from_buffer_copy(case.view)
            setattr(v, field, val)
            rc = _abi.lib().drr_project(r.ctx, r.varr, C.byref(v), _abi.ptr(out), _abi.ptr(out), None, None,
                                        _abi.stream_ptr())
            assert rc != 0, (field, val)
        with pytest.raises(ValueError):
            BatchPlan(r.ctx, [r.vols[0]], [], [])
        with pytest.raises(RuntimeError, match="tile"):
            BatchPlan(r.ctx, [r.vols[0]], [case.view], [0], neglog=False, tuning=(12, 8, 4, 8))
        with pytest.raises(RuntimeError, match="waves_per_block"):
            BatchPlan(r.ctx, [r.vols[0]], [case.view], [0], neglog=False, tuning=(8, 8, 5, 8))
        # the context still works after the refusals
        g = r.project()
        assert np.all(np.isfinite(g[0]))
    finally:
        r.close()


def test_projector_ingests_int16_hu_on_gpu(gpu):
    """A Volume made from int16 HU keeps the HU; Projector.initialize derives
    density and labels on the GPU (drr_volume_create_from_hu) without ever
    computing them on the host, and renders bit-identically to the same volume
    whose density and labels were derived on the host."""
    from deepdrr_amd import Projector, phantoms
    from deepdrr_amd.spectral_data import synthetic_120kv
    from deepdrr_amd.volume import Volume

    hu = phantoms.clinical_ct_hu((48, 44, 40), seed=3)
    kw = dict(origin=phantoms._centered_origin(hu.shape, (1.0, 1.0, 1.2)), spacing=(1.0, 1.0, 1.2),
              anatomical_coordinate_system="LPS", material_classes=phantoms.FIVE_MATERIAL_CLASSES)
    lazy = Volume.from_hu(hu, **kw)
    host = Volume.from_hu(hu, **kw)
    _ = host.data  # derived on the host
    assert lazy.hu_source() is not None and host.hu_source() is None
    dev = phantoms.ap_device(40, 36, 1.5)
    outs = []
    for vol in (lazy, host):
        with Projector(vol, device=dev, spectrum=synthetic_120kv(), neglog=False) as p:
            outs.append(p())
    assert lazy._data is None, "the lazy volume was derived on the host"
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    # default thresholds (air/soft tissue/bone) as well
    lazy3 = Volume.from_hu(hu, origin=kw["origin"], spacing=kw["spacing"], anatomical_coordinate_system="LPS")
    host3 = Volume.from_hu(hu, origin=kw["origin"], spacing=kw["spacing"], anatomical_coordinate_system="LPS")
    _ = host3.materials  # derived on the host
    assert lazy3.hu_source() is not None and host3.hu_source() is None
    o3 = []
    for vol in (lazy3, host3):
        with Projector(vol, device=dev, spectrum=synthetic_120kv(), neglog=False) as p:
            o3.append(p())
    assert np.array_equal(o3[0].view(np.uint32), o3[1].view(np.uint32))
