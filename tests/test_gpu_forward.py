"""GPU parity of the forward projector against the CPU oracle.

Bar: per-material area densities bit-identical to the FP32 transcription
(oracle/transcription.c, same operation order); intensity and photon_prob
within 1e-5 relative of it (expf differs by an ulp between ROCm and glibc);
bit-identical across runs.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from tests import scenes
from tests.scenes import GpuRunner

pytestmark = pytest.mark.gpu

REL = 1e-5


def _oracle(case):
    from oracle import ctrans

    return ctrans.project(case.volumes, case.view, case.scene, case.E, case.pdf, case.mu)


def _check(case, gpu_out, ora_out):
    I, P, AD = gpu_out[:3]
    Io, Po, ADo = ora_out
    assert np.all(np.isfinite(I)) and np.all(np.isfinite(AD))
    bad = np.argwhere(AD.view(np.uint32) != ADo.view(np.uint32))
    assert bad.size == 0, f"{case.name}: {len(bad)} area densities differ, first {bad[:3]} " \
                          f"gpu {AD[tuple(bad[0])]} ora {ADo[tuple(bad[0])]}"
    np.testing.assert_allclose(I, Io, rtol=REL, atol=0)
    np.testing.assert_allclose(P, Po, rtol=REL, atol=0)


CASES = [
    lambda: scenes.small_case(32, (24, 20)),
    lambda: scenes.small_case(32, (24, 20), atten=True),
    lambda: scenes.small_case(16, (9, 13), step=0.37),
    lambda: scenes.oblique_case(24, (16, 16), seed=1),
    lambda: scenes.oblique_case(24, (17, 15), seed=5, step=0.25),
    lambda: scenes.multi_case(),
    lambda: scenes.multi_case(same_grid=True),
    lambda: scenes.multi_case(atten=True, priorities=[0, 0]),
    lambda: scenes.many_case(3, seed=0),
    lambda: scenes.many_case(4, seed=1, priorities=[1, 0, 1, 0]),
    lambda: scenes.many_case(4, seed=2, atten=True, priorities=[0, 0, 0, 0]),
]


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_forward_matches_transcription(gpu, idx):
    case = CASES[idx]()
    print(f"case {case.name}", flush=True)
    r = GpuRunner(case)
    try:
        g = r.project()
        _check(case, g, _oracle(case))
        g2 = r.project()
        for a, b in zip(g, g2):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), "not bit-stable across runs"
    finally:
        r.close()


def test_config1_full_size(gpu):
    """128^3 phantom -> 256x256 detector at 0.8 mm, step 0.1, 20-bin 120 kV."""
    from deepdrr_amd import phantoms

    vol = phantoms.small_ct(128)
    dev = phantoms.ap_device(256, 256, 0.8)
    case = scenes.make_case("cfg1", [vol], dev.get_camera_projection())
    print("case cfg1", flush=True)
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
    finally:
        r.close()


def test_fp64_error_reported(gpu):
    """GPU vs the FP64 restatement: the FP32 arithmetic error, stated as a bound."""
    from oracle import fp64

    case = scenes.small_case(32, (24, 20))
    r = GpuRunner(case)
    try:
        I, P, AD = r.project()
    finally:
        r.close()
    I64, P64, AD64 = fp64.project_fp64(case.volumes, case.view_dict(), case.scene_dict(), case.E,
                                       case.pdf, case.mu)
    err = float(np.max(np.abs(I - I64) / np.abs(I64)))
    print(f"max relative error of FP32 intensity vs FP64: {err:.3g}")
    assert err < 1e-5


def test_solid_angle(gpu):
    from oracle import ctrans

    case = scenes.small_case(16, (12, 10))
    r = GpuRunner(case)
    try:
        out = r.project(solid_angle=True)
    finally:
        r.close()
    sa = out[3]
    ref = ctrans.solid_angle(case.view.world_from_index[:], case.view.width, case.view.height)
    np.testing.assert_allclose(sa, ref, rtol=1e-5)


def test_rays_missing_volume(gpu):
    """A view pointing away: no samples, photon_prob = sum(pdf), intensity = sum(E pdf)."""
    from deepdrr_amd import phantoms

    vol = phantoms.small_ct(16)
    dev = phantoms.ap_device(8, 8, 1.0)
    dev.set_view(point=[0, -1200, 0], direction=[0, -1, 0], up=[0, 0, 1])
    case = scenes.make_case("miss", [vol], dev.get_camera_projection())
    r = GpuRunner(case)
    try:
        I, P, AD = r.project()
    finally:
        r.close()
    assert np.all(AD == 0)
    np.testing.assert_allclose(P, np.float32(case.pdf.astype(np.float64).sum()), rtol=1e-6)
    np.testing.assert_allclose(I, np.float32((case.E.astype(np.float64) * case.pdf).sum()), rtol=1e-6)


def test_disabled_volume(gpu):
    case = scenes.multi_case()
    case.view.vol[1].enabled = 0
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
    finally:
        r.close()


def test_projector_dropin(gpu):
    """Projector(volume, device).initialize()/__call__ returns the neglog image the
    reference pipeline computes from the oracle's intensities, with both
    marches (exact: area densities bit-identical; fast: within 1e-5)."""
    from deepdrr_amd import Projector, phantoms, utils
    from deepdrr_amd.spectral_data import synthetic_120kv

    vol = phantoms.small_ct(32)
    dev = phantoms.ap_device(40, 30, 1.6)
    out = {}
    for march in ("exact", "fast"):
        with Projector(vol, device=dev, step=0.1, spectrum=synthetic_120kv(), march=march) as proj:
            img = proj()
            assert img.shape == (30, 40) and img.dtype in (np.float32, np.float64)
            out[march] = (img,) + proj.project_area_density(dev.get_camera_projection())
            case = scenes.make_case("dropin", [vol], dev.get_camera_projection(),
                                    max_ray_length=proj.max_ray_length)
    Io, Po, ADo = _oracle(case)
    img, AD, I, P = out["exact"]
    np.testing.assert_allclose(utils.neglog(Io[None])[0], img, atol=1e-5)
    assert np.array_equal(AD.view(np.uint32), ADo.view(np.uint32))
    img, AD, I, P = out["fast"]
    np.testing.assert_allclose(utils.neglog(Io[None])[0], img, atol=1e-5)
    np.testing.assert_allclose(I, Io, rtol=1e-5)
    assert np.all(np.abs(AD - ADo) <= 1e-5 * ADo.sum(0))


def test_projector_multiview_and_errors(gpu):
    from deepdrr_amd import Projector, phantoms

    vol = phantoms.small_ct(16)
    dev = phantoms.ap_device(8, 6, 4.0)
    p = Projector(vol, device=dev)
    with pytest.raises(RuntimeError):
        p()
    p.initialize()
    with pytest.raises(RuntimeError):
        p.initialize()
    projs = [dev.get_camera_projection(), dev.get_camera_projection()]
    imgs = p(*projs)
    assert imgs.shape == (2, 6, 8)
    np.testing.assert_array_equal(imgs[0], imgs[1])
    p.free()


def test_wide_slices(gpu):
    """2900 x 2900 x 4 voxels: (nx+1)(ny+1) > 2^23, so the kernels form cell
    indices the second way ((cz+1)(ny+1) + (cy+1) first); an oblique view
    crosses cells along all three axes, and labels change in a checkerboard."""
    from deepdrr_amd import phantoms
    from deepdrr_amd.volume import Volume

    nx, ny, nz = 2900, 2900, 4
    assert (nx + 1) * (ny + 1) > 1 << 23
    rng = np.random.default_rng(11)
    dens = (1.0 + 0.5 * rng.random((nx, ny, nz), dtype=np.float32)).astype(np.float32)
    i = np.arange(nx)[:, None, None] // 37
    j = np.arange(ny)[None, :, None] // 53
    bone = ((i + j) % 2 == 0) & np.ones((1, 1, nz), bool)
    spacing = (0.05, 0.05, 2.0)
    vol = Volume.from_parameters(dens, {"bone": bone, "soft tissue": ~bone},
                                 origin=phantoms._centered_origin((nx, ny, nz), spacing), spacing=spacing,
                                 anatomical_coordinate_system="LPS")
    d = np.array([0.2, 1.0, 0.05])
    proj = phantoms.pose_projection(16, 12, 0.4, [0.0, 0.0, 0.0], d / np.linalg.norm(d), 0.5)
    case = scenes.make_case("wide_slices", [vol], proj)
    print("case wide_slices", flush=True)
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
    finally:
        r.close()


@pytest.mark.parametrize("sdd", [2.0e4, 4.0e6])
def test_far_source(gpu, sdd):
    """A source far from the volume: alpha is large, so each FP32 `alpha += step`
    rounds noticeably (at 2e6 mm by up to a quarter step).  The kernels bound the
    samples they march without the maxAlpha test by that rounding and march the
    rest one at a time with the full test; the result must still equal the
    transcription's sequential sum bit for bit."""
    from deepdrr_amd import phantoms

    vol = phantoms.small_ct(32, seed=3)
    dev = phantoms.ap_device(16, 12, 4.0, sdd=sdd)
    case = scenes.make_case(f"far_source_{sdd:g}", [vol], dev.get_camera_projection(),
                            max_ray_length=4.0 * sdd)
    print(f"case {case.name}", flush=True)
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
    finally:
        r.close()


def test_source_inside_volume(gpu):
    """The source inside the volume: the slab clip starts the march at alpha 0,
    so the FP32 alpha increments change from binade to binade; bit-exact area
    densities against the transcription."""
    from deepdrr_amd import phantoms

    vol = phantoms.small_ct(32, seed=3)
    d = np.array([0.3, 1.0, 0.2])
    proj = phantoms.pose_projection(16, 12, 2.0, [1.0, -2.0, 0.5], d / np.linalg.norm(d), 0.5, sdd=20.0)
    case = scenes.make_case("inside_source", [vol], proj)
    print(f"case {case.name}", flush=True)
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
    finally:
        r.close()


def _random_scene(seed):
    """A small anisotropic volume with three labelled materials (boundaries
    everywhere), a random pose and detector, a random step, with or without
    attenuation outside the volume."""
    from deepdrr_amd import phantoms
    from deepdrr_amd.volume import Volume

    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(3, 28, 3))
    spacing = tuple(float(x) for x in rng.uniform(0.4, 2.5, 3))
    dens = rng.uniform(0.0, 2.0, shape).astype(np.float32)
    u = rng.random(shape)
    bone = u < 0.3
    air = (u >= 0.3) & (u < 0.5)
    vol = Volume.from_parameters(dens, {"bone": bone, "air": air, "soft tissue": ~(bone | air)},
                                 origin=phantoms._centered_origin(shape, spacing), spacing=spacing,
                                 anatomical_coordinate_system="LPS")
    (p, d, f), = phantoms.batch_poses(1, seed=seed, radius=0.3 * max(shape) * max(spacing), cone_deg=80.0)
    w, h = (int(x) for x in rng.integers(4, 20, 2))
    proj = phantoms.pose_projection(w, h, float(rng.uniform(0.5, 4.0)), p, d, f)
    step = float(rng.choice([0.05, 0.1, 0.37, 1.3]))
    return scenes.make_case(f"random{seed}", [vol], proj, step=step, atten=bool(rng.random() < 0.4))


@pytest.mark.parametrize("seed", range(40))
def test_random_scenes(gpu, seed):
    """Bit-exact area densities on random anisotropic volumes, poses, detectors
    and steps (the cell loads, the label identity, the FP32 cell index, the
    loop bound and the tail, with and without air outside the volume)."""
    case = _random_scene(seed)
    print(f"case {case.name}", flush=True)
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
    finally:
        r.close()


def test_clinical_size_band(gpu):
    """The clinical 512x512x400 CT (0.8x0.8x1 mm, 5 materials: 105M cells,
    rays of ~4000 samples) through bands of 1536-pixel detector rows at the
    bench's 0.194 mm pitch and poses: bit-exact area densities against the
    transcription, and batched images bit-identical to one view per launch."""
    import torch

    from deepdrr_amd import _abi, phantoms, workload
    from deepdrr_amd.batch import BatchPlan

    vol = phantoms.clinical_ct()
    poses = phantoms.batch_poses(2, seed=0)
    projs = [phantoms.pose_projection(1536, 4, 0.194, p, d, f) for p, d, f in poses]
    case = scenes.make_case("clinical_band", [vol], projs[0])
    print("case clinical_band", flush=True)
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
        views = [scenes.P.build_view([vol], [0], pr, 4080.0) for pr in projs]
        singles = [r.project(v)[:2] for v in views]
        plan = BatchPlan(r.ctx, [r.vols[0]], views, [0, 0], neglog=False, tuning=(8, 8, 4, 8))
        plan.launch()
        torch.cuda.synchronize()
        I_b, P_b = plan.fetch("intensity"), plan.fetch("photon_prob")
        plan.close()
        for i, (I, P) in enumerate(singles):
            assert np.array_equal(I.view(np.uint32), I_b[i].view(np.uint32)), f"view {i}"
            assert np.array_equal(P.view(np.uint32), P_b[i].view(np.uint32)), f"view {i}"
    finally:
        r.close()


MATERIAL_NAMES = ["air", "bone", "soft tissue", "lung", "titanium", "muscle", "blood", "teflon", "iron", "lead",
                  "copper", "polyethylene", "concrete", "Al", "Ca", "W"]


def _materials_scene(seed, k):
    """k materials (1 to 16) over a small anisotropic volume: 3x3x3 blocks of one
    label (uniform cells, material switches) with 10% of the voxels relabelled at
    random (mixed cells), two random poses, a random detector and step."""
    from deepdrr_amd import phantoms
    from deepdrr_amd.volume import Volume

    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(4, 20, 3))
    spacing = tuple(float(x) for x in rng.uniform(0.5, 2.0, 3))
    dens = rng.uniform(0.0, 3.0, shape).astype(np.float32)
    coarse = rng.integers(0, k, tuple((s + 2) // 3 for s in shape))
    lab = np.repeat(np.repeat(np.repeat(coarse, 3, 0), 3, 1), 3, 2)[:shape[0], :shape[1], :shape[2]]
    lab = np.where(rng.random(shape) < 0.1, rng.integers(0, k, shape), lab)
    names = MATERIAL_NAMES[:k]
    vol = Volume.from_parameters(dens, {m: lab == i for i, m in enumerate(names)},
                                 origin=phantoms._centered_origin(shape, spacing), spacing=spacing,
                                 anatomical_coordinate_system="LPS")
    poses = phantoms.batch_poses(2, seed=seed, radius=0.3 * max(shape) * max(spacing), cone_deg=80.0)
    w, h = (int(x) for x in rng.integers(4, 20, 2))
    pix = float(rng.uniform(0.5, 4.0))
    projs = [phantoms.pose_projection(w, h, pix, p, d, f) for p, d, f in poses]
    step = float(rng.choice([0.05, 0.1, 0.37]))
    case = scenes.make_case(f"materials{k}", [vol], projs[0], step=step, atten=(k % 3 == 0))
    assert case.scene.n_materials == k
    return case, vol, projs


@pytest.mark.parametrize("k", list(range(1, 9)) + [11, 16])
def test_material_counts(gpu, k):
    """Material counts the marches are instantiated for (1 to 16, as 4-bit
    labels; from 6 on the marches run at a lower occupancy): bit-exact area densities against the
    transcription, and batched images bit-identical to one view per launch with
    both tile shapes."""
    import torch

    from deepdrr_amd.batch import BatchPlan

    case, vol, projs = _materials_scene(40 + k, k)
    print(f"case {case.name}", flush=True)
    r = GpuRunner(case)
    try:
        _check(case, r.project(), _oracle(case))
        views = [case.view] + [scenes.P.build_view([vol], [0], pr, 4080.0) for pr in projs[1:]]
        singles = [r.project(v)[:2] for v in views]
        for tuning in ((8, 8, 4, 8), (16, 4, 2, 4)):
            plan = BatchPlan(r.ctx, [r.vols[0]], views, [0] * len(views), neglog=False, tuning=tuning)
            plan.launch()
            torch.cuda.synchronize()
            I_b, P_b = plan.fetch("intensity"), plan.fetch("photon_prob")
            plan.close()
            for i, (I, P) in enumerate(singles):
                assert np.array_equal(I.view(np.uint32), I_b[i].view(np.uint32)), f"view {i}, {tuning}"
                assert np.array_equal(P.view(np.uint32), P_b[i].view(np.uint32)), f"view {i}, {tuning}"
    finally:
        r.close()


def test_projector_collected_energy_and_upper_bound(gpu):
    """Projector(collected_energy=True): intensity x solid angle x photon_count /
    mean(solid angle) / (pixel_x pixel_y) with pixel = SDD / focal length
    (projector.py:833-853), the solid angles from the GPU checked against the
    oracle's; and intensity_upper_bound clips the images as project() does."""
    from oracle import ctrans

    from deepdrr_amd import Projector, phantoms
    from deepdrr_amd.spectral_data import synthetic_120kv

    vol = phantoms.small_ct(24)
    dev = phantoms.ap_device(20, 16, 2.0)
    cam = dev.get_camera_projection()
    kw = dict(device=dev, step=0.1, spectrum=synthetic_120kv(), neglog=False)
    with Projector(vol, **kw) as proj:
        I = proj(cam)
        view = proj.view_for(cam)
    with Projector(vol, collected_energy=True, photon_count=5000, **kw) as proj:
        E = proj(cam)
        sdd = proj.source_to_detector_distance
    assert sdd == 1020.0
    sa = ctrans.solid_angle(view.world_from_index[:], view.width, view.height)
    px = sdd / cam.index_from_camera2d.fx
    py = sdd / cam.index_from_camera2d.fy
    np.testing.assert_allclose(E, I.astype(np.float64) * sa * 5000 / np.average(sa) / (px * py), rtol=1e-5)
    ub = float(np.percentile(I, 60))
    with Projector(vol, intensity_upper_bound=ub, **kw) as proj:
        C = proj(cam)
    assert np.array_equal(C, np.clip(I, None, ub))
    assert (C < I).any()


@pytest.mark.parametrize("direction", [(1.0, 0.02, 0.01), (0.05, 1.0, 0.02)])
def test_longest_dimension(gpu, direction):
    """A 16384 x 3 x 2 rod (the largest dimension a volume may have), labels
    changing along it, viewed along its length (one ray crosses all 16384 x
    cells) and across it (the detector spans its length)."""
    from deepdrr_amd import phantoms
    from deepdrr_amd.volume import Volume

    nx, ny, nz = 16384, 3, 2
    rng = np.random.default_rng(21)
    dens = (0.5 + rng.random((nx, ny, nz), dtype=np.float32)).astype(np.float32)
    bone = ((np.arange(nx) // 97) % 2 == 0)[:, None, None] & np.ones((1, ny, nz), bool)
    spacing = (0.01, 1.0, 1.0)
    vol = Volume.from_parameters(dens, {"bone": bone, "soft tissue": ~bone},
                                 origin=phantoms._centered_origin((nx, ny, nz), spacing), spacing=spacing,
                                 anatomical_coordinate_system="LPS")
    d = np.array(direction)
    along = abs(d[0]) > abs(d[1])
    w, h, pix = (8, 8, 0.5) if along else (512, 4, 0.8)
    proj = phantoms.pose_projection(w, h, pix, [0.0, 0.0, 0.0], d / np.linalg.norm(d), 0.5)
    case = scenes.make_case(f"rod_{'along' if along else 'across'}", [vol], proj)
    print(f"case {case.name}", flush=True)
    r = GpuRunner(case)
    try:
        g = r.project()
        o = _oracle(case)
        assert (o[2] > 0).any(), "the view misses the rod"
        _check(case, g, o)
    finally:
        r.close()


def test_widest_detector(gpu):
    """A detector 65536 pixels wide (the largest a view may have) and 3 rows:
    bit-exact area densities, and batched images (both tile shapes, ragged
    last tiles) bit-identical to the single view."""
    import torch

    from deepdrr_amd import phantoms
    from deepdrr_amd.batch import BatchPlan

    vol = phantoms.small_ct(24)
    proj = phantoms.pose_projection(65536, 3, 0.002, [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], 0.5)
    case = scenes.make_case("widest_detector", [vol], proj)
    print("case widest_detector", flush=True)
    r = GpuRunner(case)
    try:
        g = r.project()
        o = _oracle(case)
        assert (o[2] > 0).sum() > 1000
        _check(case, g, o)
        for tuning in ((8, 8, 4, 8), (16, 4, 2, 4)):
            plan = BatchPlan(r.ctx, [r.vols[0]], [case.view], [0], neglog=False, tuning=tuning)
            plan.launch()
            torch.cuda.synchronize()
            I_b = plan.fetch("intensity")[0]
            plan.close()
            assert np.array_equal(g[0].view(np.uint32), I_b.view(np.uint32)), tuning
    finally:
        r.close()
