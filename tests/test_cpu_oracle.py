"""Pins for the CPU oracle (C transcription + FP64 restatement).

The reference's own goldens (tests/reference/*.png) cannot run offline: they need
the CT-chest.nrrd download, killeengeo and CUDA.  So the oracle is pinned by what
the mathematics fixes:
closed forms (empty volume, homogeneous slab Beer-Lambert), exact invariants
(density scaling by 2 doubles area densities bit-exactly), the two
independent restatements agreeing on random tiny scenes, and the reference's
documented quirks (shared label cache across volumes, half weight of the last
step).  Each check is chosen so a plausible slip -- a dropped /10 or step
factor, a transposed axis, a wrong material index, a missing half weight --
fails it.
"""
import math

import numpy as np
import pytest

from deepdrr_amd import geo, phantoms, workload
from deepdrr_amd import projector as P
from deepdrr_amd._abi_types import Scene
from deepdrr_amd.volume import Volume
from oracle import ctrans, fp64
from tests import scenes


def _run_both(case):
    c = ctrans.project(case.volumes, case.view, case.scene, case.E, case.pdf, case.mu, 4)
    f = fp64.project_fp64(case.volumes, case.view_dict(), case.scene_dict(), case.E, case.pdf, case.mu)
    return c, f


def test_empty_volume_closed_form():
    vol = phantoms.uniform_volume((6, 6, 6), density=0.0)
    dev = phantoms.ap_device(5, 4, 1.0)
    case = scenes.make_case("empty", [vol], dev.get_camera_projection())
    (I, P_, AD), (I64, P64, AD64) = _run_both(case)
    assert np.all(AD == 0) and np.all(AD64 == 0)
    np.testing.assert_allclose(P_, case.pdf.astype(np.float64).sum(), rtol=1e-6)
    np.testing.assert_allclose(I, (case.E.astype(np.float64) * case.pdf).sum(), rtol=1e-6)
    np.testing.assert_allclose(I64, (case.E.astype(np.float64) * case.pdf).sum(), rtol=1e-12)


def _slab_case(rho=1.3, n=(10, 40, 10), spacing=(2.0, 1.0, 2.0), step=0.1, material="bone"):
    """A homogeneous box; AP rays along world +y cross it along IJK k."""
    vol = phantoms.uniform_volume(n, spacing=spacing, density=rho, material=material)
    dev = phantoms.ap_device(4, 4, 0.5)
    return scenes.make_case("slab", [vol], dev.get_camera_projection(), step=step), vol


def test_homogeneous_slab_beer_lambert():
    """ad = rho * chord / 10 (g/cm^2) up to the trapezoid ends; I = sum E pdf exp(-mu ad)."""
    rho = 1.3
    case, vol = _slab_case(rho)
    (I, P_, AD), (I64, P64, AD64) = _run_both(case)
    # chord of each ray through the box from the slab interval of the ray
    mn, mx, L = workload.ray_intervals(case.view)
    chord = (mx - mn).reshape(4, 4)
    ad_closed = rho * chord / 10.0
    step = case.scene.step
    assert np.all(np.abs(AD[0] - ad_closed) <= rho * 1.5 * step / 10.0)
    mu = case.mu.reshape(len(case.E), 1)[:, 0].astype(np.float64)
    I_closed = ((case.E * case.pdf)[:, None, None] * np.exp(-mu[:, None, None] * AD64[0][None])).sum(0)
    np.testing.assert_allclose(I64, I_closed, rtol=1e-8)
    np.testing.assert_allclose(I, I_closed, rtol=2e-5)


def test_density_scaling_is_exact():
    """Doubling every density doubles every FP32 product, so area densities double bit-exactly."""
    case = scenes.small_case(12, (7, 5))
    _, _, AD = ctrans.project(case.volumes, case.view, case.scene, case.E, case.pdf, case.mu, 4)
    d, l = case.volumes[0]
    case2 = scenes.Case("x2", [(d * np.float32(2), l)], case.view, case.scene, case.E, case.pdf, case.mu)
    _, _, AD2 = ctrans.project(case2.volumes, case2.view, case2.scene, case2.E, case2.pdf, case2.mu, 4)
    assert np.array_equal((AD * 2).view(np.uint32), AD2.view(np.uint32))


def test_two_material_split_indices():
    """Left half bone, right half soft tissue: each material's area density only
    where its rays cross it (catches swapped material indices / transposed axes)."""
    n = (16, 16, 16)
    data = np.full(n, 1.0, np.float32)
    bone = np.zeros(n, bool)
    bone[:8] = True  # i < 8 -> world x < 0 (LPS: x = sx * i + origin)
    vol = Volume.from_parameters(data, {"soft tissue": ~bone, "bone": bone},
                                 origin=phantoms._centered_origin(n, (1, 1, 1)), spacing=(1, 1, 1),
                                 anatomical_coordinate_system="LPS")
    dev = phantoms.ap_device(8, 2, 2.0)  # rays spread along world x
    case = scenes.make_case("split", [vol], dev.get_camera_projection())
    (I, P_, AD), (I64, P64, AD64) = _run_both(case)
    ib, it = case.materials.index("bone"), case.materials.index("soft tissue")
    # The reference samples at q = ijk - 1 (project_kernel.cu:400-404): the
    # volume appears shifted by +1 voxel, so the bone half (i < 8) covers
    # q < 9, i.e. 9 mm of the 16 mm width.  At magnification 2 with 2 mm
    # pixels that is 5 of the 8 columns, contiguous from one edge.
    bone_cols = AD[ib].mean(0) > AD[it].mean(0)
    assert bone_cols.sum() == 5 and (bone_cols[:5].all() or bone_cols[3:].all())
    np.testing.assert_allclose(AD, AD64, rtol=1e-5, atol=1e-7)


def test_last_step_half_weight():
    """The last sample is half-weighted (project_kernel.cu:537): a 1-cell slab
    with step larger than the slab gives exactly 0.5 * rho * step / 10."""
    vol = phantoms.uniform_volume((8, 8, 1), spacing=(1, 1, 1), density=1.0, material="bone")
    dev = phantoms.ap_device(2, 2, 0.2)
    case = scenes.make_case("half", [vol], dev.get_camera_projection(), step=5.0)
    (I, P_, AD), _ = _run_both(case)
    # N steps from ~1 mm to the exit; exactly one sample lands in the 1 mm slab
    # only if it is the last one -> either 0 or the half weight
    vals = np.unique(np.round(AD[0] / (0.5 * 5.0 / 10.0), 6))
    assert set(vals.tolist()) <= {0.0, 1.0, 2.0}


def test_shared_label_cache_quirk():
    """Two volumes on the same grid share the label cache (project_kernel.cu:394-431):
    volume 1 never refetches and uses volume 0's labels."""
    n = (6, 6, 6)
    origin = phantoms._centered_origin(n, (1, 1, 1))
    v0 = Volume.from_parameters(np.zeros(n, np.float32), {"bone": np.ones(n, bool)}, origin=origin,
                                anatomical_coordinate_system="LPS")
    v1 = Volume.from_parameters(np.ones(n, np.float32), {"soft tissue": np.ones(n, bool)}, origin=origin,
                                anatomical_coordinate_system="LPS")
    dev = phantoms.ap_device(3, 3, 1.0)
    case = scenes.make_case("quirk", [v0, v1], dev.get_camera_projection(), priorities=[0, 0])
    (I, P_, AD), (I64, P64, AD64) = _run_both(case)
    ib, it = case.materials.index("bone"), case.materials.index("soft tissue")
    assert np.all(AD[it] == 0) and np.all(AD[ib] > 0)
    np.testing.assert_allclose(AD, AD64, rtol=1e-5)


def test_priority_selects_volume():
    """Where a higher-priority volume has labels, lower-priority ones are ignored."""
    big = phantoms.uniform_volume((12, 12, 12), density=1.0, material="soft tissue")
    small = phantoms.uniform_volume((4, 4, 4), spacing=(1, 1, 1), density=2.0, material="bone")
    dev = phantoms.ap_device(6, 6, 1.0)
    case_hi = scenes.make_case("prio", [big, small], dev.get_camera_projection(), priorities=[1, 0])
    (I, P_, AD), (I64, P64, AD64) = _run_both(case_hi)
    ib = case_hi.materials.index("bone")
    assert AD[ib].max() > 0
    np.testing.assert_allclose(AD, AD64, rtol=1e-5, atol=1e-7)
    case_lo = scenes.make_case("prio", [big, small], dev.get_camera_projection(), priorities=[0, 1])
    (_, _, AD_lo), _ = _run_both(case_lo)
    assert AD_lo[case_lo.materials.index("bone")].max() == 0


@pytest.mark.parametrize("mk", [lambda: scenes.small_case(16, (9, 7)),
                                lambda: scenes.small_case(16, (9, 7), atten=True),
                                lambda: scenes.oblique_case(12, (8, 8), seed=2),
                                lambda: scenes.multi_case()])
def test_transcription_matches_fp64(mk):
    case = mk()
    (I, P_, AD), (I64, P64, AD64) = _run_both(case)
    tol = 2e-4 if case.scene.attenuate_outside else 1e-5
    np.testing.assert_allclose(I, I64, rtol=tol)
    np.testing.assert_allclose(P_, P64, rtol=tol)


def test_random_tiny_scenes_agree():
    """Brute force on random 2^3..6^3 volumes and random poses."""
    rng = np.random.default_rng(5)
    for k in range(6):
        n = tuple(int(x) for x in rng.integers(2, 7, size=3))
        dens = rng.uniform(0.0, 2.0, size=n).astype(np.float32)
        lab = rng.integers(0, 3, size=n)
        mats = {"air": lab == 0, "soft tissue": lab == 1, "bone": lab == 2}
        vol = Volume.from_parameters(dens, mats, origin=phantoms._centered_origin(n, (1, 1, 1)),
                                     anatomical_coordinate_system="LPS")
        (p, d, f), = phantoms.batch_poses(1, seed=k, radius=1.0, cone_deg=60)
        proj = phantoms.pose_projection(5, 4, 1.0, p, d, f)
        case = scenes.make_case(f"rand{k}", [vol], proj, step=0.05)
        (I, P_, AD), (I64, P64, AD64) = _run_both(case)
        np.testing.assert_allclose(I, I64, rtol=1e-5)
        # FP32 sample positions shift the weights of small components slightly
        np.testing.assert_allclose(AD, AD64, rtol=1e-4, atol=1e-5 * float(AD64.max()))


def test_solid_angle_sums_to_rectangle():
    """Per-pixel solid angles sum to the detector rectangle's solid angle."""
    w, h, pix, sdd = 16, 12, 2.0, 1000.0
    dev = phantoms.ap_device(w, h, pix, sdd=sdd)
    view = P.build_view([phantoms.uniform_volume((2, 2, 2))], [0], dev.get_camera_projection(), 4000.0)
    W = np.array(view.world_from_index[:], np.float32)
    sa = ctrans.solid_angle(W, w, h)
    # rays have unit camera depth: the rectangle is at depth 1 with half-sizes a/2, b/2
    a, b = w * pix / sdd, h * pix / sdd
    omega = 4 * math.asin(math.sin(math.atan(a / 2)) * math.sin(math.atan(b / 2)))
    np.testing.assert_allclose(sa.astype(np.float64).sum(), omega, rtol=1e-4)


def test_dense_reference_matches_transcription():
    """The float64 grid_sample reference (used to check the adjoint) renders what
    the transcription renders."""
    import torch

    from deepdrr_amd.autograd import PoseGeometry
    from deepdrr_amd.material import attenuation_table
    from deepdrr_amd.spectral_data import spectrum_tables, synthetic_120kv
    from deepdrr_amd._abi_types import Scene
    from tests.dense_ref import DenseDRR

    rng = np.random.default_rng(0)
    n = 14
    dens = np.zeros((n, n, n), np.float32)
    dens[2:-2, 2:-2, 2:-2] = rng.uniform(0.2, 1.8, size=(n - 4,) * 3)
    labels = np.where(dens > 1.2, 1, np.where(dens > 0, 2, 0)).astype(np.uint8)
    W = np.eye(4)
    W[:3, :3] = [[1.0, 0, 0], [0, 0, 1.0], [0, -1.0, 0]]
    W[:3, 3] = phantoms._centered_origin((n, n, n), (1, 1, 1))
    geom = PoseGeometry((n, n, n), W)
    projs = [phantoms.pose_projection(8, 8, 3.0, p, d, f)
             for p, d, f in phantoms.batch_poses(2, seed=4, radius=2.0, cone_deg=40)]
    theta = np.array([0.3, -0.2, 0.1, 0.02, -0.01, 0.03])
    views = geom.views(projs, theta)
    E, pdf = spectrum_tables(synthetic_120kv())
    mu = attenuation_table(["air", "bone", "soft tissue"], E)
    sc = Scene(step=0.25, n_materials=3, attenuate_outside=0, air_index=0)
    d_kji = np.ascontiguousarray(np.moveaxis(dens, [0, 1, 2], [2, 1, 0]))
    l_kji = np.ascontiguousarray(np.moveaxis(labels, [0, 1, 2], [2, 1, 0]))
    ref = DenseDRR(views, W, geom.center, 0.25, E, pdf, mu, (n, n, n))
    oh = torch.tensor(np.stack([(labels == m).astype(np.float64) for m in range(3)]))
    I64 = ref.render(torch.tensor(dens.astype(np.float64)), oh,
                     torch.tensor(theta, dtype=torch.float64)).numpy()
    for i, (view, _) in enumerate(views):
        I, _, _ = ctrans.project([(d_kji, l_kji)], view, sc, E, pdf, mu, 4)
        np.testing.assert_allclose(I, I64[i], rtol=2e-5)  # f32 vs f64 geometry and integrand
