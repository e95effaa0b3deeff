"""Host-side logic and the C-ABI library (no GPU calls)."""
import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from deepdrr_amd import Projector, geo, phantoms, utils
from deepdrr_amd._abi import LIB_PATH, SIGNATURES
from deepdrr_amd.material import Material, attenuation_table, log_interp
from deepdrr_amd.spectral_data import spectrum_tables, spectrums, synthetic_120kv
from deepdrr_amd.volume import Volume, conv_hu_to_density, conv_hu_to_materials_thresholding

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    hdr = (ROOT / "include" / "deepdrr_amd.h").read_text()
    return sorted(set(re.findall(r"\b(drr_[a-z_0-9]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    assert LIB_PATH.exists(), "run python -m deepdrr_amd.build first"
    lib = C.CDLL(str(LIB_PATH))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared()) == set(SIGNATURES), "ctypes signatures out of sync with the header"


def test_spectra_match_reference_data():
    assert {k: v.shape for k, v in spectrums.items() if k != "SYNTH120KV_20"} == {
        "60KV_AL35": (91, 2), "90KV_AL40": (151, 2), "120KV_AL43": (211, 2)}
    for k in ("60KV_AL35", "90KV_AL40", "120KV_AL43"):
        assert spectrums[k][-1, 1] < 0  # the last bin of each MC-GPU spectrum is negative
    E, pdf = spectrum_tables(spectrums["90KV_AL40"])
    assert E.dtype == np.float32 and abs(float(pdf.astype(np.float64).sum()) - 1) < 1e-6
    assert E[0] == np.float32(15.0) and E[-1] == np.float32(90.0)
    s = synthetic_120kv()
    assert s.shape == (20, 2) and np.all(s[:, 1] > 0)


def test_material_log_interp_at_nodes_and_between():
    bone = Material.from_string("bone")
    e = bone.energy
    mu = bone.mu_over_rho
    # at a table node (not an edge duplicate) the interpolation returns the node value
    i = int(np.argmin(np.abs(e - 0.06)))
    assert bone.get(e[i]).mu_over_rho == pytest.approx(mu[i], rel=1e-12)
    # between nodes: straight line in log-log space
    x = math.sqrt(e[i] * e[i + 1])
    expect = math.sqrt(mu[i] * mu[i + 1])
    assert bone.get(x).mu_over_rho == pytest.approx(expect, rel=1e-12)
    # named materials map to the NIST files (mappings.py:96,100)
    assert Material.from_string("soft tissue").name == "tissue_soft"
    assert Material.from_string("titanium").name == "22_Ti_Titanium"
    with pytest.raises(AttributeError):
        Material.from_string("unobtainium")


def test_attenuation_table_layout():
    E, _ = spectrum_tables(synthetic_120kv())
    mats = ["air", "bone", "soft tissue"]
    t = attenuation_table(mats, E)
    assert t.shape == (len(E) * 3,) and t.dtype == np.float32
    for b in (0, 7, 19):
        for m, name in enumerate(mats):
            ref = Material.from_string(name).get(float(E[b]) / 1000.0).mu_over_rho
            assert t[b * 3 + m] == np.float32(ref)


def test_hu_conversion_closed_form():
    hu = np.array([-1000, -800, 0, 350, 1000, 3000], dtype=np.int16)
    rho = conv_hu_to_density(hu)
    expect = np.maximum(np.minimum(0.001029 * hu + 1.030, 0.0005886 * hu + 1.03), 0)
    np.testing.assert_array_equal(rho, expect)
    assert rho[0] == pytest.approx(0.001, abs=1e-12) and rho[2] == pytest.approx(1.03)
    m = conv_hu_to_materials_thresholding(hu)
    assert m["air"].tolist() == [True, True, False, False, False, False]
    assert m["bone"].tolist() == [False, False, False, False, True, True]


def test_camera_geometry_unit_depth():
    dev = phantoms.ap_device(64, 48, 0.5, sdd=1000.0)
    proj = dev.get_camera_projection()
    W = np.array(proj.world_from_index.data)
    assert W.shape == (4, 3) and np.all(W[3] == 0)
    # the principal point ray has unit length and points along the principal ray
    r = W[:3] @ np.array([32.0, 24.0, 1.0])
    assert np.linalg.norm(r) == pytest.approx(1.0)
    np.testing.assert_allclose(r, [0, 1, 0], atol=1e-12)
    # source at point - 0.5 * SDD * direction (source_to_point_fraction 0.5)
    np.testing.assert_allclose(np.array(proj.center_in_world), [0, -500, 0], atol=1e-9)
    # projecting a world point lands on the pixel whose ray passes through it
    x = np.array([10.0, 100.0, -5.0, 1.0])
    uvw = proj.data @ x
    u, v = uvw[0] / uvw[2], uvw[1] / uvw[2]
    ray = W[:3] @ np.array([u, v, 1.0])
    d = x[:3] - np.array(proj.center_in_world)
    np.testing.assert_allclose(np.cross(ray, d), 0, atol=1e-9)


def test_neglog_closed_form():
    x = np.array([[[1.0, 2.0], [4.0, 8.0]]], np.float32)
    y = utils.neglog(x)
    z = -np.log(x + (1.0 + 0.01))
    np.testing.assert_allclose(y, (z - z.min()) / (z.max() - z.min()), rtol=1e-6)
    assert np.all(utils.neglog(np.ones((2, 3, 3), np.float32)) == 0)


def test_projector_constructor_errors():
    vol = phantoms.small_ct(8)
    with pytest.raises(ValueError):
        Projector([vol, "not a volume"])
    with pytest.raises(AssertionError):
        Projector(vol, priorities=[3])
    with pytest.raises(ValueError):
        Projector(vol, scatter_num=-1)
    with pytest.raises(ValueError):
        Projector(vol, scatter_num=10)  # needs a device
    # as in the reference (projector.py:530-531), scatter_num > 0 is refused
    from deepdrr_amd.projector import DeprecationError

    dev = phantoms.ap_device(8, 8, 1.0)
    with pytest.raises(DeprecationError):
        Projector(vol, device=dev, scatter_num=10)
    with pytest.raises(DeprecationError):
        Projector(vol, device=dev, add_scatter=True)
    with pytest.raises(ValueError):
        Projector(vol, mc_scatter_histories=10)  # needs a device
    with pytest.raises(ValueError):
        Projector(vol, mc_scatter_histories=-1)
    with pytest.raises(ValueError):
        Projector(vol, max_mesh_hits=6)
    with pytest.raises(KeyError):
        Projector(vol, spectrum="100KV")
    with pytest.raises(ValueError):
        Projector(vol, march="approximate")
    # 16 materials are rendered (4-bit labels); a 17th is refused
    from deepdrr_amd.volume import Volume

    names = ["air", "bone", "soft tissue", "lung", "titanium", "muscle", "blood", "teflon", "iron", "lead",
             "copper", "polyethylene", "concrete", "Al", "Ca", "W", "Au"]
    lab = np.arange(17 * 8).reshape(17, 8, 1) % 17
    dens = np.ones(lab.shape, np.float32)
    for k in (16, 17):
        v = Volume.from_parameters(dens, {m: lab == i for i, m in enumerate(names[:k])}, origin=[0, 0, 0],
                                   anatomical_coordinate_system="LPS")
        if k == 16:
            assert len(Projector(v).all_materials) == 16
        else:
            with pytest.raises(ValueError, match="16"):
                Projector(v)
    p = Projector(vol)
    assert p.all_materials == ["air", "bone", "soft tissue"]
    with pytest.raises(RuntimeError):
        p.project()
    with pytest.raises(RuntimeError):
        _ = p.camera_intrinsics
    vols = [phantoms.small_ct(8), phantoms.small_ct(8)]
    assert Projector(vols).priorities == [1, 0]
    q = Projector(vol, attenuate_outside_volume=True)
    assert q.air_index == q.all_materials.index("air")


def test_volume_material_formatting():
    data = np.ones((3, 3, 3), np.float32)
    a = np.zeros((3, 3, 3), bool)
    a[0] = True
    b = np.zeros((3, 3, 3), bool)
    b[:2] = True
    v = Volume.from_parameters(data, {"air": a, "bone": b}, origin=[0, 0, 0], anatomical_coordinate_system="LPS")
    d, seg = v.materials
    assert d == {"air": 0, "bone": 1}
    assert seg[0].max() == 1 and seg[2].max() == 0  # later masks win; uncovered voxels -> 0
    with pytest.raises(NotImplementedError):
        Volume.from_parameters(data, {"air": a}, origin=[0, 0, 0])
