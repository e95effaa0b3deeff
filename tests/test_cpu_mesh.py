"""The mesh branch of the oracle against closed forms.

The oracle (oracle/mesh_hits.c + the mesh branch of oracle/transcription.c)
stands in for deepdrr's OpenGL depth peeling and peel post-processing
(projector.py:1055-1330, peel_postprocess_kernel.cu) and follows
projectKernel's mesh statements (project_kernel.cu:363-375, 498-521,
569-579).  The reference's own mesh tests compare PNG renders of a CT that
needs a download, so the oracle is pinned here by geometry instead:
* an additive box adds density x (analytic chord of the ray through the box);
  a box wound inside out adds nothing (the facing convention);
* a subtractive box inside a uniform volume removes density x chord from the
  volume's material (to within the two boundary samples) and its hidden range
  is the analytic entry/exit pair;
* a higher layer's subtractive box takes its overlap off a lower layer's
  additive box (the box intersection is a box: closed form);
* overlapping subtractive boxes of one layer, and of two layers, hide their
  union; a source inside a mesh hides nothing (the "sea level" rule);
* the mesh's volume equals the sum over nearly parallel rays of chord x
  pixel area (a closed non-convex mesh, within 2%);
* a ray through the diagonal shared by a face's two triangles counts once.
The meshes are the reference's own test resources
(tests/resources/10cmcube.stl, xyzfaces.stl).
"""
from pathlib import Path

import numpy as np
import pytest

from deepdrr_amd import phantoms
from deepdrr_amd.mesh import DRRMaterial, Mesh, TriangleMesh, read_stl, signed_volume
from tests import scenes

GOLD = Path(__file__).resolve().parent / "golden" / "meshes"


def box_mesh(center, half, material: DRRMaterial, inside_out=False) -> Mesh:
    v, f = read_stl(GOLD / "10cmcube.stl")  # the +-0.1 cube
    v = v / 0.1 * np.asarray(half, np.float64) + np.asarray(center, np.float64)
    if inside_out:
        f = f[:, ::-1]
    return Mesh(mesh=TriangleMesh.from_arrays(v, f, material))


def rays(view):
    """Unit world ray of every pixel [H*W, 3] (FP32 as the kernels form it)."""
    F32 = np.float32
    W = np.array(view.world_from_index[:], F32)
    vv, uu = np.meshgrid(np.arange(view.height), np.arange(view.width), indexing="ij")
    u = uu.reshape(-1).astype(F32) + F32(0.5)
    v = vv.reshape(-1).astype(F32) + F32(0.5)
    r = np.stack([(u * W[3 * a] + v * W[3 * a + 1]) + W[3 * a + 2] for a in range(3)], 1)
    L = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2])
    return (r * (F32(1) / L)[:, None]).astype(np.float64)


def box_chord(o, d, lo, hi):
    """Entry, exit (FP64 slab test) and the distance of the entry/exit points
    to the box's edges."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - o) / d
        t1 = (hi - o) / d
    tmin = np.max(np.minimum(t0, t1), 1)
    tmax = np.min(np.maximum(t0, t1), 1)
    hit = tmax > tmin
    edge = np.full(len(d), np.inf)
    for t in (tmin, tmax):
        p = o + t[:, None] * d
        dist = np.minimum(np.abs(p - lo), np.abs(p - hi))  # per axis; one axis is ~0 (the face)
        edge = np.minimum(edge, np.sort(dist, 1)[:, 1])
    return np.where(hit, tmin, 0.0), np.where(hit, tmax, 0.0), hit, edge


def _ap(w=48, h=40, pix=0.9, sdd=1020.0):
    return phantoms.ap_device(w, h, pix, sdd=sdd).get_camera_projection()


def test_additive_box_chord_and_facing():
    proj = _ap()
    mat = DRRMaterial("titanium", density=4.5)
    mc = scenes.make_mesh_case("add_box", [], [box_mesh([1.0, 2.0, -1.5], [8, 6, 7], mat)], proj)
    I, P, AD = mc.oracle()
    ti = mc.case.materials.index("titanium")
    d = rays(mc.case.view)
    ent, ext, hit, edge = box_chord(mc.source.astype(float), d, [-7, -4, -8.5], [9, 8, 5.5])
    got = AD[ti].reshape(-1) * 10.0
    ok = hit & (edge > 1e-3)
    assert ok.sum() > 500 and (~hit).sum() > 100
    # FP32 hit distances near 510 mm are good to a few 1e-5 mm
    np.testing.assert_allclose(got[ok], 4.5 * (ext - ent)[ok], rtol=1e-6, atol=4.5 * 5e-4)
    assert np.all(got[~hit] == 0)
    # the same box wound inside out sums to a negative value, clamped to 0
    mc2 = scenes.make_mesh_case("inside_out", [], [box_mesh([1.0, 2.0, -1.5], [8, 6, 7], mat, inside_out=True)],
                                proj)
    assert np.all(mc2.oracle()[2] == 0)


def test_shared_diagonal_counts_once():
    """The central ray of an AP view meets the box's front and back faces on the
    diagonal the faces' two triangles share (edge-inclusive tests hit both)."""
    proj = _ap(w=5, h=5, pix=0.9)
    mat = DRRMaterial("bone", density=2.0)
    mc = scenes.make_mesh_case("diag", [], [box_mesh([0, 0, 0], [5, 5, 5], mat)], proj)
    AD = mc.oracle()[2][mc.case.materials.index("bone")]
    assert AD[2, 2] * 10 == pytest.approx(2.0 * 10.0, rel=1e-5)


def test_subtractive_box_in_uniform_volume():
    proj = _ap(w=40, h=36, pix=0.9)
    vol = phantoms.uniform_volume((40, 40, 40), density=1.0, material="soft tissue")
    box = box_mesh([0.5, -1.0, 1.0], [7, 5, 6], DRRMaterial("bone", density=1.92, subtractive=True, layer=1))
    mc = scenes.make_mesh_case("sub_box", [vol], [box], proj)
    base = scenes.make_mesh_case("no_mesh", [vol], [], proj)
    _, _, AD = mc.oracle()
    _, _, AD0 = base.oracle()
    st = mc.case.materials.index("soft tissue")
    bo = mc.case.materials.index("bone")
    d = rays(mc.case.view)
    ent, ext, hit, edge = box_chord(mc.source.astype(float), d, [-6.5, -6, -5], [7.5, 4, 7])
    ok = hit & (edge > 1e-3)
    removed = (AD0[base.case.materials.index("soft tissue")] - AD[st]).reshape(-1) * 10.0
    # +-1 sample at each end of the hidden range (step 0.1 mm, density 1)
    assert np.all(np.abs(removed[ok] - (ext - ent)[ok]) <= 0.2 + 1e-4)
    assert np.all(removed[~hit] == 0)
    np.testing.assert_allclose(AD[bo].reshape(-1)[ok] * 10.0, 1.92 * (ext - ent)[ok], rtol=1e-6, atol=1.92 * 5e-4)
    # the hidden range is (entry, exit]
    iv, cnt, add, nu = mc.oracle_rays()
    cnt = cnt.reshape(-1)
    iv = iv.reshape(len(cnt), -1, 2)
    assert np.all(cnt[ok] == 1) and np.all(cnt[~hit] == 0)
    np.testing.assert_allclose(iv[ok, 0, 0], ent[ok], atol=2e-4)
    np.testing.assert_allclose(iv[ok, 0, 1], ext[ok], atol=2e-4)


def test_higher_layer_takes_overlap_off_additive():
    proj = _ap(w=44, h=40, pix=0.9)
    A = box_mesh([0, 0, 0], [8, 8, 8], DRRMaterial("bone", density=2.0, layer=0))
    B = box_mesh([5, 2, 4], [5, 9, 6], DRRMaterial("titanium", density=4.5, subtractive=True, layer=1))
    mc = scenes.make_mesh_case("layers", [], [A, B], proj)
    _, _, AD = mc.oracle()
    bo, ti = mc.case.materials.index("bone"), mc.case.materials.index("titanium")
    o, d = mc.source.astype(float), rays(mc.case.view)
    eA, xA, hA, edA = box_chord(o, d, [-8, -8, -8], [8, 8, 8])
    eB, xB, hB, edB = box_chord(o, d, [0, -7, -2], [10, 11, 10])
    eI, xI, hI, edI = box_chord(o, d, [0, -7, -2], [8, 8, 8])  # A intersect B
    ok = (edA > 1e-3) & (edB > 1e-3) & (edI > 1e-3)
    chordA, chordB, chordI = (xA - eA) * hA, (xB - eB) * hB, (xI - eI) * hI
    assert (ok & hI).sum() > 100 and (ok & hA & ~hB).sum() > 100
    np.testing.assert_allclose(AD[bo].reshape(-1)[ok] * 10, 2.0 * (chordA - chordI)[ok], rtol=1e-6, atol=2.0 * 1e-3)
    np.testing.assert_allclose(AD[ti].reshape(-1)[ok] * 10, 4.5 * chordB[ok], rtol=1e-6, atol=4.5 * 5e-4)


def test_union_of_subtractive_boxes_and_sea_level():
    proj = _ap(w=30, h=30, pix=0.9)
    sub = lambda layer: DRRMaterial("bone", density=1.0, subtractive=True, additive=False, layer=layer)
    # one layer, overlapping along the beam: one range, entry of the first to exit of the second
    mc = scenes.make_mesh_case("same_layer", [], [box_mesh([0, -3, 0], [4, 4, 4], sub(0)),
                                                  box_mesh([0, 3, 0], [4, 4, 4], sub(0))], proj)
    iv, cnt, _, _ = mc.oracle_rays()
    o, d = mc.source.astype(float), rays(mc.case.view)
    e, x, hit, edge = box_chord(o, d, [-4, -7, -4], [4, 7, 4])
    ok = hit & (edge > 1e-3)
    c = cnt.reshape(-1)
    v = iv.reshape(len(c), -1, 2)
    assert np.all(c[ok] == 1)
    np.testing.assert_allclose(v[ok, 0], np.stack([e, x], 1)[ok], atol=2e-4)
    # two layers: the same union
    mc2 = scenes.make_mesh_case("two_layers", [], [box_mesh([0, -3, 0], [4, 4, 4], sub(0)),
                                                   box_mesh([0, 3, 0], [4, 4, 4], sub(1))], proj)
    iv2, cnt2, _, _ = mc2.oracle_rays()
    assert np.array_equal(cnt2, cnt)
    np.testing.assert_allclose(iv2.reshape(len(c), -1, 2)[ok, 0], v[ok, 0], atol=2e-4)
    # separated along the beam: two ranges
    mc3 = scenes.make_mesh_case("apart", [], [box_mesh([0, -8, 0], [4, 3, 4], sub(0)),
                                              box_mesh([0, 8, 0], [4, 3, 4], sub(0))], proj)
    _, cnt3, _, _ = mc3.oracle_rays()
    assert np.all(cnt3.reshape(-1)[ok] == 2)
    # the source inside a subtractive mesh: its first hit is an exit, the
    # altitude dips below the sea level and the tide rules drop the ray's hits
    src = np.asarray(proj.center_in_world, float)
    mc4 = scenes.make_mesh_case("source_inside", [], [box_mesh(src, [30, 30, 30], sub(0))], proj)
    _, cnt4, _, _ = mc4.oracle_rays()
    assert np.all(cnt4 == 0)


def test_closed_mesh_volume_from_chords():
    """Nearly parallel rays (source 5e4 mm away): the sum of chord x pixel area
    equals the volume of a closed non-convex mesh (xyzfaces.stl, 156 faces)."""
    v, f = read_stl(GOLD / "xyzfaces.stl")
    v = (v - 0.1) * 100.0  # a 20 mm block centred on the origin
    vol_ref = signed_volume(v, f)
    sdd = 1.0e5
    n, pix_iso = 90, 0.25
    proj = phantoms.ap_device(n, n, 2 * pix_iso, sdd=sdd).get_camera_projection()
    mesh = Mesh(mesh=TriangleMesh.from_arrays(v, f, DRRMaterial("soft tissue", density=1.0)))
    mc = scenes.make_mesh_case("xyzfaces", [], [mesh], proj, sdd=sdd)
    AD = mc.oracle()[2][mc.case.materials.index("soft tissue")]
    est = float(AD.astype(np.float64).sum() * 10.0 * pix_iso ** 2)
    assert (AD > 0).sum() > 1000
    assert abs(est - vol_ref) <= 0.02 * vol_ref, (est, vol_ref)


def test_projector_accepts_meshes():
    """Mesh renderables join the projector's materials (projector.py:467-559);
    other renderables and bad mesh settings are refused."""
    from deepdrr_amd import Projector

    vol = phantoms.small_ct(16)
    m = box_mesh([0, 0, 0], [3, 3, 3], DRRMaterial("titanium", density=4.5, subtractive=True, layer=1))
    p = Projector([vol, m])
    assert p.volumes == [vol] and p.meshes == [m] and "titanium" in p.all_materials
    assert p.all_materials == sorted(p.all_materials)
    Projector([m])  # meshes alone
    with pytest.raises(ValueError):
        Projector([vol, m], max_mesh_hits=256)
    with pytest.raises(ValueError):
        Projector([vol, m], mesh_layers=5)
    with pytest.raises(ValueError):
        Projector([vol, object()])
    with pytest.raises(ValueError):
        DRRMaterial("not a material")
