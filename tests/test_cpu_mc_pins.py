"""Pins of the CPU Monte-Carlo counterpart's transport (oracle/mc_cpu.c) by
what the mathematics fixes, so that a slip shared by the GPU kernel and its
CPU counterpart (same formulas in both) cannot hide behind their agreement:

  * RANECU: the Schrage-form update equals the modular products
    s1 <- 40014 s1 mod 2147483563, s2 <- 40692 s2 mod 2147483399 and the
    output (s1 - s2) mod (2147483563 - 1) scaled by 2^-31 (L'Ecuyer 1988),
    checked with exact integer arithmetic;
  * the direction rotation: a unit vector at polar cosine ct from the incoming
    direction, and a rigid turn by the azimuth phi about it;
  * energy balance: per history batch the emitted energy equals scored +
    photoelectric + Compton recoil + cutoff + escaped;
  * single scatter from a thin slab hit by a pencil beam: the single-Compton
    and single-Rayleigh images against the analytic Klein-Nishina / Thomson
    angular densities times each pixel's solid angle, the Compton energy
    relation and the attenuation of the scattered photon on its way out.
"""
import math

import numpy as np
import pytest

from deepdrr_amd._abi_types import MCParams
from oracle import ctrans

M1, A1 = 2147483563, 40014
M2, A2 = 2147483399, 40692


@pytest.mark.parametrize("seed,h", [(12345, 0), (12345, 1), (1, 987654321), (2 ** 63 + 5, 42)])
def test_ranecu_is_lecuyer_combined_generator(seed, h):
    (s1, s2), o1, o2, u = ctrans.ranecu(seed, h, 5000)
    assert 1 <= s1 < M1 and 1 <= s2 < M2
    x1, x2 = s1, s2
    for i in range(len(u)):
        x1 = (A1 * x1) % M1
        x2 = (A2 * x2) % M2
        assert (o1[i], o2[i]) == (x1, x2), i
        z = x1 - x2
        if z < 1:
            z += M1 - 1
        want = np.float32(min(np.float32(np.float32(z) * np.float32(2.0 ** -31)), np.float32(0.99999994)))
        assert u[i] == want, i


def test_ranecu_streams_uniform_and_independent():
    _, _, _, a = ctrans.ranecu(7, 0, 200000)
    _, _, _, b = ctrans.ranecu(7, 1, 200000)
    for x in (a, b):
        assert 0.0 < x.min() and x.max() < 1.0
        assert abs(x.mean() - 0.5) < 5 * math.sqrt(1 / 12 / len(x))
        assert abs(x.var() - 1 / 12) < 0.002
        counts, _ = np.histogram(x, 50, (0, 1))
        e = len(x) / 50
        assert np.sum((counts - e) ** 2 / e) < 100  # chi^2 with 49 dof, far tail
    r = np.corrcoef(a.astype(np.float64), b.astype(np.float64))[0, 1]
    assert abs(r) < 5 / math.sqrt(len(a))
    lag = np.corrcoef(a[:-1].astype(np.float64), a[1:].astype(np.float64))[0, 1]
    assert abs(lag) < 5 / math.sqrt(len(a))


def _frame(d):
    a = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(d, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    return e1, e2


@pytest.mark.parametrize("pole", [False, True])
def test_rotate_polar_cosine_and_rigid_azimuth(pole):
    rng = np.random.default_rng(3)
    for _ in range(200):
        if pole:
            d = np.array([1e-4 * rng.standard_normal(), 1e-4 * rng.standard_normal(), rng.choice([-1.0, 1.0])])
        else:
            d = rng.standard_normal(3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        ct = float(rng.uniform(-1, 1))
        e1, e2 = _frame(d.astype(np.float64))
        offsets = []
        for phi in np.linspace(0.0, 2 * math.pi, 9)[:-1] + 0.1:
            o = ctrans.rotate(d[0], d[1], d[2], ct, phi).astype(np.float64)
            assert abs(np.linalg.norm(o) - 1.0) < 1e-6
            # FP32: 1 - w^2 cancels as |w| -> 1, so the error grows as 1 / sqrt(1 - w^2);
            # for |w| > 0.99999 the rotation takes d as exactly +-z, off by its
            # tilt |(u, v)| (the MC-GPU convention), which bounds the error there
            t = math.sqrt(max(1e-8, 1.0 - float(d[2]) ** 2))
            tilt = math.hypot(float(d[0]), float(d[1]))
            tol = 2e-5 + 1.01 * tilt if pole else 2e-6 + 2e-7 / t
            assert abs(float(o @ d.astype(np.float64)) - ct) < tol
            st = math.sqrt(max(0.0, 1 - ct * ct))
            if st > 0.05:
                psi = math.atan2(float(o @ e2), float(o @ e1))
                offsets.append((psi, phi))
        if offsets:
            # the same offset psi -/+ phi for every phi: a rigid turn about d in
            # one sense of rotation, so a uniform phi gives a uniform azimuth
            psi, phi = np.array(offsets).T
            spread = min(np.ptp(np.unwrap((psi - phi) % (2 * math.pi))),
                         np.ptp(np.unwrap((psi + phi) % (2 * math.pi))))
            assert spread < (2e-2 if pole else 1e-4), offsets


def _phantom_params(n_bins=4):
    rng = np.random.default_rng(5)
    nx, ny, nz = 24, 20, 16
    dens = rng.uniform(0.2, 2.5, (nz, ny, nx)).astype(np.float32)
    labs = rng.integers(0, 3, (nz, ny, nx)).astype(np.uint8)
    p = MCParams()
    p.n_materials = 3
    p.n_bins = n_bins
    E = np.linspace(30.0, 110.0, n_bins)
    cdf = np.cumsum(np.arange(1, n_bins + 1, dtype=np.float64))
    cdf /= cdf[-1]
    for i in range(n_bins):
        p.energies_keV[i] = float(E[i])
        p.cdf[i] = float(cdf[i])
    for m, (pe, c, r) in enumerate([(0.01, 0.09, 0.005), (0.045, 0.085, 0.0125), (0.4, 0.075, 0.03)]):
        p.a_pe[m], p.a_compton[m], p.a_rayleigh[m] = pe, c, r
    p.k_pe, p.k_compton, p.k_rayleigh = 3.0, 0.3, 1.8
    p.max_density = float(dens.max())
    Wf = np.array([[1.0, 0, 0, -12.0], [0, 1.0, 0, -10.0], [0, 0, 1.0, -8.0]])
    Wi = np.array([[1.0, 0, 0, 12.0], [0, 1.0, 0, 10.0], [0, 0, 1.0, 8.0]])
    p.world_from_ijk[:] = list(Wf.reshape(12))
    p.ijk_from_world[:] = list(Wi.reshape(12))
    p.voxel_size[:] = [1.0, 1.0, 1.0]
    p.source_world[:] = [0.0, 0.0, -300.0]
    p.detector_origin[:] = [-160.0, -160.0, 200.0]
    p.detector_u[:] = [10.0, 0.0, 0.0]
    p.detector_v[:] = [0.0, 10.0, 0.0]
    p.detector_normal[:] = [0.0, 0.0, 1.0]
    p.det_width = p.det_height = 32
    p.cone_axis[:] = [0.0, 0.0, 1.0]
    p.cone_cos = 0.97
    p.max_interactions = 2  # histories are cut at their third interaction
    return dens, labs, p


def test_energy_balance_per_history_batch():
    dens, labs, p = _phantom_params()
    tally, counts, steps = ctrans.mc_tally(dens, labs, p, 100000, seed=9)
    out = tally["scored"] + tally["photoelectric"] + tally["compton_recoil"] + tally["cutoff"] + tally["escaped"]
    assert tally["emitted"] > 0
    assert abs(out - tally["emitted"]) <= 1e-9 * tally["emitted"], tally
    # every fate occurs in this scene
    assert all(tally[k] > 0 for k in ("scored", "photoelectric", "compton_recoil", "cutoff", "escaped")), tally
    assert np.all(counts > 0), counts
    # the images score exactly the tallied energy, split by class
    img, _, steps2 = ctrans.mc(dens, labs, p, 100000, seed=9, n_threads=4)
    assert steps2 == steps
    assert abs(img.sum() - tally["scored"]) <= 1e-9 * tally["scored"]


# ---- single scatter from a thin slab --------------------------------------------------------

ME_C2 = 510.99895


def _slab_params(kind, E0=60.0, mu0=0.02):
    """128 x 128 x 2 voxels of 1 mm (slab z in [-1, 1] mm), one material with one
    interaction kind; pencil beam along +z from z = -500 mm; detector plane
    z = 100 mm, 64 x 64 pixels of 3 mm centred on the beam."""
    p = MCParams()
    p.n_materials = 1
    p.n_bins = 1
    p.energies_keV[0] = E0
    p.cdf[0] = 1.0
    a = {"compton": (0.0, mu0, 0.0), "rayleigh": (0.0, 0.0, mu0)}[kind]
    p.a_pe[0], p.a_compton[0], p.a_rayleigh[0] = a
    p.k_pe, p.k_compton, p.k_rayleigh = 3.0, 0.3, 1.8
    p.max_density = 1.0
    Wf = np.array([[1.0, 0, 0, -63.5], [0, 1.0, 0, -63.5], [0, 0, 1.0, -0.5]])
    Wi = np.array([[1.0, 0, 0, 63.5], [0, 1.0, 0, 63.5], [0, 0, 1.0, 0.5]])
    p.world_from_ijk[:] = list(Wf.reshape(12))
    p.ijk_from_world[:] = list(Wi.reshape(12))
    p.voxel_size[:] = [1.0, 1.0, 1.0]
    p.source_world[:] = [0.0, 0.0, -500.0]
    p.detector_origin[:] = [-96.0, -96.0, 100.0]
    p.detector_u[:] = [3.0, 0.0, 0.0]
    p.detector_v[:] = [0.0, 3.0, 0.0]
    p.detector_normal[:] = [0.0, 0.0, 1.0]
    p.det_width = p.det_height = 64
    p.cone_axis[:] = [0.0, 0.0, 1.0]
    p.cone_cos = 1.0
    p.max_interactions = 64
    dens = np.ones((2, 128, 128), np.float32)
    labs = np.zeros((2, 128, 128), np.uint8)
    return p, dens, labs


def _expected_single(kind, p, n_hist, E0=60.0, mu0=0.02, kexp=None):
    """Expected energy per pixel of the single-scatter class: entry at z = -1,
    first interaction density mu(E0) exp(-mu(E0)(z + 1)), angular density
    (1/sigma) dsigma/dOmega into the pixel's solid angle seen from the
    interaction point, energy E'(theta), and survival exp(-mu(E') (1 - z)/cos)
    to the slab's far face."""
    kexp = {"compton": p.k_compton, "rayleigh": p.k_rayleigh}[kind]

    def mu(E):
        return mu0 * (E / 60.0) ** (-kexp)

    k = E0 / ME_C2

    def eprime(c):
        return E0 / (1.0 + k * (1.0 - c)) if kind == "compton" else E0 + 0.0 * c

    def ang(c):
        if kind == "compton":
            P = 1.0 / (1.0 + k * (1.0 - c))
            return P * P * (P + 1.0 / P - (1.0 - c * c))
        return 1.0 + c * c

    cs = np.linspace(-1, 1, 200001)
    norm = 2 * math.pi * np.trapezoid(ang(cs), cs)  # integral over the sphere
    H = W = 64
    sub = 4
    # Gauss-Legendre nodes over the slab depth z in [-1, 1]
    zn, zw = np.polynomial.legendre.leggauss(12)
    out = np.zeros((H, W))
    out2 = np.zeros((H, W))  # the same with E'^2: sum of squared energies
    offs = (np.arange(sub) + 0.5) / sub
    gu, gv = np.meshgrid(offs, offs, indexing="xy")
    for z, wz in zip(zn, zw):
        dens_z = mu(E0) * math.exp(-mu(E0) * (z + 1.0))  # interaction density at depth z
        for iv in range(H):
            yv = -96.0 + 3.0 * (iv + gv)
            for iu in range(W):
                xu = -96.0 + 3.0 * (iu + gu)
                dx, dy, dzd = xu, yv, 100.0 - z
                r2 = dx * dx + dy * dy + dzd * dzd
                r = np.sqrt(r2)
                c = dzd / r
                dA = (3.0 / sub) ** 2
                dOmega = dA * c / r2
                e1 = eprime(c)
                surv = np.exp(-mu(e1) * (1.0 - z) / c)
                w = wz * dens_z * ang(c) / norm * dOmega * surv
                out[iv, iu] += np.sum(w * e1)
                out2[iv, iu] += np.sum(w * e1 * e1)
    return n_hist * out, n_hist * out2


@pytest.mark.parametrize("kind,cls", [("compton", 1), ("rayleigh", 2)])
def test_single_scatter_angular_distribution(kind, cls):
    n_hist = 3_000_000
    p, dens, labs = _slab_params(kind)
    img, img2, _ = ctrans.mc(dens, labs, p, n_hist, seed=21, n_threads=8)
    got, var = img[cls], img2[cls]
    expect, expect2 = _expected_single(kind, p, n_hist)
    # rings about the beam axis (radius in pixels from the detector centre)
    yy, xx = np.mgrid[0:64, 0:64]
    rad = np.hypot(xx + 0.5 - 32, yy + 0.5 - 32)
    edges = np.array([0, 2, 4, 6, 8, 11, 14, 17, 20, 24, 28, 32, 38, 46])
    z = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (rad >= lo) & (rad < hi)
        z.append((got[sel].sum() - expect[sel].sum()) / math.sqrt(var[sel].sum()))
        # mean scored energy in the ring, sum(E^2) / sum(E): E'(theta) of the
        # Compton relation (E0 for Rayleigh), nearly free of counting noise
        mean_got = var[sel].sum() / got[sel].sum()
        mean_want = expect2[sel].sum() / expect[sel].sum()
        assert abs(mean_got / mean_want - 1.0) < 2e-3, (lo, mean_got, mean_want)
    z = np.array(z)
    assert np.all(np.abs(z) < 4.5), z
    assert np.sum(z ** 2) / len(z) < 3.0, z
    # azimuthal uniformity: the four quadrants about the beam agree
    quads = [got[:32, :32].sum(), got[:32, 32:].sum(), got[32:, :32].sum(), got[32:, 32:].sum()]
    sq = math.sqrt(var.sum() / 4)
    assert max(quads) - min(quads) < 5 * sq, quads
    # total single-scatter energy on the detector
    tot = (got.sum() - expect.sum()) / math.sqrt(var.sum())
    assert abs(tot) < 4.5, tot
    # the other single class and the primary hit nothing but the beam pixel region
    other = img[3 - cls]
    assert other.sum() == 0.0
