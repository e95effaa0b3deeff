"""GPU Woodcock Monte Carlo against the CPU counterpart at matched histories.

Both use the same RANECU stream per history, so most histories are identical;
the bar is statistical: per-pixel |GPU - CPU| <= 3 sigma (sigma from the
per-pixel sums of squares of both runs) for at least 99% of the pixels with
signal, in every class image [primary, single Compton, single Rayleigh,
multiple], and the same number of Woodcock flights within 0.1%.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from deepdrr_amd import _abi, mc, phantoms, workload
from deepdrr_amd import projector as P
from deepdrr_amd._abi_types import Scene

pytestmark = pytest.mark.gpu


def _phantom(n=24):
    """Tissue block with a lung-like cavity, a bone cylinder and a titanium rod."""
    rng = np.random.default_rng(2)
    hu = (40.0 + 30.0 * rng.standard_normal((n, n, n))).astype(np.float32)
    i, j, k = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    hu[(i - n / 3) ** 2 + (j - n / 2) ** 2 < (n / 6) ** 2] = -800.0
    hu[(i - 2 * n / 3) ** 2 + (j - n / 2) ** 2 < (n / 8) ** 2] = 1200.0
    hu[(i - n / 2) ** 2 + (k - n / 2) ** 2 < 2.0] = 5900.0
    hu[:, :, :1] = -1000.0
    vol = phantoms.Volume.from_hu(hu, origin=phantoms._centered_origin((n, n, n), (2, 2, 2)), spacing=(2, 2, 2),
                                  anatomical_coordinate_system="LPS",
                                  material_classes=phantoms.FIVE_MATERIAL_CLASSES)
    return vol


def test_mc_matches_cpu(gpu):
    from oracle import ctrans
    from deepdrr_amd.spectral_data import synthetic_120kv

    vol = _phantom()
    mats = sorted(vol.materials[0].keys())
    dens, labs = P.volume_arrays(vol, mats)
    dev = phantoms.ap_device(32, 32, 4.0, sdd=600.0)
    proj = dev.get_camera_projection()
    p = mc.make_params(mats, synthetic_120kv(), proj, 600.0, vol.world_from_IJK.data,
                       max_density=float(dens.max()))
    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    try:
        sc = Scene(step=0.1, n_materials=len(mats), attenuate_outside=0, air_index=0)
        _abi.call("drr_ctx_set_scene", ctx, C.byref(sc))
        h = C.c_void_p()
        nz, ny, nx = dens.shape
        _abi.call("drr_volume_create", ctx, dens.ctypes.data, labs.ctypes.data, nx, ny, nz, 0, C.byref(h))
        n_hist = 300000
        print(f"mc: {n_hist} histories on GPU", flush=True)
        run = mc.MonteCarlo(ctx, h, p)
        run.run(n_hist, seed=12345)
        g, g2, gsteps = run.results()
        gbsteps = run.brick_steps()
        cb = []
        c, c2, csteps = ctrans.mc(dens, labs, p, n_hist, seed=12345, brick_steps=cb)
        print(f"steps gpu {gsteps} cpu {csteps}; brick steps gpu {gbsteps} cpu {cb[0]}; "
              f"energy per class gpu {g.sum((1, 2))} cpu {c.sum((1, 2))}")
        assert abs(gsteps - csteps) <= 1e-3 * csteps
        # the 24^3 volume has partial bricks on every axis: both walks cross the
        # same bricks
        assert abs(gbsteps - cb[0]) <= 1e-3 * cb[0]
        for k, name in enumerate(mc.CLASSES):
            sig = np.sqrt(g2[k] + c2[k])
            sel = sig > 0
            z = (g[k][sel] - c[k][sel]) / sig[sel]
            frac = float(np.mean(np.abs(z) <= 3.0))
            print(f"{name}: {sel.sum()} pixels, within 3 sigma {frac:.4f}")
            assert frac >= 0.99, name
        # scatter is present (the 128 mm detector subtends little of the 4 pi
        # the 48 mm object scatters into, so its share is well under 10%)
        assert 0.002 < g[1:].sum() / g.sum() < 0.5
        _abi.lib().drr_volume_destroy(h)
    finally:
        _abi.lib().drr_ctx_destroy(ctx)


def test_projector_with_scatter(gpu):
    """mc_scatter_histories > 0 adds the MC scatter estimate to the images."""
    from deepdrr_amd import Projector
    from deepdrr_amd.spectral_data import synthetic_120kv

    vol = _phantom(16)
    dev = phantoms.ap_device(16, 16, 4.0, sdd=600.0)
    with Projector(vol, device=dev, spectrum=synthetic_120kv(), neglog=False) as p0:
        base = p0()
    with Projector(vol, device=dev, spectrum=synthetic_120kv(), neglog=False, mc_scatter_histories=200000) as p1:
        withs = p1()
    assert withs.shape == base.shape
    assert np.all(withs >= base - 1e-3) and np.mean(withs - base) > 0


def test_mc_parameter_errors(gpu):
    """Invalid transport parameters are refused on the host with a message, and
    zero histories is a no-op."""
    from deepdrr_amd.spectral_data import synthetic_120kv

    vol = _phantom(8)
    mats = sorted(vol.materials[0].keys())
    dens, labs = P.volume_arrays(vol, mats)
    proj = phantoms.ap_device(8, 8, 4.0, sdd=600.0).get_camera_projection()
    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    try:
        sc = Scene(step=0.1, n_materials=len(mats), attenuate_outside=0, air_index=0)
        _abi.call("drr_ctx_set_scene", ctx, C.byref(sc))
        h = C.c_void_p()
        nz, ny, nx = dens.shape
        _abi.call("drr_volume_create", ctx, dens.ctypes.data, labs.ctypes.data, nx, ny, nz, 0, C.byref(h))
        img = torch.zeros((4, 8, 8), device=gpu)

        def run(p, n=1000):
            return _abi.lib().drr_mc_scatter(ctx, h, C.byref(p), 0, n, C.c_uint64(1), _abi.ptr(img), None, None, None,
                                             _abi.stream_ptr())

        def params(**kw):
            p = mc.make_params(mats, synthetic_120kv(), proj, 600.0, vol.world_from_IJK.data)
            for k, v in kw.items():
                setattr(p, k, v)
            return p

        assert run(params(cone_cos=1.0)) != 0 and "cone_cos" in _abi.last_error()
        assert run(params(max_interactions=0)) != 0 and "max_interactions" in _abi.last_error()
        assert run(params(n_materials=len(mats) - 1)) != 0 and "label" in _abi.last_error()
        bad_cdf = params()
        bad_cdf.cdf[bad_cdf.n_bins - 1] = 0.5
        assert run(bad_cdf) != 0 and "cdf" in _abi.last_error()
        assert run(params(), n=0) == 0
        torch.cuda.synchronize()
        assert float(img.abs().sum()) == 0.0
        _abi.lib().drr_volume_destroy(h)
    finally:
        _abi.lib().drr_ctx_destroy(ctx)
