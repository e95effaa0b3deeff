"""Adjoint (backward) kernel against a dense float64 grid_sample reference and
float64 central finite differences.

Bar: density and pose gradients within 1e-4 (relative L2) of the dense
reference's autograd gradients (same model, FP64); with one material and a
smooth density, FP64 central FD at eps = 1e-3 agrees to 1e-4 for the density
(along the gradient direction) and for the 6 pose DoF.  With material
boundaries the model is piecewise trilinear with large slope jumps at cell
faces, and the pose FD is shown to converge as eps shrinks (1e-3 at 1e-5).
"""
import numpy as np
import pytest
import torch

from deepdrr_amd import phantoms
from deepdrr_amd.autograd import DifferentiableProjector
from deepdrr_amd.spectral_data import synthetic_120kv
from tests.dense_ref import DenseDRR

pytestmark = pytest.mark.gpu

MATS = ["air", "bone", "soft tissue"]


def _scene(n=20, n_views=3, det=12, step=0.25, seed=0):
    rng = np.random.default_rng(seed)
    # smooth density (small gradient jumps at cell faces, so FP64 central FD
    # is not dominated by trilinear kinks) with a zero border (no jumps where
    # samples enter or leave the volume as the pose moves)
    i = np.arange(n - 4, dtype=np.float64)[:, None, None]
    j = np.arange(n - 4, dtype=np.float64)[None, :, None]
    k = np.arange(n - 4, dtype=np.float64)[None, None, :]
    ph = rng.uniform(0, 2 * np.pi, size=3)
    core = (1.0 + 0.4 * np.sin(0.35 * i + ph[0]) * np.cos(0.3 * j + ph[1])
            + 0.3 * np.sin(0.25 * k + ph[2])) * (np.sin(np.pi * (i + 1) / (n - 3))
                                                 * np.sin(np.pi * (j + 1) / (n - 3))
                                                 * np.sin(np.pi * (k + 1) / (n - 3))) ** 2
    dens = np.zeros((n, n, n), np.float32)
    dens[2:-2, 2:-2, 2:-2] = core.astype(np.float32)
    labels = np.zeros((n, n, n), np.uint8)
    labels[dens > 0.0] = 2
    labels[dens > 1.2] = 1
    origin = phantoms._centered_origin((n, n, n), (1, 1, 1))
    R = np.array([[1.0, 0, 0], [0, 0, 1.0], [0, -1.0, 0]])
    W = np.eye(4)
    W[:3, :3] = R
    W[:3, 3] = origin
    projs = [phantoms.pose_projection(det, det, 1.5 * 20 / det * 2, p, d, f)
             for p, d, f in phantoms.batch_poses(n_views, seed=seed, radius=2.0, cone_deg=40)]
    return dens, labels, W, projs, step


def _dense(dp, projs, theta0=np.zeros(6)):
    views = dp.views(projs, theta0)
    return DenseDRR(views, dp.world_from_ijk, dp.center, dp.step, dp.E, dp.pdf, dp.mu, dp.shape)


def _onehot(labels):
    return torch.tensor(np.stack([(labels == m).astype(np.float64) for m in range(len(MATS))]))


def test_adjoint_vs_dense_and_fd(gpu):
    dens, labels, W, projs, step = _scene()
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        rng = np.random.default_rng(1)
        H = W_ = projs[0].intrinsic.sensor_width
        gI = rng.standard_normal((len(projs), H, W_)).astype(np.float32)
        d_t = torch.tensor(dens, device=gpu, requires_grad=True)
        th_t = torch.zeros(6, device=gpu, requires_grad=True)
        print("adjoint: forward + backward on GPU", flush=True)
        I = dp(d_t, th_t, projs)
        (I * torch.tensor(gI, device=gpu)).sum().backward()
        gD = d_t.grad.detach().cpu().numpy().astype(np.float64)
        gth = th_t.grad.detach().cpu().numpy().astype(np.float64)

        ref = _dense(dp, projs)
        oh = _onehot(labels)
        d64 = torch.tensor(dens.astype(np.float64), requires_grad=True)
        th64 = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        I64 = ref.render(d64, oh, th64)
        np.testing.assert_allclose(I.detach().cpu().numpy(), I64.detach().numpy(), rtol=2e-5)
        L = (I64 * torch.tensor(gI, dtype=torch.float64)).sum()
        L.backward()
        gD_ref = d64.grad.numpy()
        gth_ref = th64.grad.numpy()
        rel = np.linalg.norm(gD - gD_ref) / np.linalg.norm(gD_ref)
        print(f"density grad rel L2 vs dense: {rel:.2e}; pose grad {gth} vs {gth_ref}")
        assert rel < 1e-4
        err_dense = np.linalg.norm(gth - gth_ref) / np.linalg.norm(gth_ref)
        print(f"pose grad rel L2 vs dense autograd: {err_dense:.2e}")
        assert err_dense < 1e-4

        # FP64 central finite differences on the dense model
        eps = 1e-3
        gIt = torch.tensor(gI, dtype=torch.float64)

        def loss(d, th):
            with torch.no_grad():
                return float((ref.render(d, oh, th) * gIt).sum())

        d0 = torch.tensor(dens.astype(np.float64))
        z = torch.zeros(6, dtype=torch.float64)
        # (a) along the reference gradient direction: the GPU adjoint's
        # directional derivative must match FP64 central FD
        v = torch.tensor(gD_ref / np.linalg.norm(gD_ref))
        fd = (loss(d0 + eps * v, z) - loss(d0 - eps * v, z)) / (2 * eps)
        an = float((gD * v.numpy()).sum())
        print(f"density FD (gradient direction) {fd:.8g} adjoint {an:.8g}")
        assert abs(an - fd) <= 1e-4 * abs(fd)
        # (b) along a random direction the FP64 dense autograd matches FD; the
        # projection of the FP32 adjoint on a random direction is ~sqrt(n) times
        # less accurate than its L2 error, which (a) and rel above bound
        v = torch.tensor(rng.standard_normal(dens.shape))
        fd = (loss(d0 + eps * v, z) - loss(d0 - eps * v, z)) / (2 * eps)
        assert abs(float((gD_ref * v.numpy()).sum()) - fd) <= 1e-6 * abs(fd)
        # Material boundaries make the model piecewise trilinear with large slope
        # jumps at cell faces; central FD converges to the derivative as eps
        # shrinks (the straddling fraction of samples is proportional to eps).
        errs = []
        for e_pose in (1e-3, 1e-4, 1e-5):
            fdp = np.zeros(6)
            for k in range(6):
                e = torch.zeros(6, dtype=torch.float64)
                e[k] = e_pose
                fdp[k] = (loss(d0, e) - loss(d0, -e)) / (2 * e_pose)
            errs.append(np.linalg.norm(gth - fdp) / np.linalg.norm(fdp))
            print(f"pose FD eps {e_pose:g}: {fdp} adjoint {gth}; relative L2 {errs[-1]:.2e}")
        assert errs[-1] < 1e-3 and errs[-1] < errs[0]
    finally:
        dp.close()


def test_adjoint_axis_aligned_rays(gpu):
    """An AP view on an 11x11 detector of a volume whose axes are world axes:
    the central ray runs exactly along a volume axis and the central row and
    column have a direction component exactly zero (the per-cell exit distance
    is infinite along those axes).  Density and pose gradients against the
    dense FP64 autograd gradients."""
    dens, labels, W, _, step = _scene()
    projs = [phantoms.ap_device(11, 11, 3.0).get_camera_projection()]
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        gI = np.random.default_rng(2).standard_normal((1, 11, 11)).astype(np.float32)
        d_t = torch.tensor(dens, device=gpu, requires_grad=True)
        th_t = torch.zeros(6, device=gpu, requires_grad=True)
        print("adjoint, axis-aligned rays: forward + backward on GPU", flush=True)
        I = dp(d_t, th_t, projs)
        (I * torch.tensor(gI, device=gpu)).sum().backward()
        gD = d_t.grad.detach().cpu().numpy().astype(np.float64)
        gth = th_t.grad.detach().cpu().numpy().astype(np.float64)
        ref = _dense(dp, projs)
        d64 = torch.tensor(dens.astype(np.float64), requires_grad=True)
        th64 = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        I64 = ref.render(d64, _onehot(labels), th64)
        np.testing.assert_allclose(I.detach().cpu().numpy(), I64.detach().numpy(), rtol=2e-5)
        (I64 * torch.tensor(gI, dtype=torch.float64)).sum().backward()
        rel = np.linalg.norm(gD - d64.grad.numpy()) / np.linalg.norm(d64.grad.numpy())
        err = np.linalg.norm(gth - th64.grad.numpy()) / np.linalg.norm(th64.grad.numpy())
        print(f"axis-aligned: density grad rel L2 {rel:.2e}, pose grad rel L2 {err:.2e}")
        assert rel < 1e-4 and err < 1e-4
    finally:
        dp.close()


def test_adjoint_far_source(gpu):
    """A source 2 km from the volume: each FP32 `alpha += step` moves alpha by
    0.125 mm at this distance, not 0.1, so a cell's sample count must use the
    actual increment.  Density and pose gradients against the dense FP64
    autograd gradients (the dense model samples the same FP32 alpha sequence)."""
    dens, labels, W, _, step = _scene()
    projs = [phantoms.pose_projection(12, 12, 4.0, p, d, f, sdd=4.0e6)
             for p, d, f in phantoms.batch_poses(2, seed=5, radius=2.0, cone_deg=40)]
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        gI = np.random.default_rng(3).standard_normal((2, 12, 12)).astype(np.float32)
        d_t = torch.tensor(dens, device=gpu, requires_grad=True)
        th_t = torch.zeros(6, device=gpu, requires_grad=True)
        print("adjoint, far source: forward + backward on GPU", flush=True)
        I = dp(d_t, th_t, projs)
        (I * torch.tensor(gI, device=gpu)).sum().backward()
        gD = d_t.grad.detach().cpu().numpy().astype(np.float64)
        gth = th_t.grad.detach().cpu().numpy().astype(np.float64)
        ref = _dense(dp, projs)
        d64 = torch.tensor(dens.astype(np.float64), requires_grad=True)
        th64 = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        I64 = ref.render(d64, _onehot(labels), th64)
        np.testing.assert_allclose(I.detach().cpu().numpy(), I64.detach().numpy(), rtol=2e-5)
        (I64 * torch.tensor(gI, dtype=torch.float64)).sum().backward()
        rel = np.linalg.norm(gD - d64.grad.numpy()) / np.linalg.norm(d64.grad.numpy())
        err = np.linalg.norm(gth - th64.grad.numpy()) / np.linalg.norm(th64.grad.numpy())
        print(f"far source: density grad rel L2 {rel:.2e}, pose grad rel L2 {err:.2e}")
        assert rel < 1e-4 and err < 1e-4
    finally:
        dp.close()


def test_adjoint_density_only_and_no_pose(gpu):
    """theta without grad: only the density gradient is produced and it matches."""
    dens, labels, W, projs, step = _scene(n=16, n_views=2, det=10, seed=3)
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        d_t = torch.tensor(dens, device=gpu, requires_grad=True)
        th = torch.zeros(6, device=gpu)
        I = dp(d_t, th, projs)
        I.sum().backward()
        ref = _dense(dp, projs)
        d64 = torch.tensor(dens.astype(np.float64), requires_grad=True)
        ref.render(d64, _onehot(labels), torch.zeros(6, dtype=torch.float64)).sum().backward()
        g = d_t.grad.cpu().numpy()
        rel = np.linalg.norm(g - d64.grad.numpy()) / np.linalg.norm(d64.grad.numpy())
        assert rel < 1e-4
    finally:
        dp.close()


def test_adjoint_fd_single_material_smooth(gpu):
    """One material and a smooth density.  The density derivative is exact under
    FD (the model is linear in density): 1e-4 at eps = 1e-3.  In the pose the
    model is piecewise trilinear (slope jumps at cell faces), so the adjoint is
    checked against the FP64 autograd of the same model to 1e-4 and FD is
    required to converge to it: error at eps = 1e-5 below 1e-3 and below the
    error at eps = 1e-3."""
    dens, labels, W, projs, step = _scene(n=20, n_views=3, det=12, seed=2)
    labels = np.where(dens > 0, 2, 0).astype(np.uint8)  # soft tissue inside, air border
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        rng = np.random.default_rng(4)
        gI = rng.standard_normal((len(projs), 12, 12)).astype(np.float32)
        d_t = torch.tensor(dens, device=gpu, requires_grad=True)
        th_t = torch.zeros(6, device=gpu, requires_grad=True)
        (dp(d_t, th_t, projs) * torch.tensor(gI, device=gpu)).sum().backward()
        gD = d_t.grad.cpu().numpy().astype(np.float64)
        gth = th_t.grad.cpu().numpy().astype(np.float64)
        ref = _dense(dp, projs)
        oh = _onehot(labels)
        gIt = torch.tensor(gI, dtype=torch.float64)

        def loss(d, th):
            with torch.no_grad():
                return float((ref.render(d, oh, th) * gIt).sum())

        eps = 1e-3
        d0 = torch.tensor(dens.astype(np.float64))
        z = torch.zeros(6, dtype=torch.float64)
        v = torch.tensor(gD / np.linalg.norm(gD))
        fd = (loss(d0 + eps * v, z) - loss(d0 - eps * v, z)) / (2 * eps)
        an = float((gD * v.numpy()).sum())
        assert abs(an - fd) <= 1e-4 * abs(fd)
        th64 = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        (ref.render(d0, oh, th64) * gIt).sum().backward()
        err_dense = np.linalg.norm(gth - th64.grad.numpy()) / np.linalg.norm(th64.grad.numpy())
        assert err_dense < 1e-4
        errs = []
        for e_pose in (1e-3, 1e-5):
            fdp = np.zeros(6)
            for k in range(6):
                e = torch.zeros(6, dtype=torch.float64)
                e[k] = e_pose
                fdp[k] = (loss(d0, e) - loss(d0, -e)) / (2 * e_pose)
            errs.append(np.linalg.norm(gth - fdp) / np.linalg.norm(fdp))
        print(f"single material: density FD {fd:.8g} vs {an:.8g}; pose vs dense {err_dense:.2e}; "
              f"pose vs FD at 1e-3 / 1e-5: {errs[0]:.2e} / {errs[1]:.2e}")
        assert errs[1] < 1e-3 and errs[1] < errs[0]
    finally:
        dp.close()


@pytest.mark.parametrize("n,det,n_views,seed", [(18, 14, 4, 5), (40, 32, 3, 9)])
def test_adjoint_jobs_equals_sum_of_single_view_adjoints(gpu, n, det, n_views, seed):
    """drr_adjoint_jobs (one persistent launch) == sum over views of drr_adjoint,
    up to float atomic ordering; drr_project_jobs == drr_project bit for bit.
    The 40^3 case is wider than the LDS box, so oblique beams leave it and their
    outer corners take the direct global path."""
    import ctypes as C

    from deepdrr_amd import _abi
    from deepdrr_amd._abi_types import Job

    dens, labels, W, projs, step = _scene(n=n, n_views=n_views, det=det, seed=seed)
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        views = dp.views(projs, np.zeros(6))
        H, Wd = views[0][0].height, views[0][0].width
        d_kji = torch.tensor(np.ascontiguousarray(np.moveaxis(dens, [0, 1, 2], [2, 1, 0])), device=gpu)
        h = dp.upload(d_kji)
        varr = (C.c_void_p * 1)(h.value)
        jobs = (Job * len(views))()
        for i, (v, _) in enumerate(views):
            jobs[i].view = v
            jobs[i].volume_index = 0
            jobs[i].out_offset = i * H * Wd
        I_j = torch.empty((len(views), H, Wd), device=gpu)
        P_j = torch.empty_like(I_j)
        AD_j = torch.empty((len(views), len(MATS), H, Wd), device=gpu)
        _abi.call("drr_project_jobs", dp.ctx, varr, 1, jobs, len(views), _abi.ptr(I_j), _abi.ptr(P_j),
                  _abi.ptr(AD_j), _abi.stream_ptr())
        rng = np.random.default_rng(2)
        gI = torch.tensor(rng.standard_normal((len(views), H, Wd)).astype(np.float32), device=gpu)
        gPp = torch.tensor(rng.standard_normal((len(views), H, Wd)).astype(np.float32), device=gpu)
        c = np.asarray(dp.center, np.float32)
        src = np.ascontiguousarray(np.stack([s for _, s in views]), np.float32)
        gD_j = torch.zeros_like(d_kji)
        gT_j = torch.zeros(6, device=gpu)
        _abi.call("drr_adjoint_jobs", dp.ctx, h, jobs, len(views), _abi.ptr(gI), _abi.ptr(gPp),
                  src.ctypes.data, c.ctypes.data, None, _abi.ptr(gD_j), _abi.ptr(gT_j), _abi.stream_ptr())
        # reusing the forward's area densities gives the same gradients
        gD_r = torch.zeros_like(d_kji)
        gT_r = torch.zeros(6, device=gpu)
        _abi.call("drr_adjoint_jobs", dp.ctx, h, jobs, len(views), _abi.ptr(gI), _abi.ptr(gPp),
                  src.ctypes.data, c.ctypes.data, _abi.ptr(AD_j), _abi.ptr(gD_r), _abi.ptr(gT_r),
                  _abi.stream_ptr())
        gT_p = torch.zeros(6, device=gpu)
        _abi.call("drr_adjoint_jobs", dp.ctx, h, jobs, len(views), _abi.ptr(gI), _abi.ptr(gPp),
                  src.ctypes.data, c.ctypes.data, _abi.ptr(AD_j), None, _abi.ptr(gT_p), _abi.stream_ptr())
        gD_s = torch.zeros_like(d_kji)
        gT_s = torch.zeros(6, device=gpu)
        for i, (v, s) in enumerate(views):
            I = torch.empty((H, Wd), device=gpu)
            Pp = torch.empty_like(I)
            AD = torch.empty((len(MATS), H, Wd), device=gpu)
            _abi.call("drr_project", dp.ctx, varr, C.byref(v), _abi.ptr(I), _abi.ptr(Pp), _abi.ptr(AD), None,
                      _abi.stream_ptr())
            torch.cuda.synchronize()
            assert torch.equal(I, I_j[i]) and torch.equal(Pp, P_j[i]), f"view {i}"
            assert torch.equal(AD, AD_j[i]), f"area densities of view {i}"
            _abi.call("drr_adjoint", dp.ctx, h, C.byref(v), _abi.ptr(gI[i]), _abi.ptr(gPp[i]),
                      s.ctypes.data, c.ctypes.data, _abi.ptr(gD_s), _abi.ptr(gT_s), _abi.stream_ptr())
        torch.cuda.synchronize()
        _abi.lib().drr_volume_destroy(h)
        a, b = gD_j.cpu().numpy(), gD_s.cpu().numpy()
        assert np.abs(a).max() > 0
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
        r_ = gD_r.cpu().numpy()
        assert np.linalg.norm(r_ - a) <= 1e-6 * np.linalg.norm(a)
        # pose alone (no density gradient): the same pose gradient
        np.testing.assert_allclose(gT_p.cpu().numpy(), gT_r.cpu().numpy(), rtol=1e-5,
                                   atol=1e-6 * float(np.abs(gT_r.cpu().numpy()).max()))
        np.testing.assert_allclose(gT_r.cpu().numpy(), gT_j.cpu().numpy(), rtol=1e-5,
                                   atol=1e-6 * float(np.abs(gT_j.cpu().numpy()).max()))
        np.testing.assert_allclose(gT_j.cpu().numpy(), gT_s.cpu().numpy(), rtol=1e-4,
                                   atol=1e-5 * float(np.abs(gT_s.cpu().numpy()).max()))
    finally:
        dp.close()


ALL_MATS = ["air", "bone", "soft tissue", "lung", "titanium", "muscle", "blood", "teflon", "iron", "lead",
            "copper", "polyethylene", "concrete", "Al", "Ca", "W"]


@pytest.mark.parametrize("k", [2, 4, 6, 8, 16])
def test_adjoint_material_counts(gpu, k):
    """k materials (the adjoint kernels are instantiated for 1 to 16) labelled in
    2x2x2 blocks, so that half the cells mix labels and the pose gradient's
    label term is exercised: density and pose gradients against dense FP64
    autograd of the same model within 1e-4 (relative L2)."""
    dens, _, W, projs, step = _scene(seed=k)
    rng = np.random.default_rng(100 + k)
    n = dens.shape[0]
    coarse = rng.integers(0, k, (n // 2,) * 3)
    labels = np.repeat(np.repeat(np.repeat(coarse, 2, 0), 2, 1), 2, 2).astype(np.uint8)
    mats = ALL_MATS[:k]
    dp = DifferentiableProjector(labels, mats, W, step=step, spectrum=synthetic_120kv())
    try:
        gI = rng.standard_normal((len(projs),) + (projs[0].intrinsic.sensor_height,
                                                  projs[0].intrinsic.sensor_width)).astype(np.float32)
        d_t = torch.tensor(dens, device=gpu, requires_grad=True)
        th_t = torch.zeros(6, device=gpu, requires_grad=True)
        print(f"adjoint with {k} materials", flush=True)
        I = dp(d_t, th_t, projs)
        (I * torch.tensor(gI, device=gpu)).sum().backward()
        gD = d_t.grad.detach().cpu().numpy().astype(np.float64)
        gth = th_t.grad.detach().cpu().numpy().astype(np.float64)
        ref = _dense(dp, projs)
        oh = torch.tensor(np.stack([(labels == m).astype(np.float64) for m in range(k)]))
        d64 = torch.tensor(dens.astype(np.float64), requires_grad=True)
        th64 = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        I64 = ref.render(d64, oh, th64)
        np.testing.assert_allclose(I.detach().cpu().numpy(), I64.detach().numpy(), rtol=2e-5)
        (I64 * torch.tensor(gI, dtype=torch.float64)).sum().backward()
        rel = np.linalg.norm(gD - d64.grad.numpy()) / np.linalg.norm(d64.grad.numpy())
        err = np.linalg.norm(gth - th64.grad.numpy()) / np.linalg.norm(th64.grad.numpy())
        print(f"{k} materials: density grad rel L2 {rel:.2e}, pose grad rel L2 {err:.2e}")
        assert rel < 1e-4
        assert err < 1e-4
    finally:
        dp.close()


def test_adjoint_air_outside(gpu):
    """The adjoint kernels with the air outside the volume attenuating (the
    scene's attenuate_outside, reachable through the C ABI): gradients against
    dense FP64 autograd of the same model, in which that outside air (read from
    the forward's area densities) is a constant, as the adjoint treats it."""
    import ctypes as C

    from deepdrr_amd import _abi
    from deepdrr_amd._abi_types import Job, Scene

    dens, labels, W, projs, step = _scene(seed=7)
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        views = dp.views(projs, np.zeros(6))
        H, Wd = views[0][0].height, views[0][0].width
        d_kji = torch.tensor(np.ascontiguousarray(np.moveaxis(dens, [0, 1, 2], [2, 1, 0])), device=gpu)
        h = dp.upload(d_kji)
        varr = (C.c_void_p * 1)(h.value)
        jobs = (Job * len(views))()
        for i, (v, _) in enumerate(views):
            jobs[i].view = v
            jobs[i].volume_index = 0
            jobs[i].out_offset = i * H * Wd

        def area_densities():
            I_j = torch.empty((len(views), H, Wd), device=gpu)
            AD_j = torch.empty((len(views), len(MATS), H, Wd), device=gpu)
            _abi.call("drr_project_jobs", dp.ctx, varr, 1, jobs, len(views), _abi.ptr(I_j),
                      _abi.ptr(torch.empty_like(I_j)), _abi.ptr(AD_j), _abi.stream_ptr())
            torch.cuda.synchronize()
            return AD_j.cpu().numpy().astype(np.float64)

        ad0 = area_densities()
        sc = Scene(step=step, n_materials=len(MATS), attenuate_outside=1, air_index=MATS.index("air"))
        _abi.call("drr_ctx_set_scene", dp.ctx, C.byref(sc))
        ad1 = area_densities()
        _abi.lib().drr_volume_destroy(h)
        extra = ad1 - ad0
        air = MATS.index("air")
        assert np.all(extra[:, [m for m in range(len(MATS)) if m != air]] == 0.0)
        # (minAlpha/step + (ray_length - maxAlpha)/step) * 0.1129 plus the steps
        # before the volume: with unit-depth rays (ray_length ~ 1 mm) the
        # reference's suffix is negative, so only its presence is asserted
        assert np.abs(extra[:, air]).max() > 0.0
        extra_t = [torch.tensor(e.reshape(len(MATS), H * Wd)) for e in extra]

        rng = np.random.default_rng(3)
        gI = rng.standard_normal((len(projs), H, Wd)).astype(np.float32)
        d_t = torch.tensor(dens, device=gpu, requires_grad=True)
        th_t = torch.zeros(6, device=gpu, requires_grad=True)
        print("adjoint with air outside the volume", flush=True)
        I = dp(d_t, th_t, projs)
        (I * torch.tensor(gI, device=gpu)).sum().backward()
        gD = d_t.grad.detach().cpu().numpy().astype(np.float64)
        gth = th_t.grad.detach().cpu().numpy().astype(np.float64)
        ref = _dense(dp, projs)
        d64 = torch.tensor(dens.astype(np.float64), requires_grad=True)
        th64 = torch.zeros(6, dtype=torch.float64, requires_grad=True)
        I64 = ref.render(d64, _onehot(labels), th64, extra_ad=extra_t)
        np.testing.assert_allclose(I.detach().cpu().numpy(), I64.detach().numpy(), rtol=2e-5)
        (I64 * torch.tensor(gI, dtype=torch.float64)).sum().backward()
        rel = np.linalg.norm(gD - d64.grad.numpy()) / np.linalg.norm(d64.grad.numpy())
        err = np.linalg.norm(gth - th64.grad.numpy()) / np.linalg.norm(th64.grad.numpy())
        print(f"air outside: density grad rel L2 {rel:.2e}, pose grad rel L2 {err:.2e}")
        assert rel < 1e-4
        assert err < 1e-4
    finally:
        dp.close()


def test_adjoint_errors(gpu):
    """Invalid adjoint calls are refused on the host with a message, before any
    launch, and leave the gradient buffers untouched."""
    import ctypes as C

    from deepdrr_amd import _abi
    from deepdrr_amd._abi_types import Job

    dens, labels, W, projs, step = _scene(n=12, n_views=1, det=8)
    dp = DifferentiableProjector(labels, MATS, W, step=step, spectrum=synthetic_120kv())
    try:
        (v, s), = dp.views(projs, np.zeros(6))
        d_kji = torch.tensor(np.ascontiguousarray(np.moveaxis(dens, [0, 1, 2], [2, 1, 0])), device=gpu)
        h = dp.upload(d_kji)
        gI = torch.zeros((v.height, v.width), device=gpu)
        gD = torch.zeros_like(d_kji)
        gT = torch.zeros(6, device=gpu)
        c = np.asarray(dp.center, np.float32)
        lib, st = _abi.lib(), _abi.stream_ptr()
        J = (Job * 1)()
        J[0].view = v
        J[0].volume_index = 0
        J[0].out_offset = 0

        def jobs(n, gi=gI, gd=gD, gt=gT, src=s.ctypes.data):
            return lib.drr_adjoint_jobs(dp.ctx, h, J, n, None if gi is None else _abi.ptr(gi), None, src,
                                        c.ctypes.data, None, None if gd is None else _abi.ptr(gd),
                                        None if gt is None else _abi.ptr(gt), st)

        assert jobs(0) != 0 and "no jobs" in _abi.last_error()
        J[0].volume_index = 1
        assert jobs(1) != 0 and "volume 0" in _abi.last_error()
        J[0].volume_index = 0
        assert jobs(1, gd=None, gt=None) != 0 and "neither" in _abi.last_error()
        assert jobs(1, gi=None) != 0 and "no output gradient" in _abi.last_error()
        assert jobs(1, src=None) != 0 and "source_world" in _abi.last_error()
        v2 = type(v).from_buffer_copy(v)
        v2.n_volumes = 2
        rc = lib.drr_adjoint(dp.ctx, h, C.byref(v2), _abi.ptr(gI), None, s.ctypes.data, c.ctypes.data,
                             _abi.ptr(gD), None, st)
        assert rc != 0 and "one volume" in _abi.last_error()
        torch.cuda.synchronize()
        assert float(gD.abs().sum()) == 0.0 and float(gT.abs().sum()) == 0.0
        lib.drr_volume_destroy(h)
    finally:
        dp.close()
