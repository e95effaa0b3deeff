"""Dense differentiable reference of the single-volume DRR (float64, CPU torch).

Same model as the kernels (project_kernel.cu semantics): samples at the
kernel's FP32 lattice alpha_t (sequential ``alpha += step`` from minAlpha,
half weight at t = 0 and t = N-1, only steps inside [minAlpha_vol,
maxAlpha_vol]), positions p = IJK(x) - 1, density and one-hot material
fractions trilinear with clamp-to-edge -- implemented with
``torch.nn.functional.grid_sample`` (align_corners=True, border padding) so
that autograd gives d/d(density) and d/d(pose).  The sample set is fixed by
the unperturbed geometry (boundary terms are not differentiated, as in the
adjoint kernel).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

F32 = np.float32


def _ray_setup(view, step, pixels=None):
    """Per ray: unit world direction (f32), minAlpha, the volume's [mn, mx] and
    the step count N; rays of every pixel [v][u], or of the flat pixel indices
    v * W + u in ``pixels`` (each computed as in the full view)."""
    W = np.array(view.world_from_index[:], F32)
    w, h = view.width, view.height
    if pixels is None:
        vv, uu = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    else:
        pix = np.asarray(pixels, np.int64)
        vv, uu = pix // w, pix % w
    u = uu.reshape(-1).astype(F32) + F32(0.5)
    v = vv.reshape(-1).astype(F32) + F32(0.5)
    rx = (u * W[0] + v * W[1]) + W[2]
    ry = (u * W[3] + v * W[4]) + W[5]
    rz = (u * W[6] + v * W[7]) + W[8]
    L = np.sqrt((rx * rx + ry * ry) + rz * rz).astype(F32)
    inv = (F32(1) / L).astype(F32)
    r = np.stack([rx * inv, ry * inv, rz * inv], 1).astype(F32)
    vol = view.vol[0]
    A = np.array(vol.ijk_from_world[:], F32).reshape(3, 4)
    s = np.array(vol.source_ijk[:], F32)
    lo = np.array(vol.bounds_lo[:], F32)
    hi = np.array(vol.bounds_hi[:], F32)
    ri = np.stack([((A[a, 0] * r[:, 0] + A[a, 1] * r[:, 1]) + A[a, 2] * r[:, 2]) + A[a, 3] * F32(0)
                   for a in range(3)], 1).astype(F32)
    mrl = F32(view.max_ray_length)
    mn = np.zeros(len(u), F32)
    mx = np.full(len(u), mrl if mrl > 0 else np.inf, F32)
    # a zero direction component divides to +-inf, which leaves that axis
    # unconstrained when the source lies between its planes and empties the
    # ray otherwise, as the reference's skip of such components does
    with np.errstate(divide="ignore", invalid="ignore"):
        for a in range(3):
            reci = (F32(1) / ri[:, a]).astype(F32)
            a0 = ((lo[a] - s[a]) * reci).astype(F32)
            a1 = ((hi[a] - s[a]) * reci).astype(F32)
            mn = np.fmax(mn, np.fmin(a0, a1))
            mx = np.fmin(mx, np.fmax(a0, a1))
    minA = np.fmin(L, mn)
    maxA = np.fmax(F32(0), mx)
    N = np.ceil(((maxA - minA) / F32(step)).astype(F32)).astype(np.int64)
    return r, minA, mn, mx, N


def _walk(minA, mn, mx, N, step):
    """The kernel's sequential FP32 alpha lattice of every ray at once: ray
    index, alpha and end-sample weight of each in-range sample."""
    rays, alphas, weights = [], [], []
    alpha = minA.copy()
    for t in range(int(N.max()) if N.size else 0):
        act = t < N
        inr = act & (alpha >= mn) & (alpha <= mx)
        idx = np.nonzero(inr)[0]
        if idx.size:
            rays.append(idx)
            alphas.append(alpha[idx].astype(np.float64))
            wt = np.where((t == 0) | (N[idx] - 1 == t), 0.5, 1.0)
            weights.append(wt)
        alpha = np.where(act, (alpha + F32(step)).astype(F32), alpha)
    if not rays:
        return np.zeros(0, np.int64), np.zeros(0), np.zeros(0)
    return np.concatenate(rays), np.concatenate(alphas), np.concatenate(weights)


def _lattice(view, step, pixels=None):
    """Per ray: unit world direction (f32), alpha_t list and weights of in-range samples."""
    r, minA, mn, mx, N = _ray_setup(view, step, pixels)
    rays, alphas, wts = _walk(minA, mn, mx, N, step)
    return r.astype(np.float64), rays, alphas, wts


def rodrigues_t(w):
    th = torch.sqrt((w * w).sum() + 1e-300)
    K = torch.zeros(3, 3, dtype=w.dtype)
    K[0, 1], K[0, 2], K[1, 0], K[1, 2], K[2, 0], K[2, 1] = -w[2], w[1], w[2], -w[0], -w[1], w[0]
    eye = torch.eye(3, dtype=w.dtype)
    return eye + torch.sin(th) / th * K + (1 - torch.cos(th)) / th ** 2 * (K @ K)


class DenseDRR:
    """views: list of (ctypes View of the unperturbed pose, source_world (3,)).
    pixels: None for every pixel (render returns [n_views, H, W]), or flat pixel
    indices v * W + u shared by all views (render returns [n_views, len(pixels)])."""

    def __init__(self, views, world_from_ijk, center, step, E, pdf, mu, shape, pixels=None):
        self.step = float(step)
        self.shape = shape
        self.W_from_ijk = torch.tensor(np.asarray(world_from_ijk, np.float64))
        self.center = torch.tensor(np.asarray(center, np.float64))
        self.E = torch.tensor(np.asarray(E, np.float64))
        self.pdf = torch.tensor(np.asarray(pdf, np.float64))
        self.M = len(mu) // len(E)
        self.mu = torch.tensor(np.asarray(mu, np.float64).reshape(len(E), self.M))
        self.views = []
        # the alpha walk of all views' rays in one pass
        setups = [_ray_setup(view, step, pixels) for view, _ in views]
        starts = np.cumsum([0] + [len(st[1]) for st in setups])
        cat = [np.concatenate([st[k] for st in setups]) for k in range(1, 5)]
        rays_all, alphas_all, wts_all = _walk(*cat, step)
        for i, ((view, src), st) in enumerate(zip(views, setups)):
            sel = (rays_all >= starts[i]) & (rays_all < starts[i + 1])
            rays, alphas, wts = rays_all[sel] - starts[i], alphas_all[sel], wts_all[sel]
            r = st[0].astype(np.float64)
            x = np.asarray(src, np.float64)[None] + alphas[:, None] * r[rays]
            n = view.width * view.height if pixels is None else len(pixels)
            hw = (view.height, view.width) if pixels is None else (len(pixels),)
            self.views.append(dict(n=n, hw=hw, rays=torch.tensor(rays), x=torch.tensor(x),
                                   wt=torch.tensor(wts)))

    def render(self, density, onehot, theta, extra_ad=None):
        """density [ni,nj,nk] float64, onehot [M,ni,nj,nk] float64, theta [6] float64;
        extra_ad: None or per view [M, H*W] area densities (g/cm^2) added to the
        marched ones, held constant (the air outside the volume)."""
        R = rodrigues_t(theta[3:])
        T = torch.eye(4, dtype=torch.float64)
        T = T.clone()
        T[:3, :3] = R
        T[:3, 3] = self.center - R @ self.center + theta[:3]
        ijk_from_w = torch.linalg.inv(T @ self.W_from_ijk)
        ni, nj, nk = self.shape
        vol = torch.cat([density[None], onehot], 0).permute(0, 3, 2, 1)[None]  # (1, 1+M, nk, nj, ni)
        outs = []
        for i, v in enumerate(self.views):
            p = (v["x"] @ ijk_from_w[:3, :3].T + ijk_from_w[:3, 3]) - 1.0
            n = torch.tensor([ni, nj, nk], dtype=torch.float64)
            g = 2.0 * p / (n - 1.0) - 1.0
            s = F.grid_sample(vol, g.view(1, 1, 1, -1, 3), mode="bilinear", padding_mode="border",
                              align_corners=True)[0, :, 0, 0, :]  # (1+M, S)
            rho, seg = s[0], s[1:]
            contrib = (v["wt"] * rho)[None] * seg  # (M, S)
            ad = torch.zeros(self.M, v["n"], dtype=torch.float64).index_add(1, v["rays"], contrib)
            ad = ad * self.step / 10.0
            if extra_ad is not None:
                ad = ad + extra_ad[i]
            ex = torch.exp(-(self.mu @ ad))  # (B, R)
            I = ((self.E * self.pdf)[:, None] * ex).sum(0)
            outs.append(I.view(*v["hw"]))
        return torch.stack(outs)
