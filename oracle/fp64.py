"""FP64 NumPy restatement of deepdrr's projectKernel.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as a checker; the product never imports it.

Follows deepdrr/projector/project_kernel.cu:195-646 (same structure as
oracle/transcription.c) vectorised over rays.  The sample lattice -- ray length,
slab interval, step count and the sequential FP32 ``alpha += step``
(project_kernel.cu:242-334, 552) -- is evaluated in float32 exactly as the
kernel does, so both implementations visit the same sample points; every
integrand quantity (positions, labels, trilinear weights, material fractions,
densities, area densities, Beer-Lambert sum) is evaluated in float64.  The
difference between this and the FP32 transcription is therefore the FP32
arithmetic error of the integrand.

Texture semantics as in oracle/transcription.c: labels at
clamp(floor(p + d)), density corners clamp(floor(p)), clamp(floor(p) + 1), the
label cache shared across volumes, disabled volumes skipped.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
AIR_DENSITY = 0.1129


def _ray_setup(view, width, height, pixels=None):
    """float32 ray directions and lengths, pixel order row-major [v][u], or of
    the flat pixel indices v * width + u in ``pixels``."""
    W = np.asarray(view["world_from_index"], dtype=F32).reshape(9)
    if pixels is None:
        vv, uu = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    else:
        pix = np.asarray(pixels, np.int64)
        vv, uu = pix // width, pix % width
    u = (uu.reshape(-1).astype(F32).astype(np.float64) + 0.5).astype(F32)
    v = (vv.reshape(-1).astype(F32).astype(np.float64) + 0.5).astype(F32)
    rx = (u * W[0] + v * W[1]) + W[2]
    ry = (u * W[3] + v * W[4]) + W[5]
    rz = (u * W[6] + v * W[7]) + W[8]
    L = np.sqrt((rx * rx + ry * ry) + rz * rz).astype(F32)
    inv = (F32(1.0) / L).astype(F32)
    return rx * inv, ry * inv, rz * inv, L


def project_fp64(volumes, view, scene, energies, pdf, mu, pixels=None):
    """volumes: list of (density [nz,ny,nx] f32, labels [nz,ny,nx] u8).
    view: dict(world_from_index[9], max_ray_length, width, height,
               vols=[dict(ijk_from_world[12], source_ijk[3], bounds_lo[3],
                          bounds_hi[3], priority, enabled)])
    scene: dict(step, n_materials, attenuate_outside, air_index)
    Returns intensity [H,W], photon_prob [H,W], area_density [M,H,W] (float64);
    with ``pixels`` (flat indices v * W + u) only those rays, as intensity [R],
    photon_prob [R], area_density [M,R].
    """
    Wd, Hd = int(view["width"]), int(view["height"])
    V = len(view["vols"])
    M = int(scene["n_materials"])
    step = F32(scene["step"])
    atten = bool(scene["attenuate_outside"])
    air = int(scene["air_index"])
    rx, ry, rz, L = _ray_setup(view, Wd, Hd, pixels)
    R = rx.shape[0]
    r32 = np.stack([rx, ry, rz], axis=1)

    minA = L.copy()
    maxA = np.zeros(R, dtype=F32)
    minAv = np.zeros((V, R), dtype=F32)
    maxAv = np.zeros((V, R), dtype=F32)
    trace = np.zeros((V, R), dtype=bool)
    rijk32 = np.zeros((V, R, 3), dtype=F32)
    mrl = F32(view["max_ray_length"])
    for i, vv in enumerate(view["vols"]):
        if not vv["enabled"]:
            continue
        A = np.asarray(vv["ijk_from_world"], dtype=F32).reshape(3, 4)
        for a in range(3):
            rijk32[i, :, a] = ((A[a, 0] * rx + A[a, 1] * ry) + A[a, 2] * rz) + A[a, 3] * F32(0)
        s = np.asarray(vv["source_ijk"], dtype=F32)
        lo = np.asarray(vv["bounds_lo"], dtype=F32)
        hi = np.asarray(vv["bounds_hi"], dtype=F32)
        mn = np.zeros(R, dtype=F32)
        mx = np.full(R, mrl if mrl > 0 else np.inf, dtype=F32)
        ok = np.ones(R, dtype=bool)
        for a in range(3):
            ra = rijk32[i, :, a]
            nz = ra != 0
            with np.errstate(divide="ignore", invalid="ignore"):
                reci = (F32(1.0) / np.where(nz, ra, F32(1.0))).astype(F32)
                a0 = ((lo[a] - s[a]) * reci).astype(F32)
                a1 = ((hi[a] - s[a]) * reci).astype(F32)
            upd = nz & ok
            mn = np.where(upd, np.fmax(mn, np.fmin(a0, a1)), mn)
            mx = np.where(upd, np.fmin(mx, np.fmax(a0, a1)), mx)
            outside = (~nz) & ((lo[a] > s[a]) | (s[a] > hi[a]))
            ok &= ~outside
        trace[i] = ok
        minAv[i], maxAv[i] = mn, mx
        minA = np.where(ok, np.fmin(minA, mn), minA)
        maxA = np.where(ok, np.fmax(maxA, mx), maxA)

    with np.errstate(invalid="ignore"):
        N = np.ceil(((maxA - minA) / step).astype(F32)).astype(np.int64)
    N = np.maximum(N, 0)

    ad = np.zeros((M, R), dtype=np.float64)
    if atten:
        ad[air] += (minA / step).astype(F32).astype(np.float64) * AIR_DENSITY

    # float64 geometry for positions
    src64 = [np.asarray(vv["source_ijk"], dtype=F32).astype(np.float64) for vv in view["vols"]]
    rijk64 = rijk32.astype(np.float64)
    dims = [np.array(d.shape[::-1]) for d, _ in volumes]  # (nx, ny, nz)

    alpha = minA.copy()
    have_prev = np.zeros(R, dtype=bool)
    prev = np.zeros((R, 3))
    vox = np.zeros((R, 8), dtype=np.int64)
    nmax = int(N.max()) if R else 0
    for t in range(nmax):
        act = t < N
        a64 = alpha.astype(np.float64)
        seg = np.zeros((V, M, R))
        P = {}
        for i, vv in enumerate(view["vols"]):
            if not vv["enabled"]:
                continue
            p = src64[i][None, :] + a64[:, None] * rijk64[i] - 1.0
            P[i] = p
            b = np.floor(p)
            change = act & ((~have_prev) | np.any(b != prev, axis=1))
            if np.any(change):
                dens, labs = volumes[i]
                n = dims[i]
                idx = np.nonzero(change)[0]
                pc = p[idx]
                for d in range(8):
                    dx, dy, dz = d & 1, (d >> 1) & 1, d >> 2
                    q = np.floor(pc + np.array([dx, dy, dz], dtype=np.float64))
                    q = np.clip(q, 0, n - 1).astype(np.int64)
                    vox[idx, d] = labs[q[:, 2], q[:, 1], q[:, 0]]
                prev[idx] = b[idx]
                have_prev[idx] = True
            f = p - b
            fi = 1.0 - f
            w = _weights(f, fi)
            for d in range(8):
                for m in range(M):
                    seg[i, m] += np.where(vox[:, d] == m, w[d], 0.0)
        cp = np.full(R, V)
        n_at = np.zeros(R, dtype=np.int64)
        for i, vv in enumerate(view["vols"]):
            if not vv["enabled"]:
                continue
            inr = trace[i] & (alpha >= minAv[i]) & (alpha <= maxAv[i]) & (seg[i].sum(0) > 0)
            pr = int(vv["priority"])
            lower = inr & (pr < cp)
            same = inr & (pr == cp)
            n_at = np.where(lower, 1, np.where(same, n_at + 1, n_at))
            cp = np.where(lower, pr, cp)
        empty = act & (n_at == 0)
        if atten:
            ad[air] += np.where(empty, AIR_DENSITY, 0.0)
        full = act & (n_at > 0)
        weight = np.where(full, 1.0 / np.maximum(n_at, 1), 0.0)
        weight = weight * np.where((t == 0) | (N - 1 == t), 0.5, 1.0)
        for i, vv in enumerate(view["vols"]):
            if not vv["enabled"]:
                continue
            sel = full & trace[i] & (cp == int(vv["priority"]))
            if not np.any(sel):
                continue
            dens, _ = volumes[i]
            n = dims[i]
            p = P[i]
            b = np.floor(p)
            f = p - b
            w = _weights(f, 1.0 - f)
            rho = np.zeros(R)
            for d in range(8):
                dx, dy, dz = d & 1, (d >> 1) & 1, d >> 2
                q = np.clip(b + np.array([dx, dy, dz]), 0, n - 1).astype(np.int64)
                rho += w[d] * dens[q[:, 2], q[:, 1], q[:, 0]].astype(np.float64)
            for m in range(M):
                ad[m] += np.where(sel, weight * rho * seg[i, m], 0.0)
        alpha = np.where(act, (alpha + step).astype(F32), alpha)

    if atten:
        ad[air] += ((L - maxA) / step).astype(F32).astype(np.float64) * AIR_DENSITY
    ad *= float(step)
    ad /= 10.0
    E = np.asarray(energies, np.float64)
    P_ = np.asarray(pdf, np.float64)
    MU = np.asarray(mu, np.float64).reshape(len(E), M)
    expo = MU @ ad  # [B, R]
    pp = np.exp(-expo) * P_[:, None]
    photon_prob = pp.sum(0)
    intensity = (E[:, None] * pp).sum(0)
    if pixels is not None:
        return intensity, photon_prob, ad
    return (intensity.reshape(Hd, Wd), photon_prob.reshape(Hd, Wd), ad.reshape(M, Hd, Wd))


def _weights(f, fi):
    fx, fy, fz = f[:, 0], f[:, 1], f[:, 2]
    gx, gy, gz = fi[:, 0], fi[:, 1], fi[:, 2]
    return [gx * gy * gz, fx * gy * gz, gx * fy * gz, fx * fy * gz,
            gx * gy * fz, fx * gy * fz, gx * fy * fz, fx * fy * fz]


def neglog_fp64(images: np.ndarray, epsilon: float = 0.01) -> np.ndarray:
    """deepdrr/utils/image_utils.py:18-59 in float64."""
    x = np.array(images, dtype=np.float64)
    single = x.ndim == 2
    if single:
        x = x[None]
    x = x + (x.min(axis=(1, 2), keepdims=True) + epsilon)
    y = -np.log(x)
    lo = y.min(axis=(1, 2), keepdims=True)
    hi = y.max(axis=(1, 2), keepdims=True)
    if np.any(hi == lo):
        y[:] = 0
    else:
        y = (y - lo) / (hi - lo)
    return y[0] if single else y
