"""DRR benchmark on one MI355X.

    python bench.py [--steps K] [--warmup W] [--batch B] [--quick]

Headline workload (the metric's): a 512x512x400 int16 HU phantom (0.8 x 0.8 x
1 mm, five materials) projected onto a 1536x1536 detector at 0.194 mm, SDD
1020 mm, step 0.1 mm, 20-bin 120 kV spectrum, neglog; synthetic seeded data.

A step is one pass of the hot path over one batch of B views: B poses (100 mm
ball / 30 degree cone / source fraction U(0.5, 0.9)) in one persistent launch
replayed from a hipGraph, device neglog, one device-to-host copy of the
images.  `value` is batched views/s.  Also reported (one JSON line):
  * plain_views_per_s -- the reference's flow per view (host geometry, one
    launch, sync, D2H of intensity and photon_prob, NumPy neglog), same AP view;
  * mixed_batch -- 256 poses over 3 volumes of uneven size and detectors of
    384^2 / 768^2 / 1536^2 in one persistent launch;
  * adjoint -- 192^3 density, 64 views of 256^2: backward ms/view and its
    ratio to the forward;
  * monte_carlo -- Woodcock transport of photon histories into the 512x512x400
    phantom onto a 256^2 detector: histories/s and flight steps/s;
  * roofline of the dominant kernel (HIP events on its stream) and the CPU
    transcription's throughput on a bounded sample of the same view.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_ctx(_abi, materials, step, march="fast"):
    from deepdrr_amd._abi_types import MARCH_EXACT, MARCH_FAST, Scene
    from deepdrr_amd.material import attenuation_table
    from deepdrr_amd.spectral_data import spectrum_tables, synthetic_120kv

    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    sc = Scene(step=step, n_materials=len(materials), attenuate_outside=0, air_index=0,
               march=MARCH_FAST if march == "fast" else MARCH_EXACT)
    _abi.call("drr_ctx_set_scene", ctx, C.byref(sc))
    E, pdf = spectrum_tables(synthetic_120kv())
    mu = attenuation_table(materials, E)
    _abi.call("drr_ctx_set_spectrum", ctx, len(E), E.ctypes.data, pdf.ctypes.data, mu.ctypes.data)
    return ctx, (E, pdf, mu, sc)


PMC_SUMMARY = os.path.join("profiles", "k_batch_march_clinical_1536_pmc.md")


def profiled_counters(views_per_launch, march, tuning):
    """Counter figures of the batched march on this workload, read from the
    committed rocprofv3 summary, for the record of this march and schedule:
    HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, each scaled by a copy-kernel
    calibration), the VALU issue figures (SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU2,
    SQ_CYCLES) and the L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS)).  Returns
    (bytes, valu dict, L2 hit rate), Nones when the summary holds no record
    for this batch size, march and schedule."""
    path = os.path.join(ROOT, PMC_SUMMARY)
    if not os.path.exists(path):
        return None, None, None
    keys = ("march", "tuning", "views_per_launch", "traffic_bytes_per_launch", "valu_issue_frac",
            "valu_dual_issue_share", "valu_lane_instructions_per_sample", "l2_hit_rate")
    records, cur = [], {}
    for ln in open(path):
        ln = ln.strip()
        if ln.startswith("march:") and cur:
            records.append(cur)
            cur = {}
        for key in keys:
            if ln.startswith(key + ":"):
                cur[key] = ln.split(":", 1)[1].strip()
    if cur:
        records.append(cur)
    want = " ".join(str(int(x)) for x in tuning)
    for vals in records:
        if (vals.get("march") == march and vals.get("tuning") == want
                and int(vals.get("views_per_launch", -1)) == views_per_launch and "traffic_bytes_per_launch" in vals):
            valu = None
            if "valu_issue_frac" in vals:
                valu = {"issue_frac": float(vals["valu_issue_frac"]),
                        "dual_issue_share": float(vals["valu_dual_issue_share"]),
                        "lane_instructions_per_sample": float(vals["valu_lane_instructions_per_sample"]),
                        "note": "VALU issue quad-cycles used / available over the launch "
                                "((SQ_INSTS_VALU - SQ_ACTIVE_INST_VALU2) / (SQ_CYCLES x 8)); the kernel's "
                                "binding resource", "source": PMC_SUMMARY}
            l2 = float(vals["l2_hit_rate"]) if "l2_hit_rate" in vals else None
            return float(vals["traffic_bytes_per_launch"]), valu, l2
    return None, None, None


def plain_views(torch, _abi, ctx, varr, views, dev, utils):
    """The reference's per-view flow (projector.py:655-707): per view, the host
    geometry (``views`` holds callables that build the drr_view_t from the
    camera projection), one launch, sync, D2H of intensity and photon_prob, and
    NumPy neglog; returns images [N,H,W] and kernel ms per view."""
    imgs, kms = [], []
    for make_view in views:
        view = make_view()
        I = torch.empty((view.height, view.width), dtype=torch.float32, device=dev)
        P = torch.empty_like(I)
        _abi.call("drr_project", ctx, varr, C.byref(view), _abi.ptr(I), _abi.ptr(P), None, None,
                  _abi.stream_ptr())
        i_h = I.cpu().numpy()
        _ = P.cpu().numpy()
        ms = C.c_float()
        _abi.call("drr_ctx_last_kernel_ms", ctx, C.byref(ms))
        kms.append(ms.value)
        imgs.append(i_h)
    return utils.neglog(np.stack(imgs)), kms


def cpu_baseline(view, case_vols, scene, E, pdf, mu, budget_s=40.0):
    """Oracle transcription on the host cores over detector rows spread evenly
    across the view (every k-th row, so short and long rays are sampled in
    their proportion), scaled to views/s."""
    from oracle import ctrans
    from deepdrr_amd._abi_types import View
    threads = ctrans.default_threads()
    H = view.height
    W = np.array(view.world_from_index[:], np.float64)
    stride = 256
    while True:
        rows = len(range(stride // 2, H, stride))
        sub = View()
        C.pointer(sub)[0] = view
        # row j of the sub-view is row v = stride * j + stride // 2 of the view:
        # (j + 0.5) * (stride * w1) + (w2 + c * w1) = (v + 0.5) * w1 + w2
        c = stride // 2 + 0.5 - 0.5 * stride
        W1 = (stride * W[[1, 4, 7]]).astype(np.float32)
        W2 = (W[[2, 5, 8]] + c * W[[1, 4, 7]]).astype(np.float32)
        for i, a in enumerate((1, 4, 7)):
            sub.world_from_index[a] = float(W1[i])
        for i, a in enumerate((2, 5, 8)):
            sub.world_from_index[a] = float(W2[i])
        sub.height = rows
        t = time.perf_counter()
        ctrans.project(case_vols, sub, scene, E, pdf, mu, threads)
        dt = time.perf_counter() - t
        if dt > budget_s / 4 or stride == 1:
            break
        # aim past the stopping time, so the last sample is the one timed
        grow = max(2.0, 1.25 * budget_s / 4 / max(dt, 1e-3))
        stride = max(1, int(stride / grow))
    frac = rows / H
    return dict(value=frac / dt, unit="views/s", cores=threads, kind="port",
                sample=f"{rows} of {H} detector rows of the AP 1536x1536 view, every {stride}-th row "
                       f"({frac:.4f} of the view), {dt:.1f} s, oracle/transcription.c with OpenMP")


def measure_mixed(torch, _abi, ctx, clinical_handle, geom_clinical, step, B=256):
    """256 poses over 3 volumes (512x512x400, 256x256x300, 128^3) and detectors
    384^2 @ 0.776, 768^2 @ 0.388, 1536^2 @ 0.194 mm in one persistent launch."""
    from deepdrr_amd import phantoms, workload
    from deepdrr_amd import projector as P
    from deepdrr_amd.batch import BatchPlan

    mats = sorted(workload.CLINICAL_CLASSES)
    vol2 = phantoms.clinical_ct((256, 256, 300), spacing=(1.0, 1.0, 1.0), seed=1)
    vol3 = phantoms.small_ct(128, seed=2)
    handles = [clinical_handle]
    for v in (vol2, vol3):
        d, l = P.volume_arrays(v, mats)
        h = C.c_void_p()
        nz, ny, nx = d.shape
        _abi.call("drr_volume_create", ctx, d.ctypes.data, l.ctypes.data, nx, ny, nz, 0, C.byref(h))
        handles.append(h)
    geoms = [geom_clinical, vol2, vol3]
    dets = [(384, 0.776), (768, 0.388), (1536, 0.194)]
    views, vidx = [], []
    for i, (p, d, f) in enumerate(phantoms.batch_poses(B, seed=0)):
        vi = i % 3
        n, pix = dets[(i // 3) % 3]
        c = np.array(geoms[vi].center_in_world)
        proj = phantoms.pose_projection(n, n, pix, p, d, f, center=c)
        views.append(P.build_view([geoms[vi]], [0], proj, 4 * 1020.0))
        vidx.append(vi)
    samples = sum(workload.count_samples(v, step, stride=4) for v in views)
    plan = BatchPlan(ctx, handles, views, vidx, neglog=True)
    plan.launch()
    plan.fetch("images")
    torch.cuda.synchronize()
    t = time.perf_counter()
    plan.launch()
    plan.fetch("images")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    ms = plan.march_ms()
    plan.close()
    for h in handles[1:]:
        _abi.lib().drr_volume_destroy(h)
    return dict(views=B, views_per_s=B / dt, seconds=dt, march_ms=ms, samples=samples,
                samples_per_s=samples / (ms * 1e-3),
                detectors="384^2/768^2/1536^2", volumes="512x512x400, 256x256x300, 128^3")


def measure_adjoint(torch, n=192, n_views=64, det=256, step=0.1, repeats=3):
    """Forward and backward of 64 views 256^2 of a 192^3 density: MSE on the
    intensity against a target rendered at a perturbed pose (+-5 mm, +-3 deg,
    seed 1); backward for density and 6-DoF pose, and for density alone."""
    from deepdrr_amd import _abi, phantoms
    from deepdrr_amd.autograd import DifferentiableProjector
    from deepdrr_amd.spectral_data import synthetic_120kv

    hu = phantoms.small_ct_hu(n, seed=0)
    vol = phantoms.Volume.from_hu(hu, origin=phantoms._centered_origin((n,) * 3, (1, 1, 1)),
                                  anatomical_coordinate_system="LPS")
    mats = sorted(vol.materials[0].keys())
    remap = np.array([mats.index(k) for k in vol.materials[0]], np.uint8)
    labels = remap[vol.materials[1]]
    dp = DifferentiableProjector(labels, mats, vol.world_from_IJK.data, step=step, spectrum=synthetic_120kv())
    ang = np.linspace(0, 2 * np.pi, n_views, endpoint=False)
    projs = [phantoms.pose_projection(det, det, 1.6, [0, 0, 0], [np.sin(a), np.cos(a), 0.0], 0.5)
             for a in ang]
    rng = np.random.default_rng(1)
    pert = np.concatenate([rng.uniform(-5, 5, 3), np.deg2rad(rng.uniform(-3, 3, 3))]).astype(np.float32)
    dens = torch.tensor(vol.data, device="cuda", requires_grad=True)
    with torch.no_grad():
        target = dp(dens.detach(), torch.tensor(pert, device="cuda"), projs)

    def kernel_ms():
        ms = C.c_float()
        _abi.call("drr_ctx_last_kernel_ms", dp.ctx, C.byref(ms))
        return ms.value

    out = {}
    for mode in ("pose", "density", "pose_only"):
        theta = torch.zeros(6, device="cuda", requires_grad=(mode != "density"))
        dens.requires_grad_(mode != "pose_only")
        fw, bw, fk, bk = [], [], [], []
        for it in range(repeats + 1):  # first one warms up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            I = dp(dens, theta, projs)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            kf = kernel_ms()
            loss = ((I - target) ** 2).mean()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            kb = kernel_ms()
            dens.grad = None
            theta.grad = None
            if it:
                fw.append(t1 - t0)
                bw.append(t3 - t2)
                fk.append(kf)
                bk.append(kb)
        out[mode] = (float(np.median(fw)) * 1e3 / n_views, float(np.median(bw)) * 1e3 / n_views,
                     float(np.median(fk)) / n_views, float(np.median(bk)) / n_views)
    dp.close()
    fwd, bwd, fk, bk = out["pose"]
    _, bwd_d, _, bk_d = out["density"]
    _, bwd_p, _, bk_p = out["pose_only"]
    return dict(volume=f"{n}^3", views=n_views, detector=f"{det}^2", forward_ms_per_view=fwd,
                adjoint_ms_per_view=bwd, adjoint_over_forward=bwd / fwd,
                adjoint_density_only_ms_per_view=bwd_d, adjoint_density_only_over_forward=bwd_d / fwd,
                forward_kernel_ms_per_view=fk, adjoint_kernel_ms_per_view=bk,
                adjoint_density_only_kernel_ms_per_view=bk_d,
                adjoint_pose_only_ms_per_view=bwd_p, adjoint_pose_only_kernel_ms_per_view=bk_p,
                note="wall time (median of 3) of one drr_project_jobs launch (forward, saving area "
                     "densities) and one drr_adjoint_jobs launch (backward) over all 64 views; kernel_ms "
                     "from HIP events around the launches; MSE against a target at a perturbed pose")


def measure_mc(torch, _abi, ctx, handle, geom, n_hist):
    from deepdrr_amd import mc, phantoms, workload
    from deepdrr_amd.spectral_data import synthetic_120kv

    mats = sorted(workload.CLINICAL_CLASSES)
    dev = phantoms.ap_device(256, 256, 1.164, sdd=1020.0)
    p = mc.make_params(mats, synthetic_120kv(), dev.get_camera_projection(), 1020.0,
                       geom.world_from_IJK.data)
    run = mc.MonteCarlo(ctx, handle, p)
    run.run(min(n_hist, 1 << 20), seed=1)  # warm-up (also builds the voxel-order copy)
    torch.cuda.synchronize()
    run = mc.MonteCarlo(ctx, handle, p)
    t = time.perf_counter()
    run.run(n_hist, seed=12345)
    img, _, steps = run.results()
    dt = time.perf_counter() - t
    bsteps = run.brick_steps()
    return dict(histories=n_hist, seconds=dt, histories_per_s=n_hist / dt, flight_steps=steps,
                flights_per_history=steps / n_hist, flight_steps_per_s=steps / dt,
                brick_steps=bsteps, brick_steps_per_history=bsteps / n_hist,
                gather_GB_per_s=5.0 * steps / dt / 1e9,
                majorant="local: one per 16x16x16-voxel brick",
                scatter_fraction=float(img[1:].sum() / img.sum()), detector="256^2 at 1.164 mm")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--tuning", type=int, nargs=4, default=None,
                    help="tile_w tile_h waves_per_block blocks_per_cu (default: autotuned)")
    ap.add_argument("--step-mm", type=float, default=0.1)
    ap.add_argument("--march", choices=("fast", "exact"), default="fast",
                    help="single-volume march: fast (nested lerps, images within 1e-5 of the "
                         "transcription) or exact (area densities bit-identical to it)")
    ap.add_argument("--plain-views", type=int, default=3)
    ap.add_argument("--mc-histories", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline and plain path only")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint measurement (192^3, 64 views of 256^2)")
    ap.add_argument("--mc-only", action="store_true",
                    help="only the Monte-Carlo measurement (--mc-histories on the 512x512x400 phantom)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration for rocprof)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return
    if args.adjoint_only:
        import torch

        print(json.dumps({"adjoint": measure_adjoint(torch)}), flush=True)
        return
    if args.mc_only:
        import torch

        from deepdrr_amd import _abi, workload

        materials = sorted(workload.CLINICAL_CLASSES)
        ctx, _ = make_ctx(_abi, materials, args.step_mm)
        vh, _ = workload.clinical_volume_device(ctx, materials)
        mc = measure_mc(torch, _abi, ctx, vh, workload.clinical_geometry_volume(), args.mc_histories)
        _abi.lib().drr_volume_destroy(vh)
        _abi.lib().drr_ctx_destroy(ctx)
        print(json.dumps({"monte_carlo": mc}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; the views would shard over GPUs, but no "
                                   "multi-GPU path is implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from deepdrr_amd import _abi, phantoms, utils, workload
    from deepdrr_amd import projector as P
    from deepdrr_amd.batch import BatchPlan

    dev = torch.device("cuda", 0)
    torch.cuda.init()
    materials = sorted(workload.CLINICAL_CLASSES)
    ctx, (E, pdf, mu, scene) = make_ctx(_abi, materials, args.step_mm, args.march)
    log("generating the 512x512x400 phantom")
    t = time.perf_counter()
    vh, hu = workload.clinical_volume_device(ctx, materials)
    torch.cuda.synchronize()
    log(f"phantom + GPU ingestion {time.perf_counter() - t:.1f} s")
    varr = (C.c_void_p * 1)(vh.value)
    geom = workload.clinical_geometry_volume()
    ap_proj = phantoms.ap_device(1536, 1536, 0.194).get_camera_projection()
    ap_view = P.build_view([geom], [0], ap_proj, 4 * 1020.0)

    # ---- plain path: the reference's flow, one view per launch ----------------
    def ap_view_host():
        return P.build_view([geom], [0], ap_proj, 4 * 1020.0)

    plain_views(torch, _abi, ctx, varr, [ap_view_host], dev, utils)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, kms_plain = plain_views(torch, _abi, ctx, varr, [ap_view_host] * args.plain_views, dev, utils)
    torch.cuda.synchronize()
    plain_vps = args.plain_views / (time.perf_counter() - t)
    plain_kernel_ms = float(np.mean(kms_plain))
    n_samples_ap = workload.count_samples(ap_view, args.step_mm, stride=2)
    log(f"plain: {plain_vps:.2f} views/s, march {plain_kernel_ms:.2f} ms/view")

    # ---- headline: batched clinical views ---------------------------------------
    B = args.batch
    views = [P.build_view([geom], [0], phantoms.pose_projection(1536, 1536, 0.194, p, d, f), 4 * 1020.0)
             for p, d, f in phantoms.batch_poses(B, seed=0)]
    samples_per_launch = sum(workload.count_samples(v, args.step_mm, stride=2) for v in views)
    from deepdrr_amd import autotune

    tune_key = autotune.workload_key(views, [0] * B, [(512, 512, 400)], len(materials),
                                     torch.cuda.get_device_name(0), march=args.march)
    t = time.perf_counter()
    plan = BatchPlan(ctx, [vh], views, [0] * B, neglog=True,
                     tuning=tuple(args.tuning) if args.tuning else "auto", tune_key=tune_key)
    log(f"schedule {plan.tuning} ({'given' if args.tuning else 'autotuned'}, {time.perf_counter() - t:.1f} s)")
    # Each step's images reach host memory; the copy of one step crosses the
    # link on the plan's copy stream while the next step marches.
    for _ in range(args.warmup):
        plan.launch()
        plan.fetch_async("images").result()
    torch.cuda.synchronize()
    kms = []
    pending = None
    t = time.perf_counter()
    for _ in range(args.steps):
        plan.launch()
        nxt = plan.fetch_async("images")
        if pending is not None:
            pending.result()
        pending = nxt
        kms.append(plan.march_ms())
    pending.result()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    plan.close()
    value = B * args.steps / dt
    kernel_ms = float(np.mean(kms))
    # the same batch marched by the other march (same schedule), for comparison
    from deepdrr_amd._abi_types import MARCH_EXACT, MARCH_FAST

    other = "exact" if args.march == "fast" else "fast"
    scene.march = MARCH_EXACT if other == "exact" else MARCH_FAST
    _abi.call("drr_ctx_set_scene", ctx, C.byref(scene))
    plan_o = BatchPlan(ctx, [vh], views, [0] * B, neglog=True, tuning=tuple(plan.tuning))
    kms_o = []
    for i in range(3):
        plan_o.launch()
        torch.cuda.synchronize()
        if i:
            kms_o.append(plan_o.march_ms())
    plan_o.close()
    scene.march = MARCH_FAST if args.march == "fast" else MARCH_EXACT
    _abi.call("drr_ctx_set_scene", ctx, C.byref(scene))
    other_march_ms = float(np.mean(kms_o))
    log(f"{other} march of the same batch: {other_march_ms / B:.2f} ms/view")
    gather_gbs = samples_per_launch * workload.BYTES_PER_SAMPLE / (kernel_ms * 1e-3) / 1e9
    traffic, valu, l2_hit = profiled_counters(B, args.march, plan.tuning)
    # HBM bytes the launch moves (rocprof FETCH_SIZE + WRITE_SIZE of the same
    # launch, copy-calibrated) over the live kernel time
    achieved = traffic / (kernel_ms * 1e-3) / 1e9 if traffic is not None else None
    log(f"batched: {value:.2f} views/s, march {kernel_ms / B:.2f} ms/view")

    result = dict(
        metric="views/s (512x512x400 CT->1536^2 det) batched vs plain; adjoint ms/view",
        value=value, unit="views/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=dt / args.steps * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32", data="synthetic",
        config={"workload": f"batched forward: {B} poses (100 mm ball, 30 deg cone, source fraction "
                            "U(0.5,0.9)) of a 512x512x400 int16 CT (0.8x0.8x1 mm, 5 materials) -> "
                            "1536x1536 detector at 0.194 mm, SDD 1020 mm, step 0.1 mm, 20-bin 120 kV, "
                            "device neglog, one persistent launch per step replayed from a hipGraph, each step's "
                            "images copied to host memory while the next step marches",
                "views_per_step": B, "tuning": list(plan.tuning),
                "tuning_source": "given" if args.tuning else "autotuned"},
        march=args.march,
        plain_views_per_s=plain_vps, plain_march_ms_per_view=plain_kernel_ms,
        batched_march_ms_per_view=kernel_ms / B, batched_over_plain=value / plain_vps,
        **{f"batched_{other}_march_ms_per_view": other_march_ms / B},
        plain_samples_per_view=n_samples_ap,
        plain_march_samples_per_s=n_samples_ap / (plain_kernel_ms * 1e-3),
        roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS if achieved is not None else None, "traffic": traffic,
                  "achieved_source": "traffic (HBM bytes per launch, rocprofv3) / kernel_ms (HIP events, live)",
                  "traffic_source": ("rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated, the same march and "
                                     f"schedule ({PMC_SUMMARY})" if traffic is not None else None),
                  "valu": valu, "l2_hit_rate": l2_hit,
                  "kernel": f"k_batch_march<{len(materials)},false,{plan.tuning[0]},"
                            f"{'true' if args.march == 'fast' else 'false'}>",
                  "kernel_ms": kernel_ms, "samples_per_launch": samples_per_launch,
                  "samples_per_s": samples_per_launch / (kernel_ms * 1e-3),
                  "gather_equivalent": {
                      "bytes_per_sample": workload.BYTES_PER_SAMPLE, "GB_per_s": gather_gbs,
                      "note": "8 f32 density + 8 u8 label corners per in-volume sample (the reference "
                              "kernel's texel footprint) x samples / kernel time; most of these bytes are "
                              "served from registers, L1 and L2, so this is not an HBM rate"}},
    )
    if not args.quick:
        log("mixed batch (3 volumes, 3 detector sizes, 256 poses)")
        result["mixed_batch"] = measure_mixed(torch, _abi, ctx, vh, geom, args.step_mm)
        log(f"mixed: {result['mixed_batch']['views_per_s']:.1f} views/s")
        log("Monte Carlo")
        result["monte_carlo"] = measure_mc(torch, _abi, ctx, vh, geom, args.mc_histories)
        log(f"mc: {result['monte_carlo']['histories_per_s']:.3g} histories/s")
    if not args.no_cpu:
        log("cpu baseline")
        del hu
        dens = np.empty((400, 512, 512), np.float32)
        labs = np.empty((400, 512, 512), np.uint8)
        _abi.call("drr_volume_download", vh, dens.ctypes.data, labs.ctypes.data)
        result["cpu_baseline"] = cpu_baseline(ap_view, [(dens, labs)], scene, E, pdf, mu)
        del dens, labs
    _abi.lib().drr_volume_destroy(vh)
    _abi.lib().drr_ctx_destroy(ctx)
    if not args.quick:
        log("adjoint (192^3, 64 views of 256^2)")
        result["adjoint"] = measure_adjoint(torch)
        a = result["adjoint"]
        log(f"adjoint: {a['adjoint_ms_per_view']:.3f} ms/view ({a['adjoint_over_forward']:.2f}x forward), "
            f"density only {a['adjoint_density_only_ms_per_view']:.3f} ms/view")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
