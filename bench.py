"""DRR benchmark on one MI355X.

    python bench.py [--steps K] [--warmup W] [--mode batched|plain] [--batch B]

Workload: 512x512x400 int16 HU phantom (0.8 x 0.8 x 1 mm, five materials)
projected onto a 1536x1536 detector at 0.194 mm, SDD 1020 mm, step 0.1 mm,
20-bin 120 kV spectrum, neglog; synthetic seeded data.

A step is one pass of the hot path over one batch of views:
  * batched: B poses (sampled in a 100 mm ball / 30 degree cone) in one
    persistent launch replayed from a hipGraph, device neglog, one D2H copy;
  * plain: the reference's flow per view -- host geometry, one launch, sync,
    D2H of intensity and photon_prob, host NumPy neglog.
Prints one JSON line (metric views/s).  Extra fields: the plain path's
views/s, the roofline of the dominant kernel (HIP events on its stream) and
the CPU transcription's throughput on a bounded sample of the same view.
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
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # non-packed FP32 lane-ops/s, 256 CUs @ 2.4 GHz


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_scene(_abi, ctx, materials, step):
    from deepdrr_amd._abi_types import Scene
    from deepdrr_amd.material import attenuation_table
    from deepdrr_amd.spectral_data import spectrum_tables, synthetic_120kv

    sc = Scene(step=step, n_materials=len(materials), attenuate_outside=0, air_index=0)
    _abi.call("drr_ctx_set_scene", ctx, C.byref(sc))
    E, pdf = spectrum_tables(synthetic_120kv())
    mu = attenuation_table(materials, E)
    _abi.call("drr_ctx_set_spectrum", ctx, len(E), E.ctypes.data, pdf.ctypes.data, mu.ctypes.data)
    return E, pdf, mu


def plain_views(torch, _abi, ctx, varr, views, dev, utils):
    """The reference's per-view flow; returns the images [N,H,W] and kernel ms per view."""
    imgs = []
    kms = []
    for view in views:
        H, W = view.height, view.width
        I = torch.empty((H, W), dtype=torch.float32, device=dev)
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


def cpu_baseline(view, case_vols, scene, E, pdf, mu, budget_s=15.0):
    """Oracle transcription on the host cores over a band of detector rows."""
    from oracle import ctrans

    from deepdrr_amd._abi_types import View
    threads = ctrans.default_threads()
    H = view.height
    rows = 8
    while True:
        sub = View()
        C.pointer(sub)[0] = view
        W = np.array(view.world_from_index[:], np.float32)
        v0 = H // 2 - rows // 2
        # shift the pixel origin to row v0 (u*W0 + (v+v0)*W1 + W2)
        W2 = (np.float32(v0) * W[[1, 4, 7]] + W[[2, 5, 8]]).astype(np.float32)
        sub.world_from_index[2], sub.world_from_index[5], sub.world_from_index[8] = map(float, W2)
        sub.height = rows
        t = time.perf_counter()
        ctrans.project(case_vols, sub, scene, E, pdf, mu, threads)
        dt = time.perf_counter() - t
        if dt > budget_s / 4 or rows * 2 > H:
            break
        rows = min(H, int(rows * max(2.0, budget_s / 4 / max(dt, 1e-3))))
    frac = rows / H
    return dict(value=frac / dt, unit="views/s", cores=threads, kind="port",
                sample=f"{rows} of {H} detector rows of the AP view ({frac:.4f} of the view), "
                       f"{dt:.1f} s, oracle/transcription.c with OpenMP")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["batched", "plain"], default="batched")
    ap.add_argument("--tuning", type=int, nargs=4, default=[8, 8, 4, 8],
                    help="tile_w tile_h waves_per_block blocks_per_cu")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--step-mm", type=float, default=0.1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-views", type=int, default=3)
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
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from deepdrr_amd import _abi, phantoms, utils, workload
    from deepdrr_amd import projector as P
    from deepdrr_amd._abi_types import Scene

    dev = torch.device("cuda", 0)
    torch.cuda.init()
    materials = sorted(workload.CLINICAL_CLASSES)
    ctx = C.c_void_p()
    _abi.call("drr_ctx_create", 0, C.byref(ctx))
    E, pdf, mu = make_scene(_abi, ctx, materials, args.step_mm)
    log("generating the 512x512x400 phantom")
    t = time.perf_counter()
    vh, hu = workload.clinical_volume_device(ctx, materials)
    torch.cuda.synchronize()
    log(f"phantom + GPU ingestion {time.perf_counter() - t:.1f} s")
    varr = (C.c_void_p * 1)(vh.value)
    geom = workload.clinical_geometry_volume()
    ap_dev = phantoms.ap_device(1536, 1536, 0.194)
    ap_view = P.build_view([geom], [0], ap_dev.get_camera_projection(), 4 * 1020.0)
    n_samples_ap = workload.count_samples(ap_view, args.step_mm, stride=2)

    # ---- plain path (reference flow), always measured -----------------------
    plain_n = args.plain_views
    log("plain warmup")
    plain_views(torch, _abi, ctx, varr, [ap_view], dev, utils)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, kms_plain = plain_views(torch, _abi, ctx, varr, [ap_view] * plain_n, dev, utils)
    torch.cuda.synchronize()
    plain_vps = plain_n / (time.perf_counter() - t)
    plain_kernel_ms = float(np.mean(kms_plain))
    log(f"plain: {plain_vps:.3f} views/s, march kernel {plain_kernel_ms:.2f} ms/view")

    result = dict(metric="views/s (512x512x400 CT->1536^2 det) batched vs plain; adjoint ms/view",
                  unit="views/s", n_gpus=1, higher_is_better=True, scaling="weak", vs_baseline=None,
                  dtype="f32", data="synthetic")
    steps, warm = args.steps, args.warmup
    if args.mode == "plain":
        for _ in range(warm):
            plain_views(torch, _abi, ctx, varr, [ap_view], dev, utils)
        torch.cuda.synchronize()
        t = time.perf_counter()
        kms = []
        for _ in range(steps):
            _, k = plain_views(torch, _abi, ctx, varr, [ap_view], dev, utils)
            kms += k
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        value = steps / dt
        kernel_ms = float(np.mean(kms))
        samples_per_launch = n_samples_ap
        kernel_name = "k_project_v1<5,false>"
        config = {"workload": "plain forward, 512x512x400 int16 CT (0.8x0.8x1 mm, 5 materials) -> "
                              "1536x1536 detector at 0.194 mm, AP view, step 0.1 mm, 20-bin 120 kV, neglog",
                  "views_per_step": 1, "mode": "plain"}
    else:
        from deepdrr_amd.batch import BatchPlan

        B = args.batch
        poses = phantoms.batch_poses(B, seed=0)
        views = [P.build_view([geom], [0], phantoms.pose_projection(1536, 1536, 0.194, p, d, f),
                              4 * 1020.0) for p, d, f in poses]
        log("counting in-volume samples of the batch")
        samples_per_launch = sum(workload.count_samples(v, args.step_mm, stride=2) for v in views)
        plan = BatchPlan(ctx, [vh], views, [0] * B, neglog=True, tuning=tuple(args.tuning))
        log(f"batched warmup ({B} views per launch)")
        for _ in range(warm):
            plan.launch()
            plan.fetch("images")
        torch.cuda.synchronize()
        kms = []
        t = time.perf_counter()
        for _ in range(steps):
            plan.launch()
            plan.fetch("images")       # one D2H of the B neglog images
            kms.append(plan.march_ms())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        value = B * steps / dt
        kernel_ms = float(np.mean(kms))
        kernel_name = "k_batch_march<5,false,8>"
        plan.close()
        config = {"workload": f"batched forward, {B} poses (100 mm ball, 30 deg cone, source fraction "
                              "U(0.5,0.9)) of a 512x512x400 int16 CT (0.8x0.8x1 mm, 5 materials) -> "
                              "1536x1536 detector at 0.194 mm, step 0.1 mm, 20-bin 120 kV, device neglog, "
                              "one persistent launch per step replayed from a hipGraph",
                  "views_per_step": B, "mode": "batched", "tuning": list(args.tuning)}
    ms_per_step = dt / steps * 1e3
    achieved = samples_per_launch * workload.BYTES_PER_SAMPLE / (kernel_ms * 1e-3) / 1e9
    result.update(value=value, steps=args.steps, warmup=args.warmup, ms_per_step=ms_per_step,
                  config=config,
                  plain_views_per_s=plain_vps, plain_kernel_ms=plain_kernel_ms,
                  batched_over_plain=value / plain_vps,
                  roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                            "kernel": kernel_name, "kernel_ms": kernel_ms,
                            "samples_per_launch": samples_per_launch,
                            "samples_per_s": samples_per_launch / (kernel_ms * 1e-3)})
    if not args.no_cpu:
        log("cpu baseline")
        from deepdrr_amd._abi_types import Scene as S
        hu_kji = np.ascontiguousarray(np.moveaxis(hu, [0, 1, 2], [2, 1, 0]))
        del hu
        dens = np.empty(hu_kji.shape, np.float32)
        labs = np.empty(hu_kji.shape, np.uint8)
        _abi.call("drr_volume_download", vh, dens.ctypes.data, labs.ctypes.data)
        del hu_kji
        sc = S(step=args.step_mm, n_materials=len(materials), attenuate_outside=0, air_index=0)
        result["cpu_baseline"] = cpu_baseline(ap_view, [(dens, labs)], sc, E, pdf, mu)
    _abi.lib().drr_volume_destroy(vh)
    _abi.lib().drr_ctx_destroy(ctx)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
