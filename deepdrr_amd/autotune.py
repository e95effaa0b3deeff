"""Schedule autotuner for the batched march (``drr_batch_tuning_t``).

The persistent batch kernel has four knobs: the ray tile (8x8 or 16x4 pixels
per wave), waves per workgroup and persistent workgroups per CU.  The best
setting depends on the detector sizes, the material count (register use and
so occupancy) and the volume sizes, so ``tune`` times a short list of
candidates on a sample of the batch's own views (the march kernel alone, HIP
events on its stream) and keeps the fastest.  Results are cached per
workload key in memory (guarded by a lock, so several host threads may plan
batches at once) and optionally in a JSON file named by the environment
variable ``DEEPDRR_AMD_TUNE_CACHE``.

Every candidate computes bit-identical images: only the schedule changes.
"""
from __future__ import annotations

import json
import os
import threading
from typing import Dict, Optional, Sequence, Tuple

Tuning = Tuple[int, int, int, int]  # tile_w, tile_h, waves_per_block, blocks_per_cu

CANDIDATES: Tuple[Tuning, ...] = (
    (8, 8, 4, 8),
    (8, 8, 4, 4),
    (8, 8, 2, 8),
    (8, 8, 2, 16),
    (8, 8, 1, 16),
    (8, 8, 4, 2),
    (16, 4, 4, 8),
    (16, 4, 2, 8),
)

_lock = threading.Lock()
_cache: Dict[str, Tuning] = {}
_file_loaded = False


def valid(t: Sequence[int]) -> bool:
    """The tunings the library accepts (checked again in drr_jobs_build)."""
    tw, th, wpb, bpc = (int(x) for x in t)
    return (tw, th) in ((8, 8), (16, 4)) and 1 <= wpb <= 4 and 1 <= bpc <= 32


def workload_key(views: Sequence, volume_indices: Sequence[int], vol_shapes: Sequence[Tuple[int, int, int]],
                 n_materials: int, device_name: str = "", march: str = "") -> str:
    """Detector sizes, volume shapes, materials and the march ("fast" or
    "exact": their loops differ, and so do their best schedules); not the
    poses (the best schedule does not depend on them to first order)."""
    dets = sorted({(int(v.width), int(v.height)) for v in views})
    vols = sorted({tuple(int(s) for s in vol_shapes[i]) for i in set(int(x) for x in volume_indices)})
    nb = 1 << max(0, (len(views) - 1).bit_length())  # batch size bucket
    return json.dumps({"dev": device_name, "M": int(n_materials), "det": dets, "vol": vols, "n": nb,
                       "march": march}, sort_keys=True)


def _load_file():
    global _file_loaded
    if _file_loaded:
        return
    _file_loaded = True
    path = os.environ.get("DEEPDRR_AMD_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                for k, v in json.load(f).items():
                    if valid(v):
                        _cache[k] = tuple(int(x) for x in v)
        except (OSError, ValueError):
            pass


def _save_file():
    path = os.environ.get("DEEPDRR_AMD_TUNE_CACHE")
    if not path:
        return
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump({k: list(v) for k, v in _cache.items()}, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def cached(key: str) -> Optional[Tuning]:
    with _lock:
        _load_file()
        return _cache.get(key)


def tune(ctx, vol_handles: Sequence, views: Sequence, volume_indices: Sequence[int], key: str,
         candidates: Sequence[Tuning] = CANDIDATES, sample: int = 8, repeats: int = 2,
         device=None) -> Tuning:
    """Fastest candidate on up to ``sample`` evenly spaced views of the batch
    (kernel time of the persistent march); cached under ``key``."""
    from .batch import BatchPlan

    hit = cached(key)
    if hit is not None:
        return hit
    n = len(views)
    k = min(n, max(1, sample))
    idx = [round(i * (n - 1) / max(1, k - 1)) for i in range(k)] if k > 1 else [0]
    sub_views = [views[i] for i in idx]
    sub_vidx = [volume_indices[i] for i in idx]
    best, best_ms = None, float("inf")
    for cand in candidates:
        if not valid(cand):
            continue
        plan = BatchPlan(ctx, vol_handles, sub_views, sub_vidx, neglog=False, tuning=cand, device=device)
        try:
            plan.launch()  # warm-up
            ms = float("inf")
            for _ in range(max(1, repeats)):
                plan.launch()
                plan.torch.cuda.synchronize(plan.device)
                ms = min(ms, plan.march_ms())
        finally:
            plan.close()
        if ms < best_ms:
            best, best_ms = tuple(cand), ms
    with _lock:
        _cache[key] = best
        _save_file()
    return best
