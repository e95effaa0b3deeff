"""Batched forward projection: many (volume, pose, detector) views in one
persistent launch replayed from a hipGraph.

This is the batched form of deepdrr's per-view loop
(``deepdrr/projector/projector.py:679-707``): the same images, computed by one
graph launch with the neglog epilogue on the device and one device-to-host
copy at the end.  A view over n volumes uses the plan's volumes
``volume_index .. volume_index + n - 1``.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _abi, geo
from ._abi_types import BatchTuning, Job

DEFAULT_TUNING = (8, 8, 4, 8)  # tile_w, tile_h, waves_per_block, blocks_per_cu


class BatchPlan:
    """A captured batch over device-resident volumes.

    ``ctx``: a drr_ctx handle; ``vol_handles``: drr_volume handles; ``views``:
    ctypes ``View`` objects and ``volume_indices`` into ``vol_handles`` (the
    first of a view's ``n_volumes`` consecutive handles).  Outputs live in device tensors ``intensity``,
    ``photon_prob`` and (with ``neglog``) ``images``, each image at
    ``offsets[i]`` with shape ``shapes[i]``.

    ``tuning``: a ``(tile_w, tile_h, waves_per_block, blocks_per_cu)`` tuple,
    None for the default, or ``"auto"`` to time the candidates of
    ``deepdrr_amd.autotune`` on a sample of these views first (cached under
    ``tune_key``, see ``autotune.workload_key``).
    """

    def __init__(self, ctx, vol_handles: Sequence, views: Sequence, volume_indices: Sequence[int],
                 neglog: bool = True, tuning=None, device=None, tune_key: Optional[str] = None):
        import torch

        self.torch = torch
        self.device = device or torch.device("cuda", 0)
        n = len(views)
        if n == 0:
            raise ValueError("empty batch")
        if len(volume_indices) != n:
            raise ValueError("one volume index per view is needed")
        if isinstance(tuning, str):
            if tuning != "auto":
                raise ValueError("tuning must be a 4-tuple, None or 'auto'")
            from . import autotune

            tuning = autotune.tune(ctx, vol_handles, views, volume_indices,
                                   tune_key or autotune.workload_key(views, volume_indices,
                                                                     [(0, 0, 0)] * len(vol_handles), 0),
                                   device=device)
        self.tuning = tuple(tuning or DEFAULT_TUNING)
        self.shapes = [(v.height, v.width) for v in views]
        sizes = [h * w for h, w in self.shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        total = int(np.sum(sizes))
        jobs = (Job * n)()
        for i, (v, vi) in enumerate(zip(views, volume_indices)):
            if not 0 <= int(vi) <= len(vol_handles) - v.n_volumes:
                raise ValueError(f"view {i}: volume index {vi} with {v.n_volumes} volumes is out of range")
            jobs[i].view = v
            jobs[i].volume_index = int(vi)
            jobs[i].out_offset = int(self.offsets[i])
        self.intensity = torch.empty(total, dtype=torch.float32, device=self.device)
        self.photon_prob = torch.empty(total, dtype=torch.float32, device=self.device)
        self.neglog = bool(neglog)
        self.images = torch.empty(total, dtype=torch.float32, device=self.device) if neglog else None
        t = BatchTuning(*self.tuning)
        self._varr = (C.c_void_p * len(vol_handles))(*[h.value if hasattr(h, "value") else h
                                                         for h in vol_handles])
        self._batch = C.c_void_p()
        _abi.call("drr_batch_create", ctx, self._varr, len(vol_handles), jobs, n, C.byref(t),
                  _abi.ptr(self.intensity), _abi.ptr(self.photon_prob), _abi.ptr(self.images),
                  1 if neglog else 0, C.byref(self._batch))
        self.n_views = n
        self.total = total
        self._copy_stream = None  # fetch_async: copy stream, staging buffer, last copy
        self._staging = None
        self._copy_done = None

    def launch(self, stream=None) -> None:
        s = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        _abi.call("drr_batch_launch", self._batch, _abi.stream_ptr(s))

    def march_ms(self) -> float:
        ms = C.c_float()
        _abi.call("drr_batch_last_march_ms", self._batch, C.byref(ms))
        return ms.value

    def fetch(self, which: str = "images") -> List[np.ndarray]:
        buf = {"images": self.images, "intensity": self.intensity, "photon_prob": self.photon_prob}[which]
        if buf is None:
            raise ValueError(f"{which} was not requested when the batch was planned")
        # A page-locked destination lets the copy run as one DMA at link speed
        # (a pageable one is staged through a bounce buffer).  Each call gets
        # its own buffer, so earlier results stay valid; torch's pinned-memory
        # cache hands back the block once those arrays are dropped.
        host = self.torch.empty(buf.numel(), dtype=buf.dtype, pin_memory=True)
        host.copy_(buf)
        host = host.numpy()
        return [host[o:o + h * w].reshape(h, w) for o, (h, w) in zip(self.offsets, self.shapes)]

    def fetch_async(self, which: str = "images") -> "PendingFetch":
        """Start copying this launch's images to host memory without waiting.

        The images go device-to-device into a staging buffer on the current
        stream (a copy at HBM speed), then to page-locked host memory on a copy
        stream of this plan's own, so the next ``launch`` runs while they cross
        the link.  ``PendingFetch.result()`` waits for that copy and returns the
        arrays as ``fetch`` does.  A later call first waits (on the device)
        until the previous host copy has left the staging buffer."""
        torch = self.torch
        buf = {"images": self.images, "intensity": self.intensity, "photon_prob": self.photon_prob}[which]
        if buf is None:
            raise ValueError(f"{which} was not requested when the batch was planned")
        if self._copy_stream is None:
            self._copy_stream = torch.cuda.Stream(self.device)  # default priority
        cur = torch.cuda.current_stream(self.device)
        if self._staging is None or self._staging.numel() < buf.numel():
            self._staging = torch.empty(buf.numel(), dtype=buf.dtype, device=self.device)
        if self._copy_done is not None:
            cur.wait_event(self._copy_done)
        staging = self._staging[:buf.numel()]
        staging.copy_(buf)
        ready = cur.record_event()
        host = torch.empty(buf.numel(), dtype=buf.dtype, pin_memory=True)
        with torch.cuda.stream(self._copy_stream):
            self._copy_stream.wait_event(ready)
            host.copy_(staging, non_blocking=True)
            done = self._copy_stream.record_event()
        # the copy stream reads the staging buffer: keep torch's allocator from
        # handing it out again before that copy is done
        self._staging.record_stream(self._copy_stream)
        self._copy_done = done
        return PendingFetch(host, done, self.offsets, self.shapes)

    def __call__(self, which: Optional[str] = None):
        self.launch()
        out = self.fetch(which or ("images" if self.neglog else "intensity"))
        if len(set(self.shapes)) == 1:
            return np.stack(out)
        return out

    def close(self) -> None:
        if self._batch:
            _abi.lib().drr_batch_destroy(self._batch)
            self._batch = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PendingFetch:
    """A host copy started by ``BatchPlan.fetch_async``."""

    def __init__(self, host, done, offsets, shapes):
        self._host, self._done, self._offsets, self._shapes = host, done, offsets, shapes

    def result(self) -> List[np.ndarray]:
        self._done.synchronize()
        host = self._host.numpy()
        return [host[o:o + h * w].reshape(h, w) for o, (h, w) in zip(self._offsets, self._shapes)]


def plan_for_projector(projector, projections: Sequence[geo.CameraProjection],
                       volume_indices: Optional[Sequence[int]] = None, neglog: Optional[bool] = None,
                       tuning=None) -> BatchPlan:
    """Batch plan over an initialised ``Projector``'s volumes."""
    from .projector import build_view

    if not projector.initialized:
        raise RuntimeError("Projector has not been initialized.")
    if not hasattr(projector, "max_ray_length"):
        projector.max_ray_length = projector.source_to_detector_distance * 4
    views = []
    sdd = projector.source_to_detector_distance
    if volume_indices is None:
        # every view over all the projector's volumes, as project() renders it
        volume_indices = [0] * len(projections)
        for proj in projections:
            views.append(build_view(projector.volumes, projector.priorities, proj, projector.max_ray_length,
                                    sdd if sdd > 0 else 1.0))
    else:
        for proj, vi in zip(projections, volume_indices):
            views.append(build_view([projector.volumes[vi]], [projector.priorities[vi]], proj,
                                    projector.max_ray_length, sdd if sdd > 0 else 1.0))
    key = None
    if tuning == "auto":
        from . import autotune

        key = autotune.workload_key(views, volume_indices, [v.shape for v in projector.volumes],
                                    len(projector.all_materials), march=projector.march)
    return BatchPlan(projector._ctx, projector._vols, views, volume_indices,
                     neglog=projector.neglog if neglog is None else neglog, tuning=tuning,
                     device=projector._torch_device, tune_key=key)
