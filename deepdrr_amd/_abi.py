"""ctypes loader for libdeepdrr_amd.so (the C ABI in include/deepdrr_amd.h).

There is no CPU fallback: if the library is missing or a call fails, a
RuntimeError says so.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

from ._abi_types import BatchTuning, Job, MCParams, MeshPrim, Scene, View

LIB_PATH = Path(__file__).resolve().parent / "libdeepdrr_amd.so"

_P = C.c_void_p
_lock = threading.Lock()
_lib = None

# name -> (restype, argtypes)
SIGNATURES = {
    "drr_last_error": (C.c_char_p, []),
    "drr_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "drr_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "drr_ctx_destroy": (C.c_int, [_P]),
    "drr_ctx_set_scene": (C.c_int, [_P, C.POINTER(Scene)]),
    "drr_ctx_set_spectrum": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "drr_ctx_last_kernel_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "drr_volume_create": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_P)]),
    "drr_volume_create_from_hu": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            _P, _P, C.POINTER(_P)]),
    "drr_volume_destroy": (C.c_int, [_P]),
    "drr_volume_download": (C.c_int, [_P, _P, _P]),
    "drr_project": (C.c_int, [_P, _P, C.POINTER(View), _P, _P, _P, _P, _P]),
    "drr_batch_create": (C.c_int, [_P, _P, C.c_int, C.POINTER(Job), C.c_int, C.POINTER(BatchTuning),
                                   _P, _P, _P, C.c_int, C.POINTER(_P)]),
    "drr_batch_launch": (C.c_int, [_P, _P]),
    "drr_batch_last_march_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "drr_batch_destroy": (C.c_int, [_P]),
    "drr_project_jobs": (C.c_int, [_P, _P, C.c_int, C.POINTER(Job), C.c_int, _P, _P, _P, _P]),
    "drr_project_jobs_noalloc": (C.c_int, [_P, _P, C.c_int, C.POINTER(Job), C.c_int, _P, _P, _P, _P]),
    "drr_adjoint_jobs": (C.c_int, [_P, _P, C.POINTER(Job), C.c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    "drr_neglog": (C.c_int, [_P, _P, _P, C.c_int, C.c_int64, C.c_float, _P]),
    "drr_add_noise": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, _P]),
    "drr_adjoint": (C.c_int, [_P, _P, C.POINTER(View), _P, _P, _P, _P, _P, _P, _P]),
    "drr_mc_scatter": (C.c_int, [_P, _P, C.POINTER(MCParams), C.c_int64, C.c_int64, C.c_uint64,
                                 _P, _P, _P, _P, _P]),
    "drr_mesh_create": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, C.POINTER(MeshPrim), C.c_int, C.c_int,
                                  C.c_int, C.POINTER(_P)]),
    "drr_mesh_destroy": (C.c_int, [_P]),
    "drr_project_meshes": (C.c_int, [_P, _P, C.POINTER(View), _P, _P, C.c_float, _P, _P, _P, _P, _P]),
    "drr_mesh_rays": (C.c_int, [_P, _P, C.POINTER(View), _P, C.c_float, _P, _P, _P, C.POINTER(C.c_int), _P]),
}


def lib() -> C.CDLL:
    """Load the HIP library once; raise RuntimeError if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m deepdrr_amd.build` (hipcc, gfx950)")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = l
        return l


def last_error() -> str:
    msg = lib().drr_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def call(name: str, *args) -> None:
    fn = getattr(lib(), name, None)
    if fn is None:
        raise RuntimeError(f"{name} is not exported by {LIB_PATH}")
    check(fn(*args), name)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().drr_device_count(C.byref(n)), "drr_device_count")
    return n.value


def ptr(t) -> int:
    """Device pointer of a torch tensor (or 0 for None)."""
    return 0 if t is None else int(t.data_ptr())


def stream_ptr(torch_stream=None) -> int:
    import torch

    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)
