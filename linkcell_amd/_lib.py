"""ctypes binding of the in-tree HIP library (``liblinkcell_amd.so``).

Loading fails loudly when the library has not been built: there is no
CPU fallback behind any entry point of this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblinkcell_amd.so"

# Every symbol the public headers under include/ declare.
EXPORTS = (
    # include/linkcell.h
    "lc_knearest", "lc_knearest_d2", "lc_knearest_many", "lc_last_error",
    "lc_version",
    # include/linkcell_gpu.h
    "lc_gpu_available", "lc_gpu_last_error", "lc_gpu_workspace_new",
    "lc_gpu_workspace_free", "lc_gpu_knearest", "lc_gpu_knearest_many",
    "lc_gpu_knearest_many_dcell", "lc_gpu_queue", "lc_gpu_wait",
    "lc_gpu_alloc", "lc_gpu_free", "lc_gpu_fill_i32", "lc_gpu_memcpy",
    # include/linkcell_amd.h
    "lc_gpu_knearest_ex", "lc_gpu_set_mode", "lc_gpu_set_timing",
    "lc_gpu_last_timing", "lc_gpu_set_dense_min", "lc_gpu_set_stats",
    "lc_gpu_last_stats", "lc_gpu_knearest_batch", "lc_gpu_graph_capture",
    "lc_gpu_graph_launch", "lc_gpu_graph_reset", "lc_gpu_autotune",
    "lc_gpu_set_schedule", "lc_gpu_sched_set_prediction",
    "lc_gpu_sched_set_tile_prediction", "lc_gpu_sched_tile_count",
    "lc_gpu_sched_tile_labels", "lc_gpu_last_retried",
    "lc_gpu_sched_collect", "lc_gpu_sched_samples", "lc_gpu_sched_clear",
    "lc_gpu_sched_train", "lc_gpu_sched_get_params", "lc_gpu_sched_set_params",
    "lc_gpu_sched_forward", "lc_gpu_sched_gradient", "lc_gpu_sched_adam",
    "lc_gpu_last_timing_ex", "lc_gpu_sched_dump", "lc_gpu_timing_history",
    "lc_gpu_knearest_noalloc", "lc_gpu_noalloc_scratch_bytes",
    "lc_gpu_set_dcell_capacity", "lc_gpu_dcell_plan",
    "lc_gpu_sched_set_source_prediction", "lc_gpu_sched_source_labels",
)


class lc_cell(C.Structure):
    """Mirror of ``struct lc_cell`` (rows a, b, c then origin)."""

    _fields_ = [(name, C.c_double) for name in (
        "ax", "ay", "az", "bx", "by", "bz", "cx", "cy", "cz", "ox", "oy", "oz")]

    @classmethod
    def from12(cls, v):
        v = [float(x) for x in v]
        if len(v) != 12:
            raise ValueError("lc_cell needs 12 values")
        return cls(*v)


class lc_gpu_frame(C.Structure):
    """Mirror of ``struct lc_gpu_frame``: one system of a batched search."""

    _fields_ = [
        ("xyz", C.c_void_p), ("n", C.c_size_t), ("cell", lc_cell),
        ("k", C.c_size_t), ("mask", C.c_void_p), ("cell_hint", C.c_double),
        ("out_nn", C.c_void_p), ("out_d2", C.c_void_p), ("out_reach", C.c_void_p),
    ]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_sz = C.c_size_t

_PROTOS = {
    "lc_knearest": (C.c_int, [_vp, _sz, C.POINTER(lc_cell), _sz, _vp, C.c_double, _vp]),
    "lc_knearest_d2": (C.c_int, [_vp, _sz, C.POINTER(lc_cell), _sz, _vp, C.c_double, _vp, _vp]),
    "lc_knearest_many": (C.c_int, [_vp, _sz, _sz, C.POINTER(lc_cell), _sz, _vp, C.c_double, _vp, _vp]),
    "lc_last_error": (C.c_char_p, []),
    "lc_version": (C.c_char_p, []),
    "lc_gpu_available": (C.c_int, []),
    "lc_gpu_last_error": (C.c_char_p, []),
    "lc_gpu_workspace_new": (_vp, []),
    "lc_gpu_workspace_free": (None, [_vp]),
    "lc_gpu_knearest": (C.c_int, [_vp, _vp, _sz, C.POINTER(lc_cell), _sz, _vp, C.c_double, _vp, _vp]),
    "lc_gpu_knearest_many": (C.c_int, [_vp, _vp, _sz, _sz, C.POINTER(lc_cell), _sz, _vp,
                                       C.c_double, _vp, _vp, C.c_int]),
    "lc_gpu_knearest_many_dcell": (C.c_int, [_vp, _vp, _sz, _sz, _vp, C.c_int, _sz, _vp,
                                             C.c_double, _vp, _vp, C.c_int]),
    "lc_gpu_queue": (_vp, [_vp]),
    "lc_gpu_wait": (None, [_vp]),
    "lc_gpu_alloc": (C.c_int, [C.POINTER(_vp), _sz]),
    "lc_gpu_free": (None, [_vp]),
    "lc_gpu_fill_i32": (C.c_int, [_vp, C.c_int, _sz]),
    "lc_gpu_memcpy": (C.c_int, [_vp, _vp, _sz, C.c_int]),
    "lc_gpu_knearest_ex": (C.c_int, [_vp, _vp, _sz, _sz, C.POINTER(lc_cell), _sz, _vp,
                                     C.c_double, _vp, _vp, _vp, C.c_int]),
    "lc_gpu_set_mode": (C.c_int, [_vp, C.c_int]),
    "lc_gpu_set_timing": (C.c_int, [_vp, C.c_int]),
    "lc_gpu_last_timing": (C.c_int, [_vp, _dp, _ip]),
    "lc_gpu_set_dense_min": (C.c_int, [_vp, C.c_int]),
    "lc_gpu_set_stats": (C.c_int, [_vp, C.c_int]),
    "lc_gpu_last_stats": (C.c_int, [_vp, C.POINTER(C.c_ulonglong), C.c_int]),
    "lc_gpu_knearest_batch": (C.c_int, [_vp, C.POINTER(lc_gpu_frame), _sz, C.c_int]),
    "lc_gpu_graph_capture": (C.c_int, [_vp, C.POINTER(lc_gpu_frame), _sz]),
    "lc_gpu_graph_launch": (C.c_int, [_vp, C.c_int]),
    "lc_gpu_graph_reset": (C.c_int, [_vp]),
    "lc_gpu_autotune": (C.c_int, [_vp, C.POINTER(lc_gpu_frame), _sz, _ip, _dp]),
    "lc_gpu_set_schedule": (C.c_int, [_vp, C.c_int]),
    "lc_gpu_sched_set_prediction": (C.c_int, [_vp, _vp, _sz]),
    "lc_gpu_sched_set_tile_prediction": (C.c_int, [_vp, _vp, _sz]),
    "lc_gpu_sched_tile_count": (C.c_int, [_vp, C.POINTER(_sz)]),
    "lc_gpu_sched_tile_labels": (C.c_int, [_vp, _vp, _sz]),
    "lc_gpu_last_retried": (C.c_int, [_vp, _ip]),
    "lc_gpu_sched_dump": (C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lc_gpu_timing_history": (C.c_int, [_vp, _dp, C.c_int, _ip]),
    "lc_gpu_knearest_noalloc": (C.c_int, [_vp, _sz, _sz, C.POINTER(lc_cell), _sz, _vp,
                                          C.c_double, _vp, _vp, _vp]),
    "lc_gpu_noalloc_scratch_bytes": (_sz, []),
    "lc_gpu_set_dcell_capacity": (C.c_int, [_vp, _sz]),
    "lc_gpu_sched_set_source_prediction": (C.c_int, [_vp, _vp, _sz]),
    "lc_gpu_sched_source_labels": (C.c_int, [_vp, _vp, _sz]),
    "lc_gpu_dcell_plan": (C.c_int, [_vp, _vp, C.c_int, C.c_double, _dp]),
    "lc_gpu_last_timing_ex": (C.c_int, [_vp, _dp, C.c_int]),
    "lc_gpu_sched_collect": (C.c_int, [_vp, C.POINTER(_sz)]),
    "lc_gpu_sched_samples": (C.c_int, [_vp, C.POINTER(_sz)]),
    "lc_gpu_sched_clear": (C.c_int, [_vp]),
    "lc_gpu_sched_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_uint,
                                     C.c_double, _dp, _dp]),
    "lc_gpu_sched_get_params": (C.c_int, [_vp, _vp, _sz, _dp]),
    "lc_gpu_sched_set_params": (C.c_int, [_vp, _vp, _sz, C.c_double]),
    "lc_gpu_sched_forward": (C.c_int, [_vp, _vp, _sz, _vp]),
    "lc_gpu_sched_gradient": (C.c_int, [_vp, _vp, _vp, _vp, _sz, _vp, _dp]),
    "lc_gpu_sched_adam": (C.c_int, [_vp, _vp, _vp, _vp, _sz, C.c_int, C.c_double, _dp]),
}

_lib = None


def load():
    """Load the library once; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m linkcell_amd.build` "
            "(the HIP library is the only implementation)")
    # PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.
    # Loading it first makes the dynamic linker bind this library to the
    # same runtime (same sonames); loading /opt/rocm's copy first would put
    # two HIP runtimes in the process and torch's device init then fails.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
