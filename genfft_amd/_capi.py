"""ctypes binding of the C ABI declared in include/genfft_amd.h.

The shared library genfft_amd/libgenfft_amd.so is built in-tree by
``make -C genfft_amd/csrc`` (what ``__graft_entry__.build()`` runs).  There is
no fallback path: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgenfft_amd.so")

# status codes
GFFT_OK = 0
GFFT_ERR_INVALID_ARG = 1
GFFT_ERR_UNSUPPORTED = 2
GFFT_ERR_HIP = 3
GFFT_ERR_ALLOC = 4
GFFT_ERR_NO_DEVICE = 5

# kinds
GFFT_C2C = 0
GFFT_R2C = 1
GFFT_C2C_2D = 2
GFFT_R2C_2D = 3
GFFT_COLS = 4
GFFT_DIT = 5
GFFT_C2R = 6
GFFT_C2R_2D = 7

GFFT_F32 = 0
GFFT_F64 = 1
GFFT_F16 = 2
GFFT_BF16 = 3
GFFT_SCALE_GIVEN = 0
GFFT_SCALE_AUTO = 1
GFFT_FLAG_HALF = 1
GFFT_FLAG_NO_WORKSPACE = 2
GFFT_NOALLOC_SCRATCH_BYTES = 64 << 20
GFFT_FORWARD = 0
GFFT_INVERSE = 1

# Every function declared in include/genfft_amd.h (checked by the CPU tests).
EXPORTS = (
    "gfft_plan_create",
    "gfft_plan_destroy",
    "gfft_execute",
    "gfft_execute_noalloc",
    "gfft_execute_host",
    "gfft_execute_scaled",
    "gfft_execute_scrambled",
    "gfft_plan_workspace",
    "gfft_plan_extents",
    "gfft_plan_describe",
    "gfft_plan_explain",
    "gfft_plan_profile",
    "gfft_plan_profile_launch",
    "gfft_separate_2x_real",
    "gfft_interleave_real",
    "gfft_interleave_real_2d",
    "gfft_is_device_pointer",
    "gfft_error_string",
    "gfft_last_error",
    "gfft_version",
    "gfft_batch_plan_create",
    "gfft_batch_execute",
    "gfft_batch_plan_describe",
    "gfft_batch_plan_destroy",
    "gfft_autotune",
    "gfft_tuning_save",
    "gfft_tuning_load",
    "gfft_tuning_clear",
    "gfft_filter_plan_create",
    "gfft_filter_forward",
    "gfft_filter_backward",
    "gfft_filter_plan_destroy",
)

GFFT_BATCH_GRAPH = 0
GFFT_BATCH_EAGER = 1
GFFT_BATCH_PER_JOB = 2


class GfftDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("n2", ctypes.c_int64),
        ("batch", ctypes.c_int64),
        ("in_stride", ctypes.c_int64),
        ("out_stride", ctypes.c_int64),
        ("in_dist", ctypes.c_int64),
        ("out_dist", ctypes.c_int64),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class GfftError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"genfft_amd error {status}: {message}")
        self.status = status


_lib = None
_lock = threading.Lock()


def lib() -> ctypes.CDLL:
    """Load libgenfft_amd.so (once) and declare the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `make -C genfft_amd/csrc` "
                "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.gfft_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(GfftDesc)]
        L.gfft_plan_create.restype = i32
        L.gfft_plan_destroy.argtypes = [vp]
        L.gfft_execute_scaled.argtypes = [vp, vp, vp, i32, vp, i32, vp]
        L.gfft_execute_scaled.restype = i32
        L.gfft_execute_scrambled.argtypes = [vp, vp, vp, i32, vp]
        L.gfft_execute_scrambled.restype = i32
        L.gfft_plan_destroy.restype = i32
        L.gfft_execute.argtypes = [vp, vp, vp, i32, vp]
        L.gfft_execute.restype = i32
        L.gfft_execute_noalloc.argtypes = [vp, vp, vp, i32, vp]
        L.gfft_execute_noalloc.restype = i32
        L.gfft_execute_host.argtypes = [vp, vp, vp, i32]
        L.gfft_execute_host.restype = i32
        L.gfft_plan_workspace.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
        L.gfft_plan_workspace.restype = i32
        L.gfft_plan_extents.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        L.gfft_plan_extents.restype = i32
        L.gfft_plan_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.gfft_plan_describe.restype = i32
        L.gfft_plan_explain.argtypes = [ctypes.POINTER(GfftDesc), ctypes.c_char_p, ctypes.c_size_t]
        L.gfft_plan_explain.restype = i32
        L.gfft_plan_profile.argtypes = [vp, vp, vp, i32, vp, ctypes.POINTER(ctypes.c_float), i32,
                                        ctypes.POINTER(i32)]
        L.gfft_plan_profile.restype = i32
        L.gfft_plan_profile_launch.argtypes = [vp, vp, vp, i32, vp, i32, i32, ctypes.POINTER(ctypes.c_float)]
        L.gfft_plan_profile_launch.restype = i32
        L.gfft_separate_2x_real.argtypes = [i32, vp, vp, vp, i64, i64, vp]
        L.gfft_separate_2x_real.restype = i32
        L.gfft_interleave_real.argtypes = [i32, vp, vp, vp, i64, i64, vp]
        L.gfft_interleave_real.restype = i32
        L.gfft_interleave_real_2d.argtypes = [i32, vp, i64, vp, i64, vp, i64, i64, i64, vp]
        L.gfft_interleave_real_2d.restype = i32
        L.gfft_is_device_pointer.argtypes = [vp]
        L.gfft_is_device_pointer.restype = i32
        L.gfft_error_string.argtypes = [i32]
        L.gfft_error_string.restype = ctypes.c_char_p
        L.gfft_last_error.argtypes = []
        L.gfft_last_error.restype = ctypes.c_char_p
        L.gfft_version.argtypes = []
        L.gfft_version.restype = i32
        p64 = ctypes.POINTER(ctypes.c_int64)
        L.gfft_batch_plan_create.argtypes = [ctypes.POINTER(vp), i32, i64, p64, p64, p64,
                                             ctypes.POINTER(ctypes.c_int32)]
        L.gfft_batch_plan_create.restype = i32
        L.gfft_batch_execute.argtypes = [vp, vp, vp, i32, vp]
        L.gfft_batch_execute.restype = i32
        L.gfft_batch_plan_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.gfft_batch_plan_describe.restype = i32
        L.gfft_batch_plan_destroy.argtypes = [vp]
        L.gfft_batch_plan_destroy.restype = i32
        L.gfft_autotune.argtypes = [i32, i64, i64, i32, ctypes.c_char_p, ctypes.c_size_t]
        L.gfft_autotune.restype = i32
        L.gfft_tuning_save.argtypes = [ctypes.c_char_p]
        L.gfft_tuning_save.restype = i32
        L.gfft_tuning_load.argtypes = [ctypes.c_char_p]
        L.gfft_tuning_load.restype = i32
        L.gfft_tuning_clear.argtypes = []
        L.gfft_tuning_clear.restype = i32
        L.gfft_filter_plan_create.argtypes = [ctypes.POINTER(vp), i32, i64, i64]
        L.gfft_filter_plan_create.restype = i32
        L.gfft_filter_forward.argtypes = [vp, vp, vp, vp, i32, vp]
        L.gfft_filter_forward.restype = i32
        L.gfft_filter_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gfft_filter_backward.restype = i32
        L.gfft_filter_plan_destroy.argtypes = [vp]
        L.gfft_filter_plan_destroy.restype = i32
        _lib = L
        return L


def check(status: int) -> None:
    if status != GFFT_OK:
        msg = lib().gfft_last_error()
        raise GfftError(status, msg.decode() if msg else "")
