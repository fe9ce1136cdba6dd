"""Loader for libwq3d.so with ctypes signatures; fails loudly when the
library has not been built (there is no host fallback)."""
import ctypes as C
import os

from . import _abi

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libwq3d.so")
_lib = None


class WQ3DError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"wq3d status {status}: {message}")
        self.status = status


def library_path():
    return _PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_PATH):
        raise ImportError(f"{_PATH} is missing: build it with "
                          "`python -m underseamodelinglibrary_amd.build` (HIP, gfx950)")
    _lib = bind(C.CDLL(_PATH))
    return _lib


def bind(lib):
    """Attach the C ABI signatures and a status checker to a loaded library."""
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    sigs = {
        "wq3d_last_error": ([], C.c_char_p),
        "wq3d_device_count": ([], i32),
        "wq3d_env_create": ([C.POINTER(_abi.EnvDesc), C.POINTER(vp)], i32),
        "wq3d_env_destroy": ([vp], i32),
        "wq3d_fan_create": ([vp, C.POINTER(_abi.FanDesc), C.POINTER(vp)], i32),
        "wq3d_fan_destroy": ([vp], i32),
        "wq3d_step": ([vp, i64], i32),
        "wq3d_sync": ([vp], i32),
        "wq3d_time": ([vp, dp], i32),
        "wq3d_source": ([vp, dp, dp, dp], i32),
        "wq3d_set_thresholds": ([vp, C.POINTER(_abi.Thresholds)], i32),
        "wq3d_set_listeners": ([vp, i32, i32], i32),
        "wq3d_drain_eigenrays": ([vp, C.POINTER(_abi.EigenrayRec), dp, dp, i64,
                                  C.POINTER(i64)], i32),
        "wq3d_drain_eigenverbs": ([vp, C.POINTER(_abi.EigenverbRec), dp, i64, C.POINTER(i64)],
                                  i32),
        "wq3d_drain_reflections": ([vp, C.POINTER(_abi.ReflectionRec), i64, C.POINTER(i64)], i32),
        "wq3d_snapshot": ([vp, i32, C.POINTER(_abi.FrontView)], i32),
        "wq3d_set_timing": ([vp, i32], i32),
        "wq3d_kernel_times": ([vp, dp, C.POINTER(i64), i32], i32),
        "wq3d_search_stats": ([vp, C.POINTER(i64), C.POINTER(i64)], i32),
        "wq3d_search_totals": ([vp, C.POINTER(i64), C.POINTER(i64)], i32),
        "wq3d_sum_eigenrays": ([C.c_void_p, dp, dp, i64, dp, i32, i64, i64, i32,
                                C.POINTER(_abi.EigenrayTotal), dp, dp], i32),
        "wq3d_sum_eigenrays_noalloc": ([C.c_void_p, dp, dp, i64, dp, i32, i64, i64, i32,
                                        C.POINTER(_abi.EigenrayTotal), dp, dp], i32),
        "wq3d_batch_create": ([C.POINTER(vp), i32, C.POINTER(vp)], i32),
        "wq3d_batch_destroy": ([vp], i32),
        "wq3d_batch_step": ([vp, i64], i32),
        "wq3d_batch_get_shape": ([vp, C.POINTER(_abi.BatchShape)], i32),
        "wq3d_batch_set_shape": ([vp, C.POINTER(_abi.BatchShape)], i32),
        "wq3d_batch_tune": ([vp, i32, dp, C.POINTER(i64)], i32),
        "wq3d_batch_set_timing": ([vp, i32], i32),
        "wq3d_batch_stage_times": ([vp, dp, C.POINTER(i64)], i32),
        "wq3d_adjoint_checkpoint": ([vp], i32),
        "wq3d_adjoint_gradient": ([vp, C.POINTER(_abi.EigenrayRec), dp, dp, dp, i64, dp, i64], i32),
        "wq3d_adjoint_release": ([vp], i32),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res

    def check(status):
        if status != _abi.WQ3D_OK:
            raise WQ3DError(status, lib.wq3d_last_error().decode())

    lib.check = check
    return lib
