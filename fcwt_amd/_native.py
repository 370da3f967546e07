"""ctypes binding of libfcwt_hip.so (the C ABI in include/fcwt_hip.h).

The library is loaded from the package's own ``_lib`` folder; if it is not
built, every call raises ``ImportError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import pathlib
import threading

LIB_PATH = pathlib.Path(__file__).resolve().parent / "_lib" / "libfcwt_hip.so"

OK = 0
EINVAL = -1
ENOMEM = -2
EDEVICE = -3
EUNSUPPORTED = -4
EIO = -5

COVERAGE_SCALAR = 0x1

# every symbol include/fcwt_hip.h declares, with (restype, argtypes)
_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_plan_t = ctypes.c_void_p
SIGNATURES = {
    "fcwt_hip_last_error": (ctypes.c_char_p, []),
    "fcwt_hip_version": (ctypes.c_char_p, []),
    "fcwt_hip_find2power": (ctypes.c_int, [ctypes.c_int]),
    "fcwt_hip_scales": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                       _c_float_p]),
    "fcwt_hip_frequencies": (ctypes.c_int, [_c_float_p, ctypes.c_int, ctypes.c_int, _c_float_p]),
    "fcwt_hip_morlet_generate": (ctypes.c_int, [ctypes.c_float, ctypes.c_int, _c_float_p]),
    "fcwt_hip_morlet_support": (ctypes.c_int, [ctypes.c_float, ctypes.c_float]),
    "fcwt_hip_morlet_generate_time": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, ctypes.c_int, _c_float_p,
                                                     _c_float_p]),
    "fcwt_hip_morlet_wavelet": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, _c_float_p, ctypes.c_int]),
    "fcwt_hip_describe_rows": (ctypes.c_int, [ctypes.c_int, _c_float_p, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                              ctypes.c_uint, _c_int_p, _c_int_p, _c_int_p]),
    "fcwt_hip_plan_create": (ctypes.c_int, [ctypes.POINTER(_plan_t), ctypes.c_int, _c_float_p, ctypes.c_int,
                                            ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_uint]),
    "fcwt_hip_plan_destroy": (ctypes.c_int, [_plan_t]),
    "fcwt_hip_plan_shape": (ctypes.c_int, [_plan_t, _c_int_p, _c_int_p, _c_int_p]),
    "fcwt_hip_plan_device_bytes": (ctypes.c_int, [_plan_t, ctypes.POINTER(ctypes.c_size_t)]),
    "fcwt_hip_plan_check": (ctypes.c_int, [_plan_t]),
    "fcwt_hip_plan_profile": (ctypes.c_int, [_plan_t, ctypes.c_int]),
    "fcwt_hip_plan_profile_read": (ctypes.c_int, [_plan_t, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                                  ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_longlong), _c_int_p]),
    "fcwt_hip_plan_row_split": (ctypes.c_int, [_plan_t, _c_int_p, _c_int_p, _c_int_p]),
    "fcwt_hip_cwt": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p]),
    "fcwt_hip_cwt_device": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fcwt_hip_cwt_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_float_p, ctypes.c_int,
                                                   ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                                   ctypes.c_void_p, ctypes.c_void_p]),
    "fcwt_hip_ccwt": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p]),
    "fcwt_hip_ccwt_device": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fcwt_hip_adjoint_device": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                               ctypes.c_void_p]),
    "fcwt_hip_adjoint": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "fcwt_hip_adjoint_batch_device": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_void_p]),
    "fcwt_hip_spectrum_device": (ctypes.c_int, [_plan_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_void_p]),
    "fcwt_hip_tune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
    "fcwt_hip_tune_load": (ctypes.c_int, [ctypes.c_char_p]),
    "fcwt_hip_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _c_int_p,
                                             ctypes.POINTER(_c_float_p), _c_int_p, ctypes.c_float, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_uint]),
    "fcwt_hip_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fcwt_hip_batch_execute_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                                     ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]),
    "fcwt_hip_batch_check": (ctypes.c_int, [ctypes.c_void_p]),
    "fcwt_hip_batch_shape": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p]),
    "fcwt_hip_batch_row_launches": (ctypes.c_int, [ctypes.c_void_p, _c_int_p]),
    "fcwt_hip_batch_graph_launches": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]),
}

_lib = None
_lock = threading.Lock()


class FcwtError(RuntimeError):
    """A libfcwt_hip call failed; ``code`` is its status."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise ImportError(
                        f"{LIB_PATH} is not built; run `python -m fcwt_amd.build` (the GPU library is required, "
                        "there is no CPU fallback)")
                # torch ships its own HIP runtime under the same soname
                # (libamdhip64.so.7) as /opt/rocm's; whichever loads first
                # serves the whole process, and torch cannot find a GPU when
                # the other one was loaded before it.  Loading torch first
                # keeps one runtime, torch's, for both.
                try:
                    import torch  # noqa: F401
                except ImportError:
                    pass
                L = ctypes.CDLL(str(LIB_PATH))
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc == OK:
        return
    msg = f"{what}: {lib().fcwt_hip_last_error().decode(errors='replace')}"
    if rc == EINVAL:
        raise ValueError(msg)
    raise FcwtError(rc, msg)


def fptr(a):
    return a.ctypes.data_as(_c_float_p)
