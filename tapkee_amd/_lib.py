"""ctypes binding of libtapkee_amd.so (include/tapkee_amd.h).

The library is built in-tree by `python -m tapkee_amd.build`.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.

One HIP runtime per process: PyTorch ships its own libamdhip64.so.7 (same
soname as ROCm's).  A process that uses torch must import it before the
engine loads, so both bind the same runtime; `tapkee_amd.autograd` checks
this.  Processes without torch use the system ROCm runtime.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtapkee_amd.so")

# tka_status
OK, E_WRONG_PARAMETER, E_NO_DATA, E_NOT_ENOUGH_MEMORY, E_CANCELLED, E_UNSUPPORTED, E_DEVICE, E_INTERNAL = range(8)
NOALLOC_MAX_POINTS = 131072  # TKA_NOALLOC_MAX_POINTS
LOG_INFO, LOG_WARNING, LOG_DEBUG, LOG_ERROR, LOG_BENCHMARK = range(5)
LOG_LEVEL_NAMES = ("info", "warning", "debug", "error", "benchmark")
LOG_DEFAULT_LEVELS = (1 << LOG_WARNING) | (1 << LOG_ERROR)
LOG_FN = C.CFUNCTYPE(None, C.c_int32, C.c_char_p, C.c_void_p)
# tka_dist_dtype
DIST_FP32, DIST_BF16, DIST_FP8 = 0, 1, 2
DIST_NAMES = {"fp32": DIST_FP32, "bf16": DIST_BF16, "fp8": DIST_FP8}
# tka_precision
FP32, FP64 = 0, 1
# tka_y0_mode
Y0_RAND_COMPAT, Y0_SEEDED, Y0_EXPLICIT = 0, 1, 2
# tka_exact_p_mode
EXACT_P_DENSE, EXACT_P_KNN = 0, 1
# tka_dd_mode
DD_TEXTBOOK, DD_TAPKEE_COMPAT = 0, 1
# tka_kl_mode
KL_AS_REFERENCE, KL_EXACT_Z = 0, 1
MAX_KL_TRACE = 64

CANCEL_FN = C.CFUNCTYPE(C.c_int32)


class Options(C.Structure):
    _fields_ = [
        ("struct_size", C.c_int32), ("max_iter", C.c_int32), ("stop_lying_iter", C.c_int32),
        ("mom_switch_iter", C.c_int32), ("momentum", C.c_double), ("final_momentum", C.c_double),
        ("eta", C.c_double), ("exaggeration", C.c_double), ("knn_k", C.c_int32), ("dist_dtype", C.c_int32),
        ("candidates", C.c_int32), ("loop_precision", C.c_int32), ("exact_p_mode", C.c_int32),
        ("dd_mode", C.c_int32), ("y0_mode", C.c_int32), ("seed", C.c_uint32),
        ("y0", C.POINTER(C.c_double)), ("log_every", C.c_int32), ("kl_mode", C.c_int32),
        ("verbose", C.c_int32), ("device", C.c_int32), ("device_mem_cap", C.c_int64),
        ("cancel", CANCEL_FN),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("ms_normalize", C.c_double), ("ms_knn", C.c_double), ("ms_perplexity", C.c_double),
        ("ms_symmetrize", C.c_double), ("ms_loop", C.c_double), ("ms_total", C.c_double),
        ("nnz", C.c_int64), ("knn_k", C.c_int32), ("bisection_unconverged", C.c_int32),
        ("certificate_failures", C.c_int32), ("kl_count", C.c_int32),
        ("kl_iter", C.c_int32 * MAX_KL_TRACE), ("kl_value", C.c_double * MAX_KL_TRACE),
        ("kl_final_exact", C.c_double), ("ms_upload", C.c_double),
    ]


class KernelTiming(C.Structure):
    _fields_ = [
        ("launches_build", C.c_int64), ("launches_repulsion", C.c_int64), ("launches_step", C.c_int64),
        ("launches_exact", C.c_int64), ("ms_build", C.c_double), ("ms_repulsion", C.c_double),
        ("ms_step", C.c_double), ("ms_exact", C.c_double), ("interactions", C.c_int64), ("nnz", C.c_int64),
        ("n", C.c_int32), ("dims", C.c_int32), ("launches_attraction", C.c_int64), ("ms_attraction", C.c_double),
    ]


class EngineError(RuntimeError):
    """Base of the engine's errors (tapkee's exceptions surface as RuntimeError)."""

    status = E_INTERNAL


class WrongParameterError(EngineError):
    status = E_WRONG_PARAMETER


class NoDataError(EngineError):
    status = E_NO_DATA


class NotEnoughMemoryError(EngineError, MemoryError):
    status = E_NOT_ENOUGH_MEMORY


class CancelledError(EngineError):
    status = E_CANCELLED


class UnsupportedMethodError(EngineError):
    status = E_UNSUPPORTED


class DeviceError(EngineError):
    status = E_DEVICE


class MissedParameterError(EngineError):
    status = -1


class MultipleParameterError(EngineError):
    status = -2


_ERRORS = {c.status: c for c in (WrongParameterError, NoDataError, NotEnoughMemoryError, CancelledError,
                                 UnsupportedMethodError, DeviceError)}

_lib = None
loaded_after_torch = None  # whether torch's HIP runtime was already loaded when the engine loaded
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)

SYMBOLS = (
    "tka_default_options", "tka_last_error", "tka_version", "tka_tsne_run", "tka_normalize", "tka_knn",
    "tka_perplexity", "tka_perplexity_dense", "tka_symmetrize", "tka_build_p", "tka_gradient", "tka_gradient_dense", "tka_kl",
    "tka_kl_dense", "tka_update", "tka_bh_tree", "tka_gradient_dev", "tka_kl_dev", "tka_session_create",
    "tka_session_step", "tka_session_synchronize", "tka_session_embedding", "tka_session_kl", "tka_session_stats",
    "tka_session_profile", "tka_session_kernel_timing", "tka_session_destroy", "tka_session_inject_sort_fault",
    "tka_gradient_dev_noalloc", "tka_kl_grad_dev", "tka_set_logger", "tka_log", "tka_neighbors_connected",
    "tka_session_create_knn", "tka_session_create_p",
)


def lib():
    """Load the engine; raises if the HIP library has not been built."""
    global _lib, loaded_after_torch
    if _lib is not None:
        return _lib
    import sys
    loaded_after_torch = "torch" in sys.modules
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tapkee_amd.build` "
                          "(the t-SNE engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.tka_default_options.argtypes = [C.POINTER(Options)]
    L.tka_default_options.restype = None
    L.tka_last_error.restype = C.c_char_p
    L.tka_version.restype = C.c_int32
    st = C.c_int
    L.tka_tsne_run.argtypes = [_dp, C.c_int32, C.c_int32, _dp, C.c_int32, C.c_double, C.c_double,
                               C.POINTER(Options), C.POINTER(Stats)]
    L.tka_normalize.argtypes = [_dp, C.c_int32, C.c_int32]
    L.tka_knn.argtypes = [_dp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _ip, _dp, _i32p]
    L.tka_perplexity.argtypes = [_dp, C.c_int32, C.c_int32, C.c_double, _dp, _dp, _ip, _i32p]
    L.tka_perplexity_dense.argtypes = [_dp, C.c_int32, C.c_int32, C.c_double, C.c_int32, _dp]
    L.tka_symmetrize.argtypes = [_ip, _dp, C.c_int32, C.c_int32, _ip, _ip, _dp, _i64p]
    L.tka_build_p.argtypes = [_dp, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32, _ip, _ip, _dp, _i64p]
    L.tka_gradient.argtypes = [_dp, C.c_int32, C.c_int32, _ip, _ip, _dp, C.c_double, C.c_int32, _dp, _f64p]
    L.tka_gradient_dense.argtypes = [_dp, _dp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp, _f64p]
    L.tka_kl.argtypes = [_dp, C.c_int32, C.c_int32, _ip, _ip, _dp, C.c_double, C.c_int32, _f64p]
    L.tka_kl_dense.argtypes = [_dp, _dp, C.c_int32, C.c_int32, C.c_int32, _f64p]
    L.tka_update.argtypes = [_dp, _dp, _dp, _dp, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32]
    L.tka_bh_tree.argtypes = [_dp, C.c_int32, C.c_int32, _i32p, _ip, _ip, _ip, _ip, _dp, _dp, _dp]
    L.tka_gradient_dev.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double, C.c_void_p, _f64p, C.c_void_p]
    L.tka_gradient_dev_noalloc.argtypes = L.tka_gradient_dev.argtypes
    L.tka_kl_grad_dev.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_double, C.c_int32, _f64p, C.c_void_p, _f64p, C.c_void_p]
    L.tka_kl_dev.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_double, _f64p, _f64p, C.c_void_p]
    L.tka_session_create.argtypes = [_dp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                     C.POINTER(Options), C.POINTER(C.c_void_p)]
    L.tka_session_create_knn.argtypes = [_ip, _dp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                         C.POINTER(Options), C.POINTER(C.c_void_p)]
    L.tka_session_create_p.argtypes = [_ip, _ip, _dp, C.c_int32, C.c_int32, C.c_double, C.POINTER(Options),
                                       C.POINTER(C.c_void_p)]
    L.tka_session_step.argtypes = [C.c_void_p, C.c_int32]
    L.tka_session_synchronize.argtypes = [C.c_void_p]
    L.tka_session_embedding.argtypes = [C.c_void_p, _dp]
    L.tka_session_kl.argtypes = [C.c_void_p, _f64p]
    L.tka_session_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.tka_session_profile.argtypes = [C.c_void_p, C.c_int32]
    L.tka_session_kernel_timing.argtypes = [C.c_void_p, C.POINTER(KernelTiming)]
    L.tka_session_inject_sort_fault.argtypes = [C.c_void_p]
    L.tka_session_destroy.argtypes = [C.c_void_p]
    L.tka_session_destroy.restype = None
    L.tka_set_logger.argtypes = [LOG_FN, C.c_void_p, C.c_uint32]
    L.tka_set_logger.restype = None
    L.tka_log.argtypes = [C.c_int32, C.c_char_p]
    L.tka_log.restype = None
    L.tka_neighbors_connected.argtypes = [_ip, C.c_int32, C.c_int32, _i32p]
    for name in SYMBOLS:
        if name not in ("tka_default_options", "tka_last_error", "tka_version", "tka_session_destroy",
                        "tka_set_logger", "tka_log"):
            getattr(L, name).restype = st
    _lib = L
    return L


def check(status: int) -> None:
    if status == OK:
        return
    msg = lib().tka_last_error().decode(errors="replace")
    raise _ERRORS.get(status, EngineError)(msg)


def default_options() -> Options:
    o = Options()
    lib().tka_default_options(C.byref(o))
    return o
