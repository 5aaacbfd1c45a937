"""ctypes loading of the in-tree HIP libraries (no fallback: a missing or
unloadable library raises)."""
from __future__ import annotations

import ctypes
import os

# SEISMIC_AMD_LIB_DIR selects another build of the libraries (e.g. a variant
# built with SEISMIC_AMD_BUILD_DIR) without touching the in-tree one.
LIB_DIR = os.environ.get("SEISMIC_AMD_LIB_DIR") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)


class SxcPlanDesc(ctypes.Structure):
    """Mirror of sxc_plan_desc (include/seismic_xcorr.h)."""
    _fields_ = [
        ("n_events", ctypes.c_int),
        ("n_channels", ctypes.c_int),
        ("trace_len", ctypes.c_int),
        ("padded_size", ctypes.c_int),
        ("lag_lo", ctypes.c_int),
        ("lag_hi", ctypes.c_int),
        ("threshold", ctypes.c_float),
        ("tie_tol", ctypes.c_float),
        ("demean", ctypes.c_int),
        ("form", ctypes.c_int),
    ]


class SxcStats(ctypes.Structure):
    _fields_ = [
        ("pairs", ctypes.c_ulonglong),
        ("near_ties", ctypes.c_ulonglong),
        ("tie_overflow", ctypes.c_ulonglong),
    ]


class SxcTiming(ctypes.Structure):
    _fields_ = [
        ("executes", ctypes.c_int),
        ("ms_forward", ctypes.c_double),
        ("ms_pairs", ctypes.c_double),
        ("ms_resolve", ctypes.c_double),
    ]


class SxcServerDesc(ctypes.Structure):
    _fields_ = [
        ("n_catalog", ctypes.c_int),
        ("n_channels", ctypes.c_int),
        ("catalog_len", ctypes.c_int),
        ("max_query_len", ctypes.c_int),
        ("top_k", ctypes.c_int),
        ("max_batch_events", ctypes.c_int),
        ("n_slots", ctypes.c_int),
        ("mem_budget", ctypes.c_ulonglong),
        ("demean", ctypes.c_int),
    ]


class SxcMatch(ctypes.Structure):
    _fields_ = [("value", ctypes.c_float), ("lag", ctypes.c_int), ("index", ctypes.c_int)]


class SxcServerStats(ctypes.Structure):
    _fields_ = [
        ("requests", ctypes.c_ulonglong),
        ("events", ctypes.c_ulonglong),
        ("batches", ctypes.c_ulonglong),
        ("queued", ctypes.c_ulonglong),
        ("slots", ctypes.c_int),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double),
        ("mean_requests_per_batch", ctypes.c_double),
        ("events_per_s", ctypes.c_double),
        ("slot_bytes", ctypes.c_ulonglong),
        ("offered", ctypes.c_ulonglong),
        ("rejected", ctypes.c_ulonglong),
        ("budget_ms", ctypes.c_double),
        ("model_tflops", ctypes.c_double),
        ("model_overhead_ms", ctypes.c_double),
    ]


class SxcBlockInfo(ctypes.Structure):
    _fields_ = [("index", ctypes.c_int), ("n_blocks", ctypes.c_int), ("row0", ctypes.c_int),
                ("col0", ctypes.c_int), ("rows", ctypes.c_int), ("cols", ctypes.c_int)]


# every symbol include/*.h declares, per library
EXPORTS = {
    "libseismic_xcorr.so": ["seismic_last_error", "sxc_plan_create", "sxc_plan_destroy",
                            "sxc_plan_execute", "sxc_plan_execute_reversed", "sxc_plan_execute_noalloc",
                            "sxc_plan_stats",
                            "sxc_allpairs_host", "sxc_allpairs_host_reversed",
                            "sxc_plan_enable_timing", "sxc_plan_timing", "sxc_device_count",
                            "sxc_device_alloc", "sxc_device_free", "sxc_memcpy_h2d",
                            "sxc_memcpy_d2h", "sxc_memset_d8", "sxc_stream_create",
                            "sxc_stream_destroy", "sxc_stream_synchronize",
                            "sxc_device_synchronize", "sxc_copy_kernel", "sxc_server_create",
                            "sxc_server_submit", "sxc_server_wait", "sxc_server_stats",
                            "sxc_server_set_latency_budget",
                            "sxc_server_reset_stats", "sxc_server_destroy", "sxc_tiled_create",
                            "sxc_tiled_start", "sxc_tiled_next", "sxc_tiled_stats",
                            "sxc_tiled_destroy"],
    "correlation_cuda.so": ["initialiseCUDA", "correlationCUDA", "seismic_last_error"],
    "correlation_multicuda.so": ["initialiseCUDA", "correlationCUDA", "seismic_last_error"],
    "correlation_c.so": ["correlationCPP", "seismic_last_error"],
}


def library_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


_engine = None


def load_engine() -> ctypes.CDLL:
    """Load libseismic_xcorr.so and declare its argtypes."""
    global _engine
    if _engine is not None:
        return _engine
    path = library_path("libseismic_xcorr.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m seismic_amd.build`")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.seismic_last_error.restype = ctypes.c_char_p
    lib.seismic_last_error.argtypes = []
    lib.sxc_plan_create.restype = ctypes.c_int
    lib.sxc_plan_create.argtypes = [ctypes.POINTER(SxcPlanDesc), ctypes.POINTER(vp)]
    lib.sxc_plan_destroy.restype = ctypes.c_int
    lib.sxc_plan_destroy.argtypes = [vp]
    lib.sxc_plan_execute.restype = ctypes.c_int
    lib.sxc_plan_execute.argtypes = [vp, vp, ctypes.c_longlong, ctypes.c_int, vp, vp, vp, vp, vp]
    lib.sxc_plan_execute_noalloc.restype = ctypes.c_int
    lib.sxc_plan_execute_noalloc.argtypes = [ctypes.POINTER(SxcPlanDesc), vp, ctypes.c_longlong, ctypes.c_int,
                                             vp, vp, vp, vp, vp]
    lib.sxc_plan_execute_reversed.restype = ctypes.c_int
    lib.sxc_plan_execute_reversed.argtypes = [vp, vp, vp, ctypes.c_longlong, ctypes.c_int, vp, vp, vp, vp,
                                              vp]
    lib.sxc_plan_stats.restype = ctypes.c_int
    lib.sxc_plan_stats.argtypes = [vp, ctypes.POINTER(SxcStats)]
    lib.sxc_plan_enable_timing.restype = ctypes.c_int
    lib.sxc_plan_enable_timing.argtypes = [vp, ctypes.c_int]
    lib.sxc_plan_timing.restype = ctypes.c_int
    lib.sxc_plan_timing.argtypes = [vp, ctypes.POINTER(SxcTiming)]
    u64 = ctypes.c_ulonglong
    for name, args in (("sxc_device_count", [c_int_p]),
                       ("sxc_device_alloc", [ctypes.POINTER(vp), u64]),
                       ("sxc_device_free", [vp]),
                       ("sxc_memcpy_h2d", [vp, vp, u64]),
                       ("sxc_memcpy_d2h", [vp, vp, u64]),
                       ("sxc_memset_d8", [vp, ctypes.c_int, u64]),
                       ("sxc_stream_create", [ctypes.POINTER(vp)]),
                       ("sxc_stream_destroy", [vp]),
                       ("sxc_stream_synchronize", [vp]),
                       ("sxc_device_synchronize", []),
                       ("sxc_copy_kernel", [vp, vp, u64, vp])):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    lib.sxc_server_create.restype = ctypes.c_int
    lib.sxc_server_create.argtypes = [ctypes.POINTER(SxcServerDesc), c_float_p, ctypes.POINTER(vp)]
    lib.sxc_server_submit.restype = ctypes.c_int
    lib.sxc_server_submit.argtypes = [vp, c_float_p, ctypes.c_longlong, c_int_p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_longlong)]
    lib.sxc_server_wait.restype = ctypes.c_int
    lib.sxc_server_wait.argtypes = [vp, ctypes.c_longlong, ctypes.POINTER(SxcMatch),
                                    ctypes.POINTER(ctypes.c_double)]
    lib.sxc_server_stats.restype = ctypes.c_int
    lib.sxc_server_stats.argtypes = [vp, ctypes.POINTER(SxcServerStats)]
    lib.sxc_server_set_latency_budget.restype = ctypes.c_int
    lib.sxc_server_set_latency_budget.argtypes = [vp, ctypes.c_double]
    lib.sxc_server_reset_stats.restype = ctypes.c_int
    lib.sxc_server_reset_stats.argtypes = [vp]
    lib.sxc_server_destroy.restype = ctypes.c_int
    lib.sxc_server_destroy.argtypes = [vp]
    lib.sxc_tiled_create.restype = ctypes.c_int
    lib.sxc_tiled_create.argtypes = [ctypes.POINTER(SxcPlanDesc), ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(vp)]
    lib.sxc_tiled_start.restype = ctypes.c_int
    lib.sxc_tiled_start.argtypes = [vp, vp, ctypes.c_longlong, ctypes.c_int]
    lib.sxc_tiled_next.restype = ctypes.c_int
    lib.sxc_tiled_next.argtypes = [vp, ctypes.POINTER(SxcBlockInfo), vp, vp, vp, vp, ctypes.c_longlong]
    lib.sxc_tiled_stats.restype = ctypes.c_int
    lib.sxc_tiled_stats.argtypes = [vp, ctypes.POINTER(SxcStats)]
    lib.sxc_tiled_destroy.restype = ctypes.c_int
    lib.sxc_tiled_destroy.argtypes = [vp]
    lib.sxc_allpairs_host.restype = ctypes.c_int
    lib.sxc_allpairs_host.argtypes = [ctypes.POINTER(SxcPlanDesc), c_float_p, ctypes.c_longlong,
                                      ctypes.c_int, c_float_p, c_int_p, c_float_p, c_int_p]
    lib.sxc_allpairs_host_reversed.restype = ctypes.c_int
    lib.sxc_allpairs_host_reversed.argtypes = [ctypes.POINTER(SxcPlanDesc), c_float_p, c_float_p,
                                               ctypes.c_longlong, ctypes.c_int, c_float_p, c_int_p,
                                               c_float_p, c_int_p]
    _engine = lib
    return lib


def last_error(lib=None) -> str:
    lib = lib or load_engine()
    s = lib.seismic_last_error()
    return s.decode() if s else ""
