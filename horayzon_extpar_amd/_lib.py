"""ctypes binding of libhorayzon.so (C-ABI of include/horayzon.h).

The library is loaded from the package directory; a missing library is an
error (there is no fallback implementation).
"""

from __future__ import annotations

import ctypes
import os

from .build import library_path

_lib = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)

HZ_OK = 0
HZ_EINVAL = -1
HZ_EDEVICE = -2
HZ_ENOMEM = -3
HZ_EKERNEL = -4

HZ_MODE_WALK = 0
HZ_MODE_CLIMB = 1
HZ_MODE_LEARNED = 2


class HzOptions(ctypes.Structure):
    _fields_ = [("search_mode", ctypes.c_int), ("leaf_size", ctypes.c_int),
                ("block_size", ctypes.c_int), ("collect_counters", ctypes.c_int),
                ("cell_order", ctypes.c_int), ("verbose", ctypes.c_int),
                ("trace_kernel", ctypes.c_int), ("refill", ctypes.c_int),
                ("bvh_morton", ctypes.c_int), ("queue_chunk", ctypes.c_int),
                ("bvh_width", ctypes.c_int), ("prologue", ctypes.c_int),
                ("whole_cells", ctypes.c_int), ("certify", ctypes.c_int),
                ("cache", ctypes.c_int), ("bvh_sah", ctypes.c_int)]


class HzStats(ctypes.Structure):
    _fields_ = [("num_rays", ctypes.c_longlong), ("node_visits", ctypes.c_longlong),
                ("tri_tests", ctypes.c_longlong), ("num_triangles", ctypes.c_longlong),
                ("num_mesh_vertices", ctypes.c_longlong), ("bvh_nodes", ctypes.c_longlong),
                ("bvh_depth", ctypes.c_int), ("kernel_error", ctypes.c_int),
                ("t_prologue_ms", ctypes.c_double), ("t_upload_ms", ctypes.c_double),
                ("t_bvh_ms", ctypes.c_double), ("t_order_ms", ctypes.c_double),
                ("t_trace_ms", ctypes.c_double), ("t_download_ms", ctypes.c_double),
                ("t_total_ms", ctypes.c_double), ("lon0", ctypes.c_double),
                ("lat0", ctypes.c_double), ("t_predict_ms", ctypes.c_double),
                ("refill_frac", ctypes.c_double), ("lanes_per_iter", ctypes.c_double),
                ("cert_walk_cells", ctypes.c_double), ("t_certify_ms", ctypes.c_double),
                ("t_cert_walk_ms", ctypes.c_double), ("t_border_walk_ms", ctypes.c_double),
                ("spec_rays", ctypes.c_double), ("t_cert_lattice_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


# (name, restype, argtypes) of every exported entry point in include/horayzon.h
_ARR6 = [c_double_p] * 5 + [c_int_p]
SIGNATURES = [
    ("horizon_svf_comp", None,
     _ARR6 + [c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    ("horizon_svf_comp_ex", ctypes.c_int,
     _ARR6 + [c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
              ctypes.POINTER(HzOptions), ctypes.POINTER(HzStats)]),
    ("horizon_svf_comp_fortran", None,
     _ARR6 + [c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
              c_int_p]),
    ("hz_last_error", ctypes.c_char_p, []),
    ("hz_cache_release", None, []),
    ("hz_default_options", None, [ctypes.POINTER(HzOptions)]),
    ("hz_domain_create", ctypes.c_int,
     _ARR6 + [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
              ctypes.POINTER(HzOptions), ctypes.POINTER(ctypes.c_void_p),
              ctypes.POINTER(HzStats)]),
    ("hz_domain_run", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
      ctypes.POINTER(HzOptions), ctypes.POINTER(HzStats)]),
    ("hz_domain_fetch", ctypes.c_int,
     [ctypes.c_void_p, c_double_p, c_double_p, ctypes.POINTER(HzStats)]),
    ("hz_domain_destroy", None, [ctypes.c_void_p]),
    ("hz_scene_export", ctypes.c_int,
     _ARR6 + [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
              ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
              ctypes.POINTER(ctypes.c_float), c_int_p, c_double_p, c_double_p,
              ctypes.POINTER(ctypes.c_float), c_double_p]),
    ("hz_domain_num_cell", ctypes.c_int, [ctypes.c_void_p]),
    ("hz_domain_slope_aspect", ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p]),
    ("hz_batch_create", ctypes.c_int,
     [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("hz_batch_run", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
      ctypes.POINTER(HzOptions), ctypes.c_int, ctypes.POINTER(HzStats)]),
    ("hz_batch_destroy", None, [ctypes.c_void_p]),
    ("hz_predictor_num_params", ctypes.c_longlong, []),
    ("hz_predictor_create", ctypes.c_int,
     [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_void_p)]),
    ("hz_predictor_destroy", None, [ctypes.c_void_p]),
    ("hz_domain_set_predictor", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("hz_domain_descriptors", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint16)]),
    ("hz_domain_predictions", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("hz_trainer_create", ctypes.c_int, [ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_void_p)]),
    ("hz_trainer_destroy", None, [ctypes.c_void_p]),
    ("hz_trainer_set_weights", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("hz_trainer_get_weights", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("hz_trainer_set_data", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint16), ctypes.POINTER(ctypes.c_float),
      ctypes.c_int]),
    ("hz_trainer_fit", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
      ctypes.POINTER(ctypes.c_float)]),
    ("hz_trainer_gradients", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint16), ctypes.POINTER(ctypes.c_float),
      ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    ("hz_trainer_predict", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint16), ctypes.c_int,
      ctypes.POINTER(ctypes.c_float)]),
    ("hz_domain_num_triangles", ctypes.c_longlong, [ctypes.c_void_p]),
    ("hz_domain_scene", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), c_int_p, c_double_p, c_double_p,
      ctypes.POINTER(ctypes.c_float)]),
    ("hz_domain_bvh", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_longlong)]),
    ("hz_sincos", ctypes.c_int,
     [c_double_p, ctypes.c_longlong, c_double_p, c_double_p, ctypes.c_int]),
    ("hz_calibrate_gather", ctypes.c_int,
     [ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, c_double_p]),
    ("hz_walk_noalloc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
      ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
]


def lib():
    """Load libhorayzon.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} not found: build it with horayzon_extpar_amd.build_library() "
                "(hipcc, gfx950); there is no CPU fallback")
        L = ctypes.CDLL(path)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().hz_last_error().decode("utf-8", "replace")


HZ_PROLOGUE = {"host": 0, "device": 1}


def default_options(**kw) -> HzOptions:
    """The library defaults with the given fields replaced (None keeps the
    default); `prologue` also takes "host" / "device"."""
    o = HzOptions()
    lib().hz_default_options(ctypes.byref(o))
    fields = {f for f, _ in HzOptions._fields_ if f != "reserved"}
    for k, v in kw.items():
        if k not in fields:
            raise TypeError(f"unknown option {k!r}")
        if v is None:
            continue
        if k == "prologue" and isinstance(v, str):
            if v not in HZ_PROLOGUE:
                raise ValueError(f"prologue must be one of {sorted(HZ_PROLOGUE)}")
            v = HZ_PROLOGUE[v]
        setattr(o, k, v)
    return o


class HzError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libhorayzon error {code}: {msg}")
        self.code = code


def check(rc: int):
    if rc != HZ_OK:
        raise HzError(rc, last_error())


def calibrate_gather(num_nodes: int, hot_nodes: int, iters: int = 2000) -> float:
    """Steps per second of dependent random 64-byte reads from `num_nodes`
    128-byte records (`hot_nodes` of them take every other step), at the
    trace kernel's occupancy: the GPU's random-gather rate that bounds the
    walk's node fetches.  Runs on the current device."""
    rate = ctypes.c_double(0.0)
    check(lib().hz_calibrate_gather(int(num_nodes), int(hot_nodes), int(iters),
                                    ctypes.byref(rate)))
    return rate.value
