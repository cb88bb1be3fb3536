"""ctypes declarations for libtcv.so (the C ABI in include/tcv.h).

The library is required: importing this module raises when libtcv.so is
missing.  There is no CPU fallback for any operation.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtcv.so")

P = ct.c_void_p
I = ct.c_int
D = ct.c_double
LL = ct.c_longlong
PI = ct.POINTER(ct.c_int)
PD = ct.POINTER(ct.c_double)
PLL = ct.POINTER(ct.c_longlong)

STATUS = {
    0: "TCV_OK", 1: "TCV_ERR_INVALID", 2: "TCV_ERR_RANGE", 3: "TCV_ERR_DEVICE",
    4: "TCV_ERR_NOMEM", 5: "TCV_ERR_BUDGET", 6: "TCV_ERR_UNSUPPORTED",
    7: "TCV_ERR_CANCELLED", 8: "TCV_ERR_TIMEOUT",
}


class Geometry(ct.Structure):
    _fields_ = [
        ("nVis", I), ("d_iu", P), ("d_iv", P), ("d_cOffset", P), ("d_wPlane", P),
        ("gSize", I), ("nPlanes", I), ("sSize", P), ("cSize", LL),
    ]


class Job(ct.Structure):
    _fields_ = [("plan", P), ("d_data", P), ("d_C", P), ("d_grid", P), ("d_out", P)]


class ServerConfig(ct.Structure):
    _fields_ = [("arenaBytes", ct.c_size_t), ("maxInflight", I), ("maxBatch", I),
                ("tile_w", I), ("tile_h", I), ("chunk", I), ("slice_keys", I), ("phases", I),
                ("autoRelease", I), ("maxBatchPoints", ct.c_longlong), ("item_entries", I),
                ("grid_cols", I), ("exclusivePoints", ct.c_longlong)]


class Request(ct.Structure):
    _fields_ = [("op", I), ("geom", Geometry), ("d_data", P), ("d_C", P), ("d_grid", P)]


class Result(ct.Structure):
    _fields_ = [("submitted", D), ("admitted", D), ("launched", D), ("completed", D),
                ("status", I), ("d_result", P), ("points", LL)]


class Tuning(ct.Structure):
    _fields_ = [("tile_w", I), ("tile_h", I), ("chunk", I), ("slice_keys", I),
                ("phases", I), ("xcd_split", I), ("kernel", I), ("degrid_kernel", I),
                ("grid_phases", I), ("tap_bytes", I), ("item_entries", I), ("grid_cols", I)]


# name -> (restype, argtypes)
SIGS = {
    "tcv_create": (I, [I, ct.POINTER(P)]),
    "tcv_destroy": (None, [P]),
    "tcv_last_error": (ct.c_char_p, [P]),
    "tcv_status_string": (ct.c_char_p, [I]),
    "tcv_device_cus": (I, [P]),
    "tcv_synchronize": (I, [P]),
    "tcv_malloc": (I, [P, ct.c_size_t, ct.POINTER(P)]),
    "tcv_free": (I, [P, P]),
    "tcv_upload": (I, [P, P, P, ct.c_size_t]),
    "tcv_download": (I, [P, P, P, ct.c_size_t]),
    "tcv_zero": (I, [P, P, ct.c_size_t]),
    "tcv_timer_start": (I, [P]),
    "tcv_timer_stop": (I, [P, PD]),
    "tcv_set_profiling": (I, [P, I]),
    "tcv_kernel_stats": (I, [P, ct.c_char_p, ct.POINTER(ct.c_long), PD]),
    "tcv_reset_stats": (I, [P]),
    "tcv_random_int": (I, [ct.POINTER(ct.c_uint64)]),
    "tcv_attic_uvw": (None, [I, I, ct.POINTER(ct.c_uint64), P, P, P]),
    "tcv_attic_initC": (I, [D, D, I, I, I, PI, PD, P]),
    "tcv_current_params": (I, [I, PI, PD, PI, PI, PI, PD, ct.POINTER(ct.c_float)]),
    "tcv_current_uvw": (I, [ct.c_float, ct.POINTER(ct.c_uint64), P, P, P]),
    "tcv_current_initC": (I, [D, I, I, I, D, P, P, P, PLL]),
    "tcv_init_offsets": (I, [P, P, P, P, I, P, I, D, D, I, I, I, P, P, P, P, P, P, P, PLL]),
    "tcv_grid_plain": (I, [P, ct.POINTER(Geometry), P, P, P]),
    "tcv_degrid_plain": (I, [P, ct.POINTER(Geometry), P, P, P]),
    "tcv_grid_plain_f64": (I, [P, ct.POINTER(Geometry), P, P, P]),
    "tcv_degrid_plain_f64": (I, [P, ct.POINTER(Geometry), P, P, P]),
    "tcv_grid_noalloc": (I, [P, ct.POINTER(Geometry), P, P, P]),
    "tcv_range_errors": (I, [P, PLL]),
    "tcv_plan_create": (I, [P, ct.POINTER(Geometry), ct.POINTER(Tuning), ct.POINTER(P)]),
    "tcv_plan_build": (I, [P, P, ct.POINTER(Geometry)]),
    "tcv_plan_destroy": (I, [P, P]),
    "tcv_plan_info": (I, [P, P, ct.POINTER(Tuning), PLL, PLL, PLL]),
    "tcv_grid_tiled": (I, [P, P, P, P, P]),
    "tcv_degrid_tiled": (I, [P, P, P, P, P]),
    "tcv_ctable_to_half": (I, [P, P, LL, I, P, LL, P, P, P]),
    "tcv_vis_to_half": (I, [P, P, I, P, ct.POINTER(ct.c_float)]),
    "tcv_grid_tiled_half": (I, [P, P, P, ct.c_float, P, P, LL, P]),
    "tcv_degrid_tiled_half": (I, [P, P, P, P, P, LL, P]),
    "tcv_batch_create": (I, [P, I, ct.POINTER(Job), ct.POINTER(P)]),
    "tcv_batch_destroy": (I, [P, P]),
    "tcv_batch_rebuild": (I, [P, P]),
    "tcv_batch_grid": (I, [P, P]),
    "tcv_batch_degrid": (I, [P, P]),
    "tcv_batch_run_graph": (I, [P, P, I]),
    "tcv_attic_initC_device": (I, [P, D, D, I, I, I, PI, PD, P]),
    "tcv_mem_info": (I, [P, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t)]),
    "tcv_copy": (I, [P, P, P, ct.c_size_t]),
    "tcv_read_bandwidth": (I, [P, P, ct.c_size_t, ct.c_int, ct.POINTER(ct.c_double)]),
    "tcv_grid_psf": (I, [P, P, P, P]),
    "tcv_fft_exec": (I, [P, P, I, I]),
    "tcv_fft_fix": (I, [P, P, LL, ct.c_float]),
    "tcv_fft_exec_fix": (I, [P, P, I, I, ct.c_float]),
    "tcv_find_peak": (I, [P, P, LL, ct.POINTER(ct.c_float), ct.POINTER(LL)]),
    "tcv_hogbom": (I, [P, P, LL, P, LL, P, I, ct.c_float, ct.POINTER(ct.c_float), ct.POINTER(LL),
                       ct.POINTER(ct.c_float), ct.POINTER(LL)]),
    "tcv_vis_subtract": (I, [P, P, P, LL]),
    "tcv_autotune": (I, [P, ct.POINTER(Geometry), P, P, P, P, I, ct.POINTER(Tuning), I, I,
                         ct.POINTER(Tuning), PD]),
    "tcv_server_create": (I, [P, ct.POINTER(ServerConfig), ct.POINTER(P)]),
    "tcv_server_destroy": (I, [P]),
    "tcv_server_footprint": (I, [P, ct.POINTER(Request), ct.POINTER(ct.c_size_t)]),
    "tcv_server_submit": (I, [P, ct.POINTER(Request), PLL]),
    "tcv_server_cancel": (I, [P, LL]),
    "tcv_server_wait": (I, [P, LL, D, ct.POINTER(Result)]),
    "tcv_server_release": (I, [P, LL]),
    "tcv_server_stats": (I, [P, PLL, PLL, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t)]),
    "tcv_server_clock": (D, [P]),
    "tcv_server_last_error": (ct.c_char_p, [P]),
}

_lib = None


def lib():
    """Load libtcv.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                "g.build()' or make -C askap-benchmarks_amd/csrc). There is no CPU fallback.")
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names of the C ABI entry points this binding declares."""
    return sorted(SIGS)
