"""ctypes binding of the gfx950 library (hydroflux_amd/lib/libhf2d.so).

The library is built in-tree by hydroflux_amd.build (hipcc, gfx950). There is
no CPU fallback: if the library is missing or cannot be loaded, importing the
solver raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhf2d.so")

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)


class HfPointSource(C.Structure):
    _fields_ = [("row", C.c_int64), ("col", C.c_int64), ("q_mass", C.c_double)]


class HfConfig(C.Structure):
    _fields_ = [
        ("cfl", C.c_double),
        ("integrator", C.c_int32),
        ("cfl_with_bcs", C.c_int32),
        ("max_dt", C.c_double),
        ("friction", C.c_int32),
        ("options", C.c_int32),
        ("friction_dry_tol", C.c_double),
        ("max_steps", C.c_int64),
        ("rain_rate", C.c_double),
    ]


# Every symbol include/hydroflux_amd.h declares, with its ctypes signature.
SIGNATURES = {
    "hf2d_create": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int32, C.POINTER(C.c_void_p)]),
    "hf2d_destroy": (C.c_int, [C.c_void_p]),
    "hf2d_last_error": (C.c_char_p, [C.c_void_p]),
    "hf2d_upload_mesh": (C.c_int, [C.c_void_p, _D, _D, _D, _D]),
    "hf2d_upload_state": (C.c_int, [C.c_void_p, _D, _D]),
    "hf2d_download_state": (C.c_int, [C.c_void_p, _D, _D]),
    "hf2d_set_boundaries": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), _D]),
    "hf2d_max_wave_speeds": (C.c_int, [C.c_void_p, _D, _D]),
    "hf2d_cfl_time_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _D]),
    "hf2d_forward_euler_step": (C.c_int, [C.c_void_p, C.c_double]),
    "hf2d_ssprk2_step": (C.c_int, [C.c_void_p, C.c_double]),
    "hf2d_manning_friction_step": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "hf2d_apply_rain": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "hf2d_apply_point_sources": (C.c_int, [C.c_void_p, C.POINTER(HfPointSource), C.c_int64, C.c_double]),
    "hf2d_configure": (C.c_int, [C.c_void_p, C.POINTER(HfConfig), C.POINTER(HfPointSource), C.c_int64]),
    "hf2d_step": (C.c_int, [C.c_void_p, C.c_double, _D]),
    "hf2d_set_forcing_series": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _D, _D, C.c_double, C.c_double,
                                          C.c_double]),
    "hf2d_set_gauges": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int64]),
    "hf2d_read_gauges": (C.c_int, [C.c_void_p, C.c_int64, _D, C.POINTER(C.c_int64)]),
    "hf2d_run_until": (C.c_int, [C.c_void_p, C.c_double]),
    "hf2d_run_steps": (C.c_int, [C.c_void_p, C.c_int64]),
    "hf2d_get_time": (C.c_int, [C.c_void_p, _D, _I64]),
    "hf2d_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_int64]),
    "hf2d_enqueue_steps": (C.c_int, [C.c_void_p, C.c_int64]),
    "hf2d_synchronize": (C.c_int, [C.c_void_p]),
    "hf2d_set_kernel_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "hf2d_kernel_timing": (C.c_int, [C.c_void_p, _D, _I64]),
    "hf2d_total_volume": (C.c_int, [C.c_void_p, _D]),
    "hf2d_total_volume_noalloc": (C.c_int, [C.c_void_p, _D]),
    "hf2d_copy_probe": (C.c_int, [C.c_int64, C.c_int32, _D]),
    "hf2d_math_probe": (C.c_int, [_D, _D, C.c_int64, _D]),
    "hf2d_batch_create": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "hf2d_batch_destroy": (C.c_int, [C.c_void_p]),
    "hf2d_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "hf2d_batch_footprint": (C.c_int, [C.c_int32, C.c_int64, C.c_int64, C.c_int64, _I64]),
    "hf2d_batch_free_bytes": (C.c_int, [C.c_void_p, _I64, _I64]),
    "hf2d_batch_add": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_double, _D, _D, _D, _D, _D,
                                 C.POINTER(C.c_int32), _D, C.POINTER(HfConfig), C.POINTER(HfPointSource),
                                 C.c_int64, C.c_double, C.POINTER(C.c_int32)]),
    "hf2d_batch_remove": (C.c_int, [C.c_void_p, C.c_int32]),
    "hf2d_batch_step": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32]),
    "hf2d_batch_synchronize": (C.c_int, [C.c_void_p]),
    "hf2d_batch_query": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "hf2d_batch_poll": (C.c_int, [C.c_void_p, C.c_int32, _D, _I64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]),
    "hf2d_batch_download": (C.c_int, [C.c_void_p, C.c_int32, _D, _D]),
    "hf2d_batch_download_done": (C.c_int, [C.c_void_p, C.c_int32, _D, _D]),
    "hf2d_batch_config_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "hf2d_batch_config_info": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]),
    "hf2d_batch_set_config": (C.c_int, [C.c_void_p, C.c_int32]),
    "hf2d_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), _D]),
}

_LIB = None


def lib():
    """Load libhf2d.so once; raise if it is absent (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"hydroflux_amd: gfx950 library not built ({LIB_PATH}); run "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB
