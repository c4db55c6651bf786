"""ctypes binding of libcaustica_hip.so (declarations in include/caustica_hip.h).

The library is the only compute path of this package: loading fails loudly when it
has not been built, and there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcaustica_hip.so"

CST_OK = 0
CST_ERR_ARG = 1
CST_ERR_HIP = 2
CST_ERR_UNSUPPORTED = 3
CST_ERR_STATE = 4
MAX_BATCH = 16
MAX_HARMONICS = 3
SCHEME_REFERENCE = 0
SCHEME_STAGGERED = 1

#: Every symbol include/caustica_hip.h declares (checked by the CPU test suite).
EXPORTS = (
    "cst_version",
    "cst_last_error",
    "cst_device_count",
    "cst_supported_length",
    "cst_session_create",
    "cst_session_destroy",
    "cst_session_stream",
    "cst_member_add",
    "cst_member_add_dl",
    "cst_session_create_scheme",
    "cst_session_scheme",
    "cst_member_add_staggered_dl",
    "cst_member_add_staggered",
    "cst_member_get_density_into",
    "cst_member_set_density_dl",
    "cst_member_remove",
    "cst_member_reset",
    "cst_session_trim",
    "cst_session_pooled_bytes",
    "cst_device_memory",
    "cst_diag_copy",
    "cst_member_schedule",
    "cst_member_step",
    "cst_advance",
    "cst_sync",
    "cst_session_query",
    "cst_member_peaks",
    "cst_member_result",
    "cst_member_result_into",
    "cst_member_result_export",
    "cst_member_get_state",
    "cst_member_set_state",
    "cst_member_get_state_into",
    "cst_member_set_state_dl",
    "cst_dl_alloc",
    "cst_dl_copy",
    "cst_member_state_size",
    "cst_member_get_raw",
    "cst_member_set_raw",
    "cst_session_set_variant",
    "cst_session_set_graphs",
    "cst_session_graph_stats",
    "cst_session_set_profiling",
    "cst_session_kernel_time",
    "cst_thermal_create",
    "cst_thermal_destroy",
    "cst_thermal_setup",
    "cst_thermal_run",
    "cst_thermal_set_chunk",
    "cst_thermal_set_uniform_maps",
    "cst_thermal_status",
    "cst_thermal_read",
    "cst_thermal_heat_from_member",
    "cst_thermal_read_source",
    "cst_thermal_set_profiling",
    "cst_thermal_kernel_time",
    "cst_thermal_run_noalloc",
)


class NativeError(RuntimeError):
    """A call into libcaustica_hip returned a nonzero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


class UnsupportedShape(NativeError):
    pass


class MemberDesc(C.Structure):
    _fields_ = [
        ("dx", C.c_double),
        ("pml_vox", C.c_int),
        ("pml_edge", C.c_double),
        ("dt", C.c_double),
        ("c_ref", C.c_double),
        ("spp", C.c_int),
        ("f0", C.c_double),
        ("amplitude", C.c_double),
        ("ramp_periods", C.c_double),
        ("dt_over_rho", C.c_void_p),
        ("rhoc2_dt", C.c_void_p),
        ("absorb", C.c_void_p),
        ("beta2_dt", C.c_void_p),
        ("ktab", C.c_void_p),
        ("sponge", C.c_void_p),
        ("n_src", C.c_int),
        ("src_idx", C.c_void_p),
        ("src_phase", C.c_void_p),
        ("src_scale", C.c_void_p),
        ("conv", C.c_int * 6),
        ("rec", C.c_int * 6),
        ("n_harmonics", C.c_int),
        ("harmonics", C.c_int * MAX_HARMONICS),
        ("max_periods", C.c_int),
    ]


class StaggeredDesc(C.Structure):
    """cst_staggered_desc: tables and flags of a member of the opt-in staggered scheme."""

    _fields_ = [
        ("kfac", C.c_void_p),
        ("pml", C.c_void_p),
        ("alpha_power", C.c_double),
        ("absorbing", C.c_int),
        ("src_axes", C.c_int),
    ]


class ThermalDesc(C.Structure):
    _fields_ = [
        ("dx", C.c_double),
        ("k", C.c_void_p),
        ("perf", C.c_void_p),
        ("coeff", C.c_void_p),
        ("q", C.c_void_p),
        ("theta0", C.c_void_p),
        ("dose0", C.c_void_p),
        ("boundary", C.c_int),
        ("theta_wall", C.c_float),
        ("t_ref", C.c_float),
        ("half_dt_min", C.c_double),
        ("dose", C.c_int),
    ]


class ThermalDevDesc(C.Structure):
    """cst_thermal_dev_desc: device pointers for cst_thermal_run_noalloc."""

    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("dx", C.c_double),
        ("theta", C.c_void_p),
        ("theta_max", C.c_void_p),
        ("dose", C.c_void_p),
        ("k", C.c_void_p),
        ("perf", C.c_void_p),
        ("coeff", C.c_void_p),
        ("q", C.c_void_p),
        ("boundary", C.c_int),
        ("theta_wall", C.c_float),
        ("t_ref", C.c_float),
        ("half_dt_min", C.c_double),
        ("dose_on", C.c_int),
    ]


NOALLOC_MAX_VOXELS = 1 << 24

_lib = None


def _bind(lib):
    i, v, d, p = C.c_int, C.c_void_p, C.c_double, C.c_void_p
    i64 = C.c_int64
    sig = {
        "cst_version": ([], i),
        "cst_last_error": ([C.c_char_p, C.c_size_t], i),
        "cst_device_count": ([C.POINTER(i)], i),
        "cst_supported_length": ([i, i], i),
        "cst_session_create": ([i, i, i, i, i, C.POINTER(v)], i),
        "cst_session_destroy": ([v], i),
        "cst_session_stream": ([v, C.POINTER(v)], i),
        "cst_member_add": ([v, C.POINTER(MemberDesc), C.POINTER(i)], i),
        "cst_member_add_dl": ([v, C.POINTER(MemberDesc), p, p, p, p, C.POINTER(i)], i),
        "cst_session_create_scheme": ([i, i, i, i, i, i, C.POINTER(v)], i),
        "cst_session_scheme": ([v, C.POINTER(i)], i),
        "cst_member_add_staggered_dl": ([v, C.POINTER(MemberDesc), C.POINTER(StaggeredDesc), p,
                                         C.POINTER(i)], i),
        "cst_member_add_staggered": ([v, C.POINTER(MemberDesc), C.POINTER(StaggeredDesc), p,
                                      C.POINTER(i)], i),
        "cst_member_get_density_into": ([v, i, p, p, p], i),
        "cst_member_set_density_dl": ([v, i, p, p, p], i),
        "cst_member_remove": ([v, i], i),
        "cst_member_reset": ([v, i], i),
        "cst_session_trim": ([v, C.POINTER(C.c_size_t)], i),
        "cst_session_pooled_bytes": ([v, C.POINTER(C.c_size_t)], i),
        "cst_device_memory": ([i, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], i),
        "cst_diag_copy": ([i, C.c_size_t], i),
        "cst_member_schedule": ([v, i, i64], i),
        "cst_member_step": ([v, i, C.POINTER(i64)], i),
        "cst_advance": ([v, i, C.POINTER(i), i], i),
        "cst_sync": ([v], i),
        "cst_session_query": ([v, C.POINTER(i)], i),
        "cst_member_peaks": ([v, i, i, i, p], i),
        "cst_member_result": ([v, i, i, d, p, p], i),
        "cst_member_result_into": ([v, i, i, d, p, p], i),
        "cst_member_result_export": ([v, i, i, d, i, C.POINTER(v), C.POINTER(v)], i),
        "cst_member_get_state": ([v, i, p, p, p, p, C.POINTER(i64)], i),
        "cst_member_set_state": ([v, i, i64, p, p, p, p], i),
        "cst_member_get_state_into": ([v, i, p, p, p, p, C.POINTER(i64)], i),
        "cst_member_set_state_dl": ([v, i, i64, p, p, p, p], i),
        "cst_dl_alloc": ([i, i, i, i, p, i, C.POINTER(v)], i),
        "cst_dl_copy": ([p, p], i),
        "cst_member_state_size": ([v, C.POINTER(C.c_size_t)], i),
        "cst_member_get_raw": ([v, i, p, C.c_size_t, C.POINTER(i64)], i),
        "cst_member_set_raw": ([v, i, p, C.c_size_t, i64], i),
        "cst_session_set_variant": ([v, i, i], i),
        "cst_session_set_graphs": ([v, i], i),
        "cst_session_graph_stats": ([v, C.POINTER(i), C.POINTER(i64), C.POINTER(i64)], i),
        "cst_session_set_profiling": ([v, i], i),
        "cst_session_kernel_time": ([v, i, C.POINTER(d), C.POINTER(i64)], i),
        "cst_thermal_create": ([i, i, i, i, C.POINTER(v)], i),
        "cst_thermal_destroy": ([v], i),
        "cst_thermal_setup": ([v, C.POINTER(ThermalDesc)], i),
        "cst_thermal_run": ([v, i], i),
        "cst_thermal_set_chunk": ([v, i], i),
        "cst_thermal_set_uniform_maps": ([v, i, C.POINTER(i)], i),
        "cst_thermal_status": ([v, C.POINTER(i)], i),
        "cst_thermal_read": ([v, p, p, p], i),
        "cst_thermal_heat_from_member": ([v, v, i, i, p, p, p, d, C.c_float], i),
        "cst_thermal_read_source": ([v, p], i),
        "cst_thermal_set_profiling": ([v, i], i),
        "cst_thermal_kernel_time": ([v, C.POINTER(d), C.POINTER(i64)], i),
        "cst_thermal_run_noalloc": ([C.POINTER(ThermalDevDesc), i, v, C.POINTER(i)], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


def lib():
    """The loaded library; raises if it was never built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m caustica_amd.build` "
                f"(hipcc, gfx950). caustica_amd has no CPU fallback."
            )
        _lib = _bind(C.CDLL(os.fspath(LIB_PATH)))
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().cst_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code: int) -> None:
    if code != CST_OK:
        msg = last_error()
        if code == CST_ERR_UNSUPPORTED:
            raise UnsupportedShape(code, msg)
        raise NativeError(code, msg)


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_memory(device: int = 0) -> tuple[int, int]:
    free, total = C.c_size_t(0), C.c_size_t(0)
    check(lib().cst_device_memory(int(device), C.byref(free), C.byref(total)))
    return int(free.value), int(total.value)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().cst_device_count(C.byref(n)))
    return int(n.value)
