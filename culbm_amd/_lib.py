"""ctypes binding of libculbm.so (the C ABI in include/culbm.h).

The product path has no fallback: if the HIP library is missing this module
raises, it never substitutes a CPU implementation.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libculbm.so"

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_STATE = 3
ERR_UNSUPPORTED = 4


class CulbmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{_code_name(code)}: {message}")
        self.code = code
        self.message = message


def _code_name(code: int) -> str:
    return {0: "CULBM_OK", 1: "CULBM_ERR_INVALID", 2: "CULBM_ERR_HIP", 3: "CULBM_ERR_STATE",
            4: "CULBM_ERR_UNSUPPORTED"}.get(code, f"code {code}")


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("inv_tau", ctypes.c_float),
        ("stream_policy", ctypes.c_int32),
        ("opt_policy", ctypes.c_int32),
        ("inner_loop", ctypes.c_int32),
        ("block_dim", ctypes.c_int32 * 3),
        ("grid_dim", ctypes.c_int32 * 3),
        ("use_ring", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("storage_scale", ctypes.c_float),
        ("wave_depth", ctypes.c_int32),
        ("rounding", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 6),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("step", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("wave_depth", ctypes.c_int32),
        ("launches_per_pass", ctypes.c_int32),
        ("bytes_per_pop", ctypes.c_int32),
        ("storage_scale", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 8),
    ]


class AdjointSpec(ctypes.Structure):
    _fields_ = [
        ("nsteps", ctypes.c_int32),
        ("t_begin", ctypes.c_int32), ("t_end", ctypes.c_int32),
        ("checkpoint_every", ctypes.c_int32),
        ("want_f0", ctypes.c_int32),
        ("mem_cap", ctypes.c_int64),
        ("timing", ctypes.c_int32),
        ("max_slots", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class AdjointStats(ctypes.Structure):
    _fields_ = [
        ("objective", ctypes.c_double),
        ("grad_inv_tau", ctypes.c_double),
        ("forward_steps", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("slots", ctypes.c_int32),
        ("launches", ctypes.c_int32),
        ("ms_sweep", ctypes.c_float),
        ("ms_recompute", ctypes.c_float),
        ("ms_reverse", ctypes.c_float),
        ("ms_total", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 6),
    ]


_F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_U32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_U16 = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_VP = ctypes.c_void_p
_I32 = ctypes.c_int32

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "culbm_config_default": (None, [ctypes.POINTER(Config)]),
    "culbm_validate": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_char_p, ctypes.c_size_t]),
    "culbm_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(_VP)]),
    "culbm_set_state": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "culbm_init_ddf": (ctypes.c_int, [_VP, _I32]),
    "culbm_step": (ctypes.c_int, [_VP, _I32]),
    "culbm_run_window": (ctypes.c_int, [_VP, _I32, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_float)]),
    "culbm_time_steps": (ctypes.c_int, [_VP, _I32, ctypes.POINTER(ctypes.c_float)]),
    "culbm_profile_steps": (ctypes.c_int, [_VP, _I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_I32)]),
    "culbm_dump": (ctypes.c_int, [_VP]),
    "culbm_dump_deferred": (ctypes.c_int, [_VP]),
    "culbm_get_macros": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "culbm_get_ddf": (ctypes.c_int, [_VP, _F32]),
    "culbm_set_ddf": (ctypes.c_int, [_VP, _F32]),
    "culbm_get_info": (ctypes.c_int, [_VP, ctypes.POINTER(Info)]),
    "culbm_moment_sums": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "culbm_moment_sums_noalloc": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "culbm_destroy": (ctypes.c_int, [_VP]),
    "culbm_last_error": (ctypes.c_char_p, [_VP]),
    "culbm_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "culbm_adjoint": (ctypes.c_int, [_VP, ctypes.POINTER(AdjointSpec), _VP, ctypes.POINTER(AdjointStats),
                                     _VP, _VP, _VP, _VP, _VP]),
    "culbm_adjoint_plan": (ctypes.c_int, [_I32, _I32, _I32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_I32)]),
    "culbm_probe_cells": (ctypes.c_int, [_I32, _I32, ctypes.c_float, _VP, _VP, _VP, ctypes.c_int64]),
    "culbm_probe_codec": (ctypes.c_int, [_I32, _I32, ctypes.c_float, _F32, _VP, _F32, ctypes.c_int64]),
    "culbm_blocking_axis": (ctypes.c_int, [_I32, _I32, _I32, ctypes.POINTER(_I32), _VP, _I32]),
    "culbm_wave_schedule": (ctypes.c_int, [_I32, _I32, _I32, _VP, _I32, ctypes.POINTER(_I32)]),
    "culbm_probe_bandwidth": (ctypes.c_int, [_I32, _I32, ctypes.c_int64, _I32, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libculbm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m culbm_amd.build` "
                "(there is no CPU fallback for the HIP path)")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, ctx=None) -> None:
    if rc != OK:
        msg = lib().culbm_last_error(ctx)
        raise CulbmError(rc, msg.decode() if msg else "")


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)
