"""ctypes binding of libwfs.so (declarations in include/wfs.h).

The library is the only compute path: if it is not built, or no GPU is
visible, creating a solver raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
# WFS_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("WFS_LIB") or os.path.join(PKG, "libwfs.so")
HEADER = os.path.join(ROOT, "include", "wfs.h")

WFS_OK, WFS_EINVAL, WFS_ENOMEM, WFS_EHIP, WFS_ENODEV, WFS_ENONFINITE = 0, -1, -2, -3, -4, -5
STORAGE = {"f32": 0, "f16dev": 1, "bf16dev": 2, "f64": 3, "f32fast": 4}
FIELD = {"rho": 0, "ux": 1, "uy": 2, "uz": 3, "sum_ux": 4, "sum_uy": 5, "sum_uz": 6, "sum_rho": 7,
         "sum_sp": 8, "sum_uxx": 9, "one": 10}


class WfsError(RuntimeError):
    """Failure reported by libwfs.so (status code and message)."""

    def __init__(self, code, msg):
        super().__init__(f"libwfs: {msg} (status {code})")
        self.code = code


class NonFiniteState(WfsError):
    """The watchdog found NaN/Inf in the solver state."""


class wfs_desc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("tau0", ctypes.c_double), ("cs", ctypes.c_double), ("storage", ctypes.c_int32),
                ("block_x", ctypes.c_int32), ("device", ctypes.c_int32)]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_SIGS = {
    "wfs_abi_version": ([], ctypes.c_int),
    "wfs_device_count": ([_P], ctypes.c_int),
    "wfs_last_error": ([], ctypes.c_char_p),
    "wfs_create": ([_P, _P], ctypes.c_int),
    "wfs_destroy": ([_P], ctypes.c_int),
    "wfs_upload": ([_P, _P, _P, _P], ctypes.c_int),
    "wfs_set_force": ([_P, _P, _P, _P], ctypes.c_int),
    "wfs_set_force_scale": ([_P, ctypes.c_float], ctypes.c_int),
    "wfs_set_uin": ([_P, _P], ctypes.c_int),
    "wfs_set_uref": ([_P, _P], ctypes.c_int),
    "wfs_set_bedwall": ([_P, _P], ctypes.c_int),
    "wfs_update_uref": ([_P], ctypes.c_int),
    "wfs_step": ([_P, _I32], ctypes.c_int),
    "wfs_synchronize": ([_P], ctypes.c_int),
    "wfs_nstep": ([_P], ctypes.c_int64),
    "wfs_get_macro": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "wfs_get_macro_f64": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "wfs_get_uref": ([_P, _P], ctypes.c_int),
    "wfs_set_macro": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "wfs_get_populations": ([_P, _P], ctypes.c_int),
    "wfs_get_populations_f64": ([_P, _P], ctypes.c_int),
    "wfs_set_populations": ([_P, _P], ctypes.c_int),
    "wfs_accumulate": ([_P], ctypes.c_int),
    "wfs_reset_sums": ([_P], ctypes.c_int),
    "wfs_get_sums": ([_P] * 8, ctypes.c_int),
    "wfs_get_sums_f64": ([_P] * 8, ctypes.c_int),
    "wfs_profile_x": ([_P, _I32, _I32, _I32, _P], ctypes.c_int),
    "wfs_total": ([_P, _I32, _P], ctypes.c_int),
    "wfs_check_finite": ([_P], ctypes.c_int),
    "wfs_timing_enable": ([_P, _I32], ctypes.c_int),
    "wfs_timing_read": ([_P, _P, _P], ctypes.c_int),
    "wfs_batch_create": ([_P, _I32, _P], ctypes.c_int),
    "wfs_batch_destroy": ([_P], ctypes.c_int),
    "wfs_batch_set_schedule": ([_P, _I32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double],
                               ctypes.c_int),
    "wfs_batch_set_budget": ([_P, _I32, ctypes.c_int64], ctypes.c_int),
    "wfs_batch_budget": ([_P, _I32, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "wfs_batch_run": ([_P, _I32, _I32], ctypes.c_int),
    "wfs_batch_synchronize": ([_P], ctypes.c_int),
    "wfs_batch_time": ([_P, _I32, _P, _P], ctypes.c_int),
    "wfs_bandwidth_probe": ([_I32, ctypes.c_int64, _I32, _P], ctypes.c_int),
    "wfs_stream_probe": ([_I32, _I32, _I32, _I32, _I32, ctypes.c_int64, _I32, _P], ctypes.c_int),
    "wfs_device_bytes": ([_P], ctypes.c_int64),
    "wfs_create_tiled": ([_P, _P, _P], ctypes.c_int),
    "wfs_stored_cells": ([_P], ctypes.c_int64),
    "wfs_set_force_cells": ([_P, ctypes.c_int64, _P, _P, _P, _P], ctypes.c_int),
    "wfs_set_body_mask": ([_P, _P], ctypes.c_int),
    "wfs_set_column_u": ([_P, ctypes.c_int32], ctypes.c_int),
    "wfs_set_inplace": ([_P, ctypes.c_int32], ctypes.c_int),
    "wfs_set_pop_stagger": ([_P, ctypes.c_int64], ctypes.c_int),
    "wfs_pop_stride": ([_P], ctypes.c_int64),
    "wfs_state_bytes": ([_P, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "wfs_state_save": ([_P, ctypes.c_void_p, ctypes.c_int64], ctypes.c_int),
    "wfs_state_load": ([_P, ctypes.c_void_p, ctypes.c_int64], ctypes.c_int),
    "wfs_particles_set": ([_P, ctypes.c_int64, _P, ctypes.c_int64, _P, ctypes.c_uint64], ctypes.c_int),
    "wfs_particles_move": ([_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P], ctypes.c_int),
    "wfs_particles_get": ([_P, _P], ctypes.c_int),
    "wfs_forces": ([_P, _P], ctypes.c_int),
    "wfs_forces_noalloc": ([_P, _P], ctypes.c_int),
    "wfs_set_block_x": ([_P, _I32], ctypes.c_int),
    "wfs_block_x": ([_P], _I32),
    "wfs_set_closure": ([_P, _I32, _P, _P, _P, ctypes.c_double], ctypes.c_int),
    "wfs_get_closure_field": ([_P, _P], ctypes.c_int),
    "wfs_set_closure_field": ([_P, ctypes.c_int32], ctypes.c_int),
    "wfs_closure_samples": ([_P, _P, _P, _I32, _I32, _I32, ctypes.c_double, ctypes.c_double, _P, _P, _P],
                            ctypes.c_int),
    "wfs_mlp_work_floats": ([ctypes.c_int64], ctypes.c_int64),
    "wfs_mlp_eval": ([_I32, _P, ctypes.c_int64, _P, _P, _P, ctypes.c_int64, _P], ctypes.c_int),
    "wfs_mlp_grad": ([_I32, _P, _P, ctypes.c_int64, _P, _P, _P, _P, ctypes.c_int64, _P], ctypes.c_int),
    "wfs_mlp_adam": ([_I32, _P, _P, _P, _P, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                      ctypes.c_double, _P], ctypes.c_int),
}

CLOSURE = {"smagorinsky": 0, "mlp": 1}
MLP_PARAMS = 1377
MLP_FEATURES = 8
C2_MIN = 0.01
C2_SCALE = 0.0289

_lib = None


def header_functions(path: str = HEADER) -> list[str]:
    """Names of the functions include/wfs.h declares."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(wfs_\w+)\s*\(", text, re.M)))


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m water_flow_simulation_amd.build` "
                               "(there is no other compute path)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        if L.wfs_abi_version() != 1:
            raise RuntimeError("libwfs.so ABI version mismatch; rebuild it")
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().wfs_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = ""):
    if rc == WFS_OK:
        return
    msg = last_error() or what
    if rc == WFS_EINVAL:
        raise ValueError(f"libwfs: {msg}")
    if rc == WFS_ENONFINITE:
        raise NonFiniteState(rc, msg)
    raise WfsError(rc, msg)


def device_count() -> int:
    n = ctypes.c_int32(0)
    check(lib().wfs_device_count(ctypes.byref(n)))
    return int(n.value)


def bandwidth_probe(nbytes=2 << 30, iters=20, device=0) -> float:
    """Device-to-device copy bandwidth in GB/s (see wfs_bandwidth_probe)."""
    out = ctypes.c_double(0.0)
    check(lib().wfs_bandwidth_probe(int(device), int(nbytes), int(iters), ctypes.byref(out)))
    return float(out.value)


def stream_probe(px, ny, nz, elem_bytes=4, stagger_bytes=0, iters=20, device=0) -> float:
    """Streaming ceiling of the step's 19-population pull pattern over a
    px x ny x nz box in GB/s (see wfs_stream_probe)."""
    out = ctypes.c_double(0.0)
    check(lib().wfs_stream_probe(int(device), int(px), int(ny), int(nz), int(elem_bytes), int(stagger_bytes),
                                 int(iters), ctypes.byref(out)))
    return float(out.value)


def ptr(a):
    """Data pointer of a numpy array (None passes NULL)."""
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)
