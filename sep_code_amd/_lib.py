"""ctypes binding of the C ABI in include/born_rtm.h.

The library is ``sep_code_amd/lib/libborn_rtm.so`` (built by
``sep_code_amd.build``).  There is no fallback: when the library is missing
or cannot be loaded, importing the operator classes raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

LIBPATH = os.path.join(_build.LIBDIR, _build.LIBNAME)

BRTM_OK = 0
STORE_NONE, STORE_F32, STORE_BF16, STORE_F16, STORE_CKPT = 0, 1, 2, 3, 4
STORE_NAMES = {"none": STORE_NONE, "f32": STORE_F32, "bf16": STORE_BF16, "f16": STORE_F16,
               "ckpt": STORE_CKPT}
FORM_VL, FORM_LV = 0, 1


class BrtmError(RuntimeError):
    """A C ABI call returned a non-zero status."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


_lib = None

# (name, argtypes) of every exported function; the restype is int unless
# listed in _RESTYPES.
_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
SIGNATURES = {
    # born_rtm.h
    "brtm_last_error": [],
    "brtm_version": [],
    "brtm_create": [_I, C.POINTER(_P)],
    "brtm_destroy": [_P],
    "brtm_create_space": [_P, _F, _F, _F, _F, _F, _I, _I, _I],
    "brtm_set_velocity": [_P, _P],
    "brtm_set_sinc_table": [_P, _I, _I, _I, _P],
    "brtm_set_source": [_P, _I, _I, _P, _P],
    "brtm_set_receivers": [_P, _I, _I, _I, _P, _P],
    "brtm_get_receivers": [_P, _P],
    "brtm_set_model": [_P, _P],
    "brtm_get_image": [_P, _P],
    "brtm_zero_image": [_P],
    "brtm_select_shot": [_P, _I],
    "brtm_release_store": [_P, _I],
    "brtm_source_prop": [_P, _I, _I, _I, _I, _I, _P, _P],
    "brtm_rtm_forward": [_P, _I, _I, _I, _I],
    "brtm_rtm_adjoint": [_P, _I, _I, _I],
    "brtm_set_imaging_scale": [_P, _F],
    "brtm_get_snapshot": [_P, _I, _P],
    "brtm_propagate": [_P, _I, _I, _P, _P],
    "brtm_time_step": [_P, _I, _I, _I, _I, C.POINTER(_D)],
    "brtm_set_tuning": [_P, _I, _I],
    "brtm_set_x_per_lane": [_P, _I],
    "brtm_set_halo_ring": [_P, _I],
    "brtm_set_receiver_fusion": [_P, _I],
    "brtm_set_prefetch": [_P, _I],
    "brtm_set_stream_loads": [_P, _I],
    "brtm_set_born_path": [_P, _I],
    "brtm_set_graphs": [_P, _I],
    "brtm_set_two_step": [_P, _I],
    "brtm_kernel_timing": [_P, _I, _I, _I, _I],
    "brtm_kernel_timing_result": [_P, C.POINTER(_D), C.POINTER(_I)],
    "brtm_copy_bench": [_P, _D, _I, C.POINTER(_D)],
    "brtm_stream_bench": [_P, _I, _D, _I, _I, C.POINTER(_D)],
    "brtm_dot_noalloc": [_P, _P, _P, C.c_long, C.POINTER(_D)],
    "brtm_sync": [_P],
    "brtm_device_bytes": [_P, C.POINTER(_D)],
}
_RESTYPES = {"brtm_last_error": C.c_char_p, "brtm_version": C.c_char_p}


def load() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            f"{LIBPATH} is missing: build it with `python -m sep_code_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(LIBPATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def register(signatures: dict, restypes: dict | None = None) -> None:
    """Declare further exported functions (used by the operator and conv modules)."""
    lib = load()
    for name, args in signatures.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = (restypes or {}).get(name, C.c_int)
    SIGNATURES.update(signatures)


def check(fn: str, rc: int) -> None:
    if rc != BRTM_OK:
        msg = load().brtm_last_error().decode(errors="replace")
        raise BrtmError(fn, rc, msg)


def call(fn: str, *args) -> None:
    check(fn, getattr(load(), fn)(*args))


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the C ABI must be C-contiguous"
    return a.ctypes.data


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)
