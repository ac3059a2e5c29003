"""ctypes binding of ``libccx.so`` (declared in ``include/ccx.h``).

The library is loaded from this package directory. torch is imported first so that the HIP
runtime it bundles is the one the library binds to. A missing library raises ``PyCCError``: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

from .exceptions import DeviceError, PyCCError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libccx.so")

MODE_DP = 0
MODE_MIXED = 1
MODE_MIXED_BF16X3 = 2
MODE_MIXED_F16X3 = 3
HBAR_STORED = 1
VJP_REUSE_HBAR = 2

_lock = threading.Lock()
_lib = None

c_dp = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "ccx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "ccx_destroy": (ctypes.c_int, [_vp]),
    "ccx_set_integrals_dense": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "ccx_set_integrals_df": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "ccx_residual": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ccx_energy": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ccx_update": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ccx_diis_reset": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "ccx_diis_add": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "ccx_diis_extrapolate": (ctypes.c_int, [_vp, _vp, _vp, c_dp, _vp]),
    "ccx_hbar_build": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "ccx_hbar_block": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "ccx_lambda_residual": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "ccx_pseudoenergy": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "ccx_vjp": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "ccx_jvp": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "ccx_densities": (ctypes.c_int, [_vp] * 13),
    "ccx_density_energy": (ctypes.c_int, [_vp] * 13),
    "ccx_triples": (ctypes.c_int, [_vp] * 6),
    "ccx_triples_range": (ctypes.c_int, [_vp] * 4 + [ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "ccx_t3_denominator": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "ccx_timing_start": (ctypes.c_int, [_vp]),
    "ccx_timing_read": (ctypes.c_int, [_vp, c_dp, ctypes.c_int]),
    "ccx_device_bytes": (ctypes.c_int64, [_vp]),
    "ccx_contract": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, _vp, c_i64p, c_i64p, _vp, c_i64p, c_i64p,
                                    ctypes.c_double, _vp, c_i64p, c_i64p, _vp]),
    "ccx_contract_noalloc": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, _vp, c_i64p, c_i64p, _vp, c_i64p,
                                            c_i64p, ctypes.c_double, _vp, c_i64p, c_i64p, _vp]),
    "ccx_contract_f32": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_float, _vp, c_i64p, c_i64p, _vp, c_i64p,
                                        c_i64p, ctypes.c_float, _vp, c_i64p, c_i64p, _vp]),
    "ccx_gemm_f16x3": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_float, _vp, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_float, _vp,
                                      ctypes.c_int64, _vp]),
    "ccx_last_error": (ctypes.c_char_p, []),
}


def load(path: str = LIB_PATH):
    """Load (once) and return the bound library."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise PyCCError("HIP library %s is missing: build it with `python -c \"import __graft_entry__ as g; "
                            "g.build()\"` (or `make -C pycc_amd/csrc`)" % path)
        import torch  # noqa: F401  (binds the HIP runtime torch ships before dlopen of the library)
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(status: int, call: str):
    if status != 0:
        msg = load().ccx_last_error()
        raise DeviceError(status, msg.decode() if msg else "", call)


def call(name: str, *args):
    """Invoke an int-returning entry point and raise DeviceError on failure."""
    fn = getattr(load(), name)
    check(fn(*args), name)
