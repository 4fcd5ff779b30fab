"""ctypes binding of libpqmc_hip.so (declarations in include/pqmc_hip.h).

The library is built in-tree by ``pqmc_allinone_amd.build.build()``.  There
is no fallback: if the shared library is missing or does not load, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpqmc_hip.so")

PQMC_OK, PQMC_EINVAL, PQMC_EHIP, PQMC_ENUMERIC, PQMC_ESYNC, PQMC_ENODEV = range(6)


class PQMCError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pqmc error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("L1", C.c_int), ("L2", C.c_int), ("Lt", C.c_int), ("ngroup", C.c_int), ("ne", C.c_int),
                ("block_bonds", C.c_int), ("helper_bonds", C.c_int), ("finite_temperature", C.c_int)]


class ModelTables(C.Structure):
    _fields_ = [("sites", C.c_void_p), ("bfi", C.c_void_p), ("cross", C.c_void_p), ("bcph", C.c_void_p),
                ("kcoe", C.c_void_p), ("vcoe", C.c_void_p), ("k12", C.c_void_p), ("v12", C.c_void_p),
                ("bc", C.c_void_p), ("dtau", C.c_double), ("M", C.c_double), ("D", C.c_double),
                ("biased", C.c_double), ("jump_distance", C.c_double), ("hop", C.c_double), ("lam", C.c_double),
                ("ncopy", C.c_int32), ("tr_weight", C.c_int32)]


class Ran2(C.Structure):
    _fields_ = [(n, C.c_int32) for n in "ijkmn"]


_P = C.c_void_p
_I = C.c_int
# name -> (restype, argtypes); every symbol of include/pqmc_hip.h
SIGNATURES = {
    "pqmc_create": (_I, [C.POINTER(Config), C.POINTER(_P)]),
    "pqmc_destroy": (_I, [_P]),
    "pqmc_last_error": (C.c_char_p, []),
    "pqmc_get_dims": (_I, [_P, _P]),
    "pqmc_set_model": (_I, [_P, C.POINTER(ModelTables)]),
    "pqmc_set_fields": (_I, [_P, _P]),
    "pqmc_get_fields": (_I, [_P, _P]),
    "pqmc_set_trial": (_I, [_P, _P]),
    "pqmc_init_g": (_I, [_P, _I]),
    "pqmc_set_precision": (_I, [_P, _I]),
    "pqmc_set_qr_mode": (_I, [_P, _I]),
    "pqmc_update_family": (_I, [_P, _I, _I, _P, _P]),
    "pqmc_evolve_g": (_I, [_P, _I, _I, _I]),
    "pqmc_evolve_matrix": (_I, [_P, _I, _I, _I, _I, _P, _I]),
    "pqmc_g_scratch": (_I, [_P, _I, _I, _P]),
    "pqmc_measure": (_I, [_P, _I, _P]),
    "pqmc_sweep": (_I, [_P, _I, _P, _I, _I, _P]),
    "pqmc_sweep_batch": (_I, [_P, _I, _P, _P, _I, _P]),
    "pqmc_sweep_many": (_I, [_P, _I, _P, _P, _I, _P, _I]),
    "pqmc_get_g": (_I, [_P, _P]),
    "pqmc_set_g": (_I, [_P, _P]),
    "pqmc_get_qstring": (_I, [_P, _I, _P]),
    "pqmc_get_dstring": (_I, [_P, _P]),
    "pqmc_get_lncw": (_I, [_P, _P]),
    "pqmc_global_weight": (_I, [_P, _P, _P]),
    "pqmc_gradient": (_I, [_P, _P]),
    "pqmc_gradient_noalloc": (_I, [_P, _P]),
    "pqmc_get_accepts": (_I, [_P, _P, C.c_int64]),
    "pqmc_get_counters": (_I, [_P, _P, _P]),
    "pqmc_get_drift": (_I, [_P, _P]),
    "pqmc_get_drift_history": (_I, [_P, _P, C.c_int32, _P]),
    "pqmc_set_counters": (_I, [_P, _P, C.c_double]),
    "pqmc_synchronize": (_I, [_P]),
    "pqmc_ran2_seed": (None, [C.POINTER(Ran2), C.c_int32]),
    "pqmc_ran2_fill": (None, [C.POINTER(Ran2), C.c_int64, _P]),
    "pqmc_mfma_selftest": (_I, [_P]),
    "pqmc_dense_check": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P, _I]),
    "pqmc_expkv_epsoc": (_I, [C.c_double, C.c_int32, _P, _P, _P, _P, C.c_int64, _P, _P, _P, _P, C.c_int32, _P]),
    "pqmc_timing": (_I, [_P, _I, _I, _P, _P]),
    "pqmc_update_profile": (_I, [_P, _I, _P, _P]),
}

_lib = None


def lib():
    """Load libpqmc_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run pqmc_allinone_amd.build.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != PQMC_OK:
        msg = lib().pqmc_last_error()
        raise PQMCError(rc, msg.decode() if msg else "")


def ptr(a) -> int:
    """Address of a contiguous numpy array."""
    return a.ctypes.data
