"""LAPACK's ZGEQRF and ZGELQF, from the OpenBLAS build that scipy ships,
called through ctypes.  They pin the sign conventions of the factorizations
the reference calls through qdr / ldq (src/mod_matrixlib.f90:323-360,
:470-503): the diagonal of R is (beta, +0), the diagonal of L is its
conjugate (beta, -0), and the sign of that zero decides the branch of
log(d) for a negative d.  `available()` is False when no such library is
found; the tests that need it then skip."""
import ctypes as C
import glob
import importlib
import os

import numpy as np

_fns = None


def _load():
    global _fns
    if _fns is not None:
        return _fns
    _fns = {}
    try:
        mod = importlib.import_module("scipy")
    except ImportError:
        return _fns
    libdir = os.path.join(os.path.dirname(mod.__file__), os.pardir, "scipy.libs")
    for path in sorted(glob.glob(os.path.join(libdir, "libscipy_openblas*.so"))):
        lib = C.CDLL(path)
        for key in ("zgeqrf", "zgelqf"):
            for name in (f"scipy_{key}_", f"{key}_"):  # LP64 (32-bit integer) interface only
                if key not in _fns and hasattr(lib, name):
                    _fns[key] = getattr(lib, name)
    return _fns


def available() -> bool:
    f = _load()
    return "zgeqrf" in f and "zgelqf" in f


def _factor(key, a):
    a = np.asfortranarray(a, np.complex128).copy(order="F")
    m, n = a.shape
    tau = np.zeros(min(m, n), np.complex128)
    lwork = 64 * max(m, n)
    work = np.zeros(lwork, np.complex128)
    M, N, LDA, LW, INFO = (C.c_int(v) for v in (m, n, m, lwork, 0))
    _load()[key](C.byref(M), C.byref(N), a.ctypes.data_as(C.c_void_p), C.byref(LDA),
                 tau.ctypes.data_as(C.c_void_p), work.ctypes.data_as(C.c_void_p), C.byref(LW), C.byref(INFO))
    if INFO.value != 0:
        raise RuntimeError(f"{key} info={INFO.value}")
    return a, tau


def zgeqrf(a):
    """Factored a (R on and above the diagonal) and tau."""
    return _factor("zgeqrf", a)


def zgelqf(a):
    """Factored a (L on and below the diagonal) and tau."""
    return _factor("zgelqf", a)
