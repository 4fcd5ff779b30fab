"""The C-ABI library loads and exports every function include/pqmc_hip.h declares.

No compute call is made here (these tests run without a GPU): only the
header/export/binding agreement, the host-side ran2 generator and the error
path of pqmc_create when no gfx950 device is visible.
"""
import ctypes as C
import os
import re

import pytest

from pqmc_allinone_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pqmc_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pqmc_\w+)\s*\(", src, flags=re.M))


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"pqmc_create", "pqmc_sweep", "pqmc_update_family", "pqmc_evolve_g", "pqmc_g_scratch",
            "pqmc_measure", "pqmc_init_g"} <= names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_match_header():
    assert set(_lib.SIGNATURES) == declared_functions()
    _lib.lib()  # sets restype/argtypes for every symbol


def test_enum_values_match_header():
    src = open(HEADER).read()
    vals = dict((k, int(v)) for k, v in re.findall(r"(PQMC_E?\w+)\s*=\s*(\d+)", src))
    for k, v in vals.items():
        assert getattr(_lib, k) == v, k


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_create_without_device_fails_cleanly():
    cfg = _lib.Config(4, 4, 20, 10, 16, 0)
    h = C.c_void_p()
    rc = _lib.lib().pqmc_create(C.byref(cfg), C.byref(h))
    assert rc == _lib.PQMC_ENODEV
    assert h.value is None
    assert _lib.lib().pqmc_last_error()


def test_create_rejects_bad_sizes():
    cfg = _lib.Config(3, 3, 20, 10, 16, 0)  # odd number of cells
    h = C.c_void_p()
    assert _lib.lib().pqmc_create(C.byref(cfg), C.byref(h)) == _lib.PQMC_EINVAL
    assert _lib.lib().pqmc_create(None, C.byref(h)) == _lib.PQMC_EINVAL
