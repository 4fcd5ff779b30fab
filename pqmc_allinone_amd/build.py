"""Build steps: the HIP library for gfx950, the oracle, and (when the
reference sources are present) the reference RNG check binary."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["pqmc_abi.hip", "k_pass.hip", "k_wrap.hip", "k_dense.hip", "k_meas.hip", "k_mixed.hip", "k_qrflow.hip"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wno-unused-value",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip() -> str:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = os.path.join(HERE, "libpqmc_hip.so")
    tmp = out + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-o", tmp, *SOURCES], cwd=CSRC)
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_reference_rng():
    """oracle/_ref/ref_rng from the reference's own RNG sources, if present."""
    if not os.path.isdir("/root/reference/src") or shutil.which("amdflang") is None:
        return None
    _run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile.ref")])
    return os.path.join(ROOT, "oracle", "_ref", "ref_rng")


def build_fortran_host():
    """Fortran host driver over the C ABI (fortran/), if amdflang is present."""
    mk = os.path.join(HERE, "fortran", "Makefile")
    if not os.path.exists(mk) or shutil.which("amdflang") is None:
        return None
    _run(["make", "-s", "-C", os.path.join(HERE, "fortran")])
    return os.path.join(HERE, "fortran", "pqmc_driver")


def build():
    build_hip()
    build_oracle()
    build_reference_rng()
    build_fortran_host()
