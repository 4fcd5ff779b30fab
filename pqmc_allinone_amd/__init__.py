"""MI355X-native projector-DQMC sweep (drop-in for the PQMC_allinone hot path).

The compute path is the HIP library ``libpqmc_hip.so`` (csrc/, C ABI in
include/pqmc_hip.h); this package holds the host-side mirror of the
reference's module interfaces and the EPSOCZ model set-up.
"""
from .model import ModelParams, Model, build_model, synthetic_inputs  # noqa: F401

__all__ = ["ModelParams", "Model", "build_model", "synthetic_inputs"]
