"""ran2 restatement (oracle and library host code) vs the reference's RNG.

tests/golden/ran2_stream.npz holds the streams of the reference's own
src/mod_ranstate.f90 compiled by oracle/Makefile.ref (see
tests/golden/make_ran2_golden.py): per seed, pairs (rands(), ran_sysm()).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ran2_stream.npz")


def _oracle_pairs(seed, n):
    u = O.ran2(seed, 2 * n)
    rands = u[0::2]
    sysm = (u[1::2] * np.float32(2) - np.float32(1)).astype(np.float32)
    return np.stack([rands.view(np.int32), sysm.view(np.int32)], axis=1)


@pytest.mark.parametrize("seed", [1, 7, 11, 42, 122, 1000, 1234, 2147483647])
def test_oracle_ran2_matches_reference_stream(seed):
    gold = np.load(GOLD)[f"seed_{seed}"]
    got = _oracle_pairs(seed, gold.shape[0])
    assert np.array_equal(got, gold)


def test_library_ran2_matches_oracle():
    """The host-side ran2 exported by libpqmc_hip.so (no GPU needed)."""
    from pqmc_allinone_amd import _lib
    L = _lib.lib()
    gold = np.load(GOLD)["seed_1234"]
    s = _lib.Ran2()
    L.pqmc_ran2_seed(C.byref(s), 1234)
    out = np.empty(2 * gold.shape[0], np.float32)
    # fill in two calls to check the carried state
    L.pqmc_ran2_fill(C.byref(s), 1000, _lib.ptr(out))
    L.pqmc_ran2_fill(C.byref(s), out.size - 1000, out[1000:].ctypes.data)
    assert np.array_equal(out[0::2].view(np.int32), gold[:, 0])
    sysm = (out[1::2] * np.float32(2) - np.float32(1)).astype(np.float32)
    assert np.array_equal(sysm.view(np.int32), gold[:, 1])


def test_ran2_range_and_negative_seed():
    u = O.ran2(-5, 100000)
    assert u.min() > 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01


@pytest.mark.skipif(not os.path.exists(os.path.join(O.ROOT, "oracle", "_ref", "ref_rng")),
                    reason="reference RNG build not present (only where the reference sources are)")
def test_oracle_ran2_matches_live_reference_build():
    binp = os.path.join(O.ROOT, "oracle", "_ref", "ref_rng")
    for seed in (99, -5, 0):
        r = subprocess.run([binp, str(seed), "2000"], capture_output=True, text=True, check=True)
        ref = np.array([[int(v) for v in ln.split()] for ln in r.stdout.split("\n") if ln.strip()], np.int64)
        assert np.array_equal(_oracle_pairs(seed, 2000), ref.astype(np.uint32).view(np.int32)), seed
