"""get_expKV_epsoc (src/mod_phonon_field.f90:572-604), the EPSOC model's 2x2
bond exponential, on the device (pqmc_expkv_epsoc) against the oracle for a
batch of (bond, slice) pairs over four families, with and without boundary
crossings, E and E^-1.  The oracle itself is pinned against scipy's expm in
tests/test_oracle_physics.py."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("inv", [False, True])
def test_expkv_epsoc_matches_oracle(inv):
    from pqmc_allinone_amd.host import expkv_epsoc
    rng = np.random.default_rng(17)
    dtau = 0.1
    # EPSOC.nml: K_coe, V_coe per family (here with -hop and lambda folded
    # in, as the reference's pf_type holds them), Kmatrix(1,2) = 1,
    # Vmatrix(1,2) = -i; one family with a common complex phase
    kcoe = np.array([-1.0, -1.0, -0.8, -0.9 * np.exp(0.2j)])
    vcoe = np.array([0.64, 0.64, 1.1, 0.7 * np.exp(0.2j)])
    k12 = np.ones(4, complex)
    v12 = np.full(4, -1j)
    n = 4096
    fam = rng.integers(0, 4, n)
    x = rng.standard_normal(n) * 1.5
    x[:8] = 0.0
    cross = (rng.random(n) < 0.25).astype(np.int32)
    bcph = np.exp(1j * np.pi * rng.random(n))
    got = expkv_epsoc(dtau, kcoe, vcoe, k12, v12, fam, x, bcph=bcph, cross=cross, inv=inv)
    print(f"{n} pairs, {int(cross.sum())} crossing, inv={inv}", flush=True)
    worst = 0.0
    for i in range(n):
        f = fam[i]
        ref = O.expkv_epsoc(dtau, kcoe[f], vcoe[f], k12[f], v12[f], x[i], bcph[i] if cross[i] else None, inv=inv)
        worst = max(worst, float(np.abs(got[i] - ref).max()))
    print(f"  max |E_device - E_oracle| = {worst:.2e}", flush=True)
    assert worst < 1e-14
    # a pair without crossings takes no phase
    plain = expkv_epsoc(dtau, kcoe, vcoe, k12, v12, fam[:16], x[:16], inv=inv)
    for i in range(16):
        f = fam[i]
        assert np.abs(plain[i] - O.expkv_epsoc(dtau, kcoe[f], vcoe[f], k12[f], v12[f], x[i], None, inv=inv)).max() < 1e-14


def test_expkv_epsoc_refuses_bad_family():
    from pqmc_allinone_amd import _lib
    from pqmc_allinone_amd.host import expkv_epsoc
    with pytest.raises(_lib.PQMCError) as e:
        expkv_epsoc(0.1, [1.0], [1.0], [1.0], [-1j], [0, 1], [0.0, 0.0])
    assert e.value.code == _lib.PQMC_EINVAL
