"""A family pass with an ill-conditioned accept, on the device, against the
same updates in extended precision.  In the backward sweep of a chain with
8x8 cells, Lt=800 and seed 1018, slice 449, family 4, bond 1 is accepted with
|det(I + G_SS Delta)| = 1.9e-4 (the bosonic action lets the small ratio
through); the next accepted bond then sees a G with entries near 2.6e5 and a
4x4 matrix M with max|M| = 5.7e3.  The oracle's float64 pass is 1.9e-7 from
the extended-precision result, which is also the spread that one-ulp
perturbations of the inputs give.  Inverting M through its adjugate (C^T /
det) left the device 4.4e-3 away (5e-2 with one bond per decision block);
the pivoted elimination that takes over when max|M|^4 / |det| exceeds 1e4
brings it back to the oracle's level."""
import numpy as np
import pytest

import bruteforce as BF
import oracle as O
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

L, LT, SEED, TC, FAM = 8, 800, 1018, 352, 4


@pytest.mark.gpu
@pytest.mark.parametrize("nb", [1, 8])
def test_ill_conditioned_pass_matches_extended_precision(nb):
    from pqmc_allinone_amd.host import Chain
    m = build_model(ModelParams(L1=L, L2=L, Lt=LT, ngroup=10, ne=L * L))
    x, sl = synthetic_inputs(m, SEED)
    orc = O.Oracle(L, L, LT, 10, L * L)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init(1)
    u = O.Ran2Stream(SEED).draws(2 * m.Nb * LT)
    print(f"oracle: backward sweep to slice {LT - TC + 1}", flush=True)
    off = 0
    for tc in range(1, TC + 1):
        t = LT - tc + 1
        for f in (4, 3, 2, 1):
            orc.wrap_g(t, f, False)
            if tc == TC and f == FAM:
                break
            orc.update_phonon_field(f, t, u[off:off + 2 * m.npla])
            off += 2 * m.npla
        if tc % 10 == 0:
            orc.get_g_scratch(t, False)
    t = LT - TC + 1
    uu = u[off:off + 2 * m.npla]
    g0, xb = orc.g, orc.get_fields()
    acc_o = orc.update_phonon_field(FAM, t, uu)
    xa = orc.get_fields()
    deltas = []
    for b in np.nonzero(acc_o)[0]:
        fld = m.bfi[FAM - 1, b]
        e_old_inv = orc.expkv(FAM, b, xb[t - 1, fld], 1)
        deltas.append((m.sites[FAM - 1, b], e_old_inv @ orc.expkv(FAM, b, xa[t - 1, fld], 0) - np.eye(4)))
    ref = BF.sequential_updates_extended(g0, deltas)
    e_orc = np.abs(orc.g - ref).max()
    ch = Chain(m, block_bonds=nb)
    ch.set_fields(xb)
    ch.set_trial(sl)
    ch.init_g(1)
    ch.g_h = g0
    acc_g = ch.update_phonon_field(FAM, t, uu)
    e_gpu = np.abs(ch.g_h - ref).max()
    print(f"  max|G| before {np.abs(g0).max():.1f}; pass error: oracle {e_orc:.2e}, device {e_gpu:.2e}", flush=True)
    assert np.array_equal(acc_g, acc_o)
    assert e_orc < 1e-5 and e_gpu < 1e-5
    ch.close()
