"""The family-pass shape autotuner: it times every accepted
(block_bonds, helper_bonds) pair on the caller's lattice and returns the
fastest; whichever shape it picks, a sweep keeps the accept sequence of the
default shape."""
import numpy as np
import pytest

from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs


@pytest.mark.gpu
def test_autotuned_shape_keeps_the_accept_sequence():
    from pqmc_allinone_amd.host import Chain, Ran2, autotune_pass
    m = build_model(ModelParams(L1=6, L2=6, Lt=40, ngroup=10, ne=36))
    x, sl = synthetic_inputs(m, 12)
    print("autotune 6x6 Lt=40", flush=True)
    best, times = autotune_pass(m, x, sl, block_bonds=(2, 8), helper_bonds=(1, 4), slices=10)
    print(f"  best {best} times {times}", flush=True)
    assert best in times and len(times) == 4 and all(v > 0 for v in times.values())
    u = Ran2(3).draws(2 * m.Nb * m.p.Lt)
    acc = []
    for bb, hb in ((0, 0), best):
        ch = Chain(m, block_bonds=bb, helper_bonds=hb)
        ch.set_fields(x)
        ch.set_trial(sl)
        ch.init_g(1)
        ch.sweep(False, u)
        acc.append(ch.accepts())
        ch.close()
    assert np.array_equal(acc[0], acc[1]) and acc[0].any()
