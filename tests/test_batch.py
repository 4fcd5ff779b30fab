"""The batched form (pqmc_sweep_batch): chains of uneven lattice and Lt
advanced together, each against the oracle run alone on the same inputs and
draws: identical accepts, G and measurements to 1e-10.  A batch whose family
passes cannot be resident together is refused before anything is launched."""
import numpy as np
import pytest

import oracle as O
from pqmc_allinone_amd._lib import PQMCError
from pqmc_allinone_amd.host import Chain, Ran2, sweep_batch
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.gpu
def test_batch_of_uneven_chains_matches_oracle():
    cases = [(4, 100, 10, 1234), (6, 60, 10, 99), (4, 100, 10, 55)]
    chains, orcs, rngs, models = [], [], [], []
    for L, Lt, ng, seed in cases:
        m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=L * L))
        x, sl = synthetic_inputs(m, seed)
        ch = Chain(m)
        ch.set_fields(x)
        ch.set_trial(sl)
        ch.init_g(1)
        o = O.Oracle(L, L, Lt, ng, L * L)
        o.set_fields(x)
        o.set_trial(sl)
        o.init(1)
        chains.append(ch)
        orcs.append(o)
        rngs.append(Ran2(seed))
        models.append(m)
    forward = True
    for sweep in range(3):
        forward = not forward
        draws = [r.draws(2 * m.Nb * m.p.Lt) for r, m in zip(rngs, models)]
        print(f"batch sweep {sweep} forward={forward}", flush=True)
        outs = sweep_batch(chains, [forward] * 3, draws, measure=True)
        for i, (ch, o) in enumerate(zip(chains, orcs)):
            acc_o, mo = o.sweep(forward, draws[i], measure=True)
            assert np.array_equal(ch.accepts(), acc_o), (sweep, i)
            assert np.array_equal(ch.get_fields(), o.get_fields()), (sweep, i)
            assert rel(ch.g_h, o.g) < 1e-10, (sweep, i)
            assert rel(outs[i], mo) < 1e-10, (sweep, i)
    for ch in chains:
        ch.close()


@pytest.mark.gpu
def test_batch_that_does_not_fit_is_refused():
    """Three 14x14 chains need 3 x 99 resident family-pass workgroups of one
    compute unit each, more than the device has; the host check refuses the
    batch before any launch."""
    m = build_model(ModelParams(L1=14, L2=14, Lt=20, ngroup=10, ne=196))
    x, sl = synthetic_inputs(m, 3)
    chains = []
    for _ in range(3):
        ch = Chain(m)
        ch.set_fields(x)
        ch.set_trial(sl)
        ch.init_g(1)
        chains.append(ch)
    u = np.zeros(2 * m.Nb * m.p.Lt, np.float32)
    with pytest.raises(PQMCError, match="compute units"):
        sweep_batch(chains, [False] * 3, [u] * 3)
    with pytest.raises(PQMCError, match="twice"):
        sweep_batch([chains[0], chains[0]], [False] * 2, [u] * 2)
    for ch in chains:
        ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bph", [2, 4])
def test_helper_bonds_match_oracle(bph):
    """A family pass whose column-owning workgroups hold 2 or 4 bonds each:
    the same accepts, fields and G as the oracle over alternating sweeps,
    with decision blocks of 8 and of 3 bonds (blocks that straddle helpers)."""
    for L, Lt, ng, nb in [(4, 100, 10, 0), (6, 60, 10, 3)]:
        m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=L * L))
        x, sl = synthetic_inputs(m, 31)
        ch = Chain(m, block_bonds=nb, helper_bonds=bph)
        ch.set_fields(x)
        ch.set_trial(sl)
        ch.init_g(1)
        o = O.Oracle(L, L, Lt, ng, L * L)
        o.set_fields(x)
        o.set_trial(sl)
        o.init(1)
        rng = Ran2(31)
        fw = True
        for sweep in range(3):
            fw = not fw
            u = rng.draws(2 * m.Nb * Lt)
            print(f"bph={bph} L={L} nb={nb} sweep {sweep}", flush=True)
            ch.sweep(fw, u)
            acc_o, _ = o.sweep(fw, u)
            assert np.array_equal(ch.accepts(), acc_o), (L, sweep)
            assert rel(ch.g_h, o.g) < 1e-10, (L, sweep)
        assert np.array_equal(ch.get_fields(), o.get_fields())
        assert ch.counters()["sync_timeouts"] == 0
        ch.close()


@pytest.mark.gpu
def test_many_uneven_chains_match_lone_chains():
    """pqmc_sweep_many over more chains than fit together (24 projector
    chains of L in {4, 6, 8, 10}, Lt in {40, 60, 80}, one or four bonds per
    helper, and two finite-temperature chains; 8 host threads): each chain
    ends two sweeps bit for bit where its twin, swept alone, ends."""
    from pqmc_allinone_amd.host import sweep_many
    specs = [(L, Lt, bph, seed, False) for i, (L, Lt) in
             enumerate([(L, Lt) for L in (4, 6, 8, 10) for Lt in (40, 60, 80)])
             for bph, seed in [(1, 100 + i), (4, 200 + i)]]
    specs += [(4, 40, 1, 300, True), (6, 60, 4, 301, True)]

    def make(L, Lt, bph, seed, ft):
        m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=10, ne=L * L))
        x, sl = synthetic_inputs(m, seed)
        ch = Chain(m, helper_bonds=bph, finite_temperature=ft)
        ch.set_fields(x)
        if not ft:
            ch.set_trial(sl)
        ch.init_g(1)
        return m, ch

    pairs = [(make(*sp), make(*sp)) for sp in specs]
    rngs = [Ran2(sp[3]) for sp in specs]
    fw = True
    for sweep in range(2):
        fw = not fw
        draws = [r.draws(2 * m.Nb * m.p.Lt) for r, ((m, _), _) in zip(rngs, pairs)]
        print(f"sweep_many: {len(pairs)} chains, sweep {sweep}", flush=True)
        outs = sweep_many([a for (_, a), _ in pairs], [fw] * len(pairs), draws, measure=True, max_concurrent=8)
        for i, ((m, a), (_, b)) in enumerate(pairs):
            mb = b.sweep(fw, draws[i], measure=True)
            assert np.array_equal(a.accepts(), b.accepts()), (i, sweep)
            assert np.array_equal(a.g_h, b.g_h), (i, sweep)
            assert np.array_equal(outs[i], mb), (i, sweep)
    for (_, a), (_, b) in pairs:
        assert a.counters()["sync_timeouts"] == 0
        a.close()
        b.close()
