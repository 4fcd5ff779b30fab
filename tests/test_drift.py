"""The drift check of get_g_scratch_T0 (src/mod_evolution.f90:179-193) on
the device: every stabilization of a sweep counts, not only the last one,
and a non-finite Green's function or log-weight fails the call with
PQMC_ENUMERIC (the reference would carry on and print; this library stops,
as it does on other non-finite results)."""
import numpy as np
import pytest

import oracle as O
from pqmc_allinone_amd import _lib
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

pytestmark = pytest.mark.gpu

L, LT, NG, NE, SEED = 4, 100, 10, 16, 1234


def _chain():
    from pqmc_allinone_amd.host import Chain
    m = build_model(ModelParams(L1=L, L2=L, Lt=LT, ngroup=NG, ne=NE))
    x, sl = synthetic_inputs(m, SEED)
    ch = Chain(m)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    return m, x, sl, ch


def test_nan_in_g_before_the_first_stabilization_is_reported():
    """G carries a NaN into the first ten slices of a backward sweep; the
    first stabilization replaces it with a finite G from scratch, and the
    nine after it are finite.  Only a check that keeps every stabilization
    sees it."""
    m, x, sl, ch = _chain()
    g = ch.g_h.copy()
    g[3, 5] = np.nan
    ch.g_h = g
    u = O.Ran2Stream(SEED).draws(2 * m.Nb * LT)
    with pytest.raises(_lib.PQMCError) as e:
        ch.sweep(False, u)
    assert e.value.code == _lib.PQMC_ENUMERIC, e.value
    dg, _ = ch.drift()
    assert dg == np.inf
    # the chain itself ended finite: the later stabilizations were finite
    assert np.isfinite(ch.g_h).all()


def test_nan_field_is_reported():
    m, x, sl, ch = _chain()
    x = x.copy()
    x[LT // 2, 3] = np.nan
    ch.set_fields(x)
    ch.init_g(1)
    u = O.Ran2Stream(SEED).draws(2 * m.Nb * LT)
    with pytest.raises(_lib.PQMCError) as e:
        ch.sweep(False, u)
    assert e.value.code == _lib.PQMC_ENUMERIC, e.value


def test_finite_chain_reports_small_drift():
    m, x, sl, ch = _chain()
    u = O.Ran2Stream(SEED).draws(2 * m.Nb * LT)
    ch.sweep(False, u)
    dg, dl = ch.drift()
    assert 0.0 < dg < 1e-9 and 0.0 <= dl < 1e-9
    assert ch.counters()["err_fast_max"] == 1e-6
