"""Finite-temperature Green's function (proj = .false.): init_g_FT,
get_g_scratch_FT and cal_g_and_det_FT (src/mod_evolution.f90:411-785).

The oracle is pinned against dense products on small lattices:
g_h = 1 - (1 + B(tau,0) B(beta,tau))^-1 and Re ln_cw = ln|det(1 + B(beta,0))|,
after init, at stabilizations in both directions and after full sweeps
(including accepted updates)."""
import numpy as np
import pytest

import bruteforce as BF
import oracle as O
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs


def ft_pair(L, Lt, ng, seed=3, lam=0.64):
    p = ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=L * L, lam=lam)
    m = build_model(p)
    x, sl = synthetic_inputs(m, seed)
    orc = O.Oracle(L, L, Lt, ng, L * L, lam=lam, ft=True)
    orc.set_fields(x)
    orc.set_trial(sl)
    return m, x, orc


def dense_b(m, x, t1, t0=0):
    """B(t1, t0) = S_t1 ... S_{t0+1}."""
    B = np.eye(m.Ns, dtype=complex)
    for t in range(t0 + 1, t1 + 1):
        B = BF.slice_matrix(m, x, t) @ B
    return B


def g_ft(m, x, tau):
    """1 - (1 + B(tau,0) B(beta,tau))^-1 and a bound on the dense reference's
    own error: the inverse of the unstabilized 1 + B is good to about
    eps * cond(1 + B)."""
    Lt = m.p.Lt
    M = np.eye(m.Ns) + dense_b(m, x, tau) @ dense_b(m, x, Lt, tau)
    tol = max(1e-10, 1e-13 * np.linalg.cond(M))
    return np.eye(m.Ns) - np.linalg.inv(M), tol


@pytest.mark.parametrize("L,Lt,ng", [(2, 12, 3), (2, 40, 10), (4, 20, 5)])
def test_ft_init_matches_dense(L, Lt, ng):
    m, x, orc = ft_pair(L, Lt, ng)
    orc.init(1)
    ref, tol = g_ft(m, x, Lt)
    assert tol < 1e-5  # the dense reference still holds five digits
    assert np.abs(orc.g - ref).max() < tol
    B = dense_b(m, x, Lt)
    lw = np.linalg.slogdet(np.eye(m.Ns) + B)[1]
    assert abs(orc.ln_cw.real - lw) < 1e-9 * max(1.0, abs(lw))


@pytest.mark.parametrize("L,Lt,ng", [(2, 30, 5), (4, 20, 5)])
def test_ft_sweeps_match_dense(L, Lt, ng):
    """Two sweeps with updates: G at the end from the strings equals the dense
    formula for the final fields, stabilizations on both directions see small
    drift, and some bonds were accepted."""
    m, x, orc = ft_pair(L, Lt, ng)
    orc.init(1)
    rng = O.Ran2Stream(17)
    fw = True
    tot = 0
    for _ in range(2):
        fw = not fw
        acc, _ = orc.sweep(fw, rng.draws(2 * m.Nb * Lt))
        tot += int(acc.sum())
        xf = orc.get_fields()
        ref, tol = g_ft(m, xf, 0 if not fw else Lt)
        assert np.abs(orc.g - ref).max() < max(tol, 1e-9)
    assert tot > 0
    pa, na, efm = orc.counters()
    assert efm <= 1e-6


def test_ft_midsweep_stabilization_matches_dense():
    """Partial backward sweep: G right after the stabilization at time t is
    1 - (1 + B(t-1, 0) B(beta, t-1))^-1 for the current fields."""
    L, Lt, ng = 2, 30, 5
    m, x, orc = ft_pair(L, Lt, ng)
    orc.init(1)
    rng = O.Ran2Stream(5)
    orc.sweep(False, rng.draws(2 * m.Nb * Lt), ntc=2 * ng)
    xf = orc.get_fields()
    t = Lt - 2 * ng + 1
    ref, tol = g_ft(m, xf, t - 1)
    assert np.abs(orc.g - ref).max() < max(tol, 1e-9)
