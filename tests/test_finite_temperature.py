"""Finite-temperature Green's function (proj = .false.): init_g_FT,
get_g_scratch_FT and cal_g_and_det_FT (src/mod_evolution.f90:411-785).

The oracle is pinned against dense products on small lattices:
g_h = 1 - (1 + B(tau,0) B(beta,tau))^-1, Re ln_cw = ln|det(1 + B(beta,0))| and
Im ln_cw = arg det(1 + B(beta,0)) (mod 2 pi), after init, at stabilizations in
both directions and after full sweeps (including accepted updates).  The
device path is compared with the oracle (marked gpu)."""
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
    check_lncw_dense(m, x, orc.ln_cw)


def check_lncw_dense(m, x, lncw):
    """Re ln_cw = ln|det(1 + B(beta,0))|; Im ln_cw = arg det(1 + B) up to
    2 pi (tr_weight off, ncopy 1).  The final mod 2 pi_SP reduction leaves
    n (2 pi_SP - 2 pi), n below (Ns + 3) / 2 (one +-pi per log term)."""
    sign, lw = np.linalg.slogdet(np.eye(m.Ns) + dense_b(m, x, m.p.Lt))
    assert abs(lncw.real - lw) < 1e-9 * max(1.0, abs(lw))
    dphi = np.angle(np.exp(1j * lncw.imag) / sign)
    assert abs(dphi) < 1e-9 + (m.Ns + 3) / 2 * 1.8e-7


def qrp_chain_g(m, x):
    """G = (1 + B(beta,0))^-1, Re ln det(1 + B) and the log of the largest
    scale, from an independent UDT chain: LAPACK's column-pivoted QR
    (ZGEQP3 through scipy, overflow-safe norms) per ngroup slices,
    Q D T <- QRP(B_blk Q D) with T accumulated, then the scale-split formula
    of cal_g_and_det_FT with the identity as the left string.  (A chain of
    plain SVDs, the reference's udv, loses the small scales of these graded
    products and is not used as a check.)"""
    import scipy.linalg as sla
    Ns, Lt, ng = m.Ns, m.p.Lt, m.p.ngroup
    Q, d, Tm = np.eye(Ns, dtype=complex), np.ones(Ns, dtype=complex), np.eye(Ns, dtype=complex)
    for b in range(Lt // ng):
        X = np.eye(Ns, dtype=complex)
        for t in range(b * ng + 1, (b + 1) * ng + 1):
            X = BF.slice_matrix(m, x, t) @ X
        q, r, piv = sla.qr((X @ Q) * d, pivoting=True)
        d = np.diag(r).copy()
        t = np.empty_like(r)
        t[:, piv] = r / d[:, None]  # T' P^T
        Q, Tm = q, t @ Tm
    ad = np.abs(d)
    db = np.where(ad > 1, d, 1.0)
    ds = np.where(ad > 1, 1.0, d)
    # 1 + Q D T = Q Db (Db^-1 Q^H T^-1 + Ds) T
    Mm = np.linalg.solve(Tm.T, Q.conj()).T / db[:, None] + np.diag(ds)
    G = np.linalg.solve(Tm, np.linalg.solve(Mm, Q.conj().T / db[:, None]))
    lw = np.log(np.abs(db)).sum() + np.linalg.slogdet(Mm)[1] + np.linalg.slogdet(Tm)[1]
    return np.eye(Ns) - G, lw, np.log(ad.max())


@pytest.mark.parametrize("L,Lt,ng,ng2", [(2, 1000, 10, 20), (4, 1000, 10, 25), (4, 500, 5, 1), (2, 100, 10, 1)])
def test_ft_large_beta_matches_pivoted_qr_chain(L, Lt, ng, ng2):
    """beta = 100, 50 and 10 (dense inverses are useless from about 10): at
    beta = 100 the largest scale of B is about exp(400), so its square
    overflows a double; the pre-pivoted, exactly rescaled QR strings give
    the G and Re ln det(1 + B) of a column-pivoted QR chain, and the same G
    for another block length ng2."""
    m, x, orc = ft_pair(L, Lt, ng, seed=7)
    orc.init(1)
    ref, lw, lmax = qrp_chain_g(m, x)
    if Lt == 1000:
        assert lmax > 354
    assert np.isfinite(orc.g).all()
    assert np.abs(orc.g - ref).max() < 1e-9
    assert abs(orc.ln_cw.real - lw) < 1e-10 * abs(lw)
    _, _, o2 = ft_pair(L, Lt, ng2, seed=7)
    o2.init(1)
    assert np.abs(o2.g - orc.g).max() < 1e-12


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
        if fw:  # the last stabilization ends the sweep: no accepts after it
            check_lncw_dense(m, xf, orc.ln_cw)
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


# ---- the device path against the oracle -------------------------------------

SP_TWO_PI = float(np.float32(2 * np.pi))


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def ft_chain(m, x):
    from pqmc_allinone_amd.host import Chain
    ch = Chain(m, finite_temperature=True)
    ch.set_fields(x)
    return ch


def same_lncw(a, b, ns):
    """Re to 1e-9 relative; Im as a phase.  det(1 + B) is real and positive
    here and B is real at zero boundary phases, so det M, det Q and the
    scales are real too: each log term is 0 or +-i pi, the sign of a
    rounded zero choosing the branch, and the unreduced sums of the two
    sides may differ by multiples of 2 pi.  The final mod 2 pi_SP turns
    those into multiples of 2 pi - 2 pi_SP (1.75e-7), at most one per term."""
    if abs(a.real - b.real) > 1e-9 * max(1.0, abs(b.real)):
        return False
    d = abs(a.imag - b.imag) % (2 * np.pi)
    return min(d, 2 * np.pi - d) < 1e-9 + (ns + 3) / 2 * 1.8e-7


@pytest.mark.gpu
@pytest.mark.parametrize("L,Lt,ng", [(2, 12, 3), (4, 20, 5), (4, 40, 10)])
def test_gpu_ft_init_matches_oracle(L, Lt, ng):
    m, x, orc = ft_pair(L, Lt, ng)
    orc.init(1)
    ch = ft_chain(m, x)
    print(f"ft init L={L} Lt={Lt} ng={ng}", flush=True)
    ch.init_g(1)
    e = rel(ch.g_h, orc.g)
    print(f"  rel G err {e:.3e}  lncw gpu {ch.ln_cw} oracle {orc.ln_cw}", flush=True)
    assert e < 1e-10
    assert same_lncw(ch.ln_cw, orc.ln_cw, m.Ns)
    ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("L,Lt,ng", [(2, 30, 5), (4, 20, 5), (6, 40, 10)])
def test_gpu_ft_sweeps_match_oracle(L, Lt, ng):
    """Four sweeps, alternating direction, with measurements: identical
    accepts and fields, G, ln_cw and measurements to 1e-9."""
    m, x, orc = ft_pair(L, Lt, ng)
    orc.init(1)
    ch = ft_chain(m, x)
    ch.init_g(1)
    rng = O.Ran2Stream(17)
    fw, tot = True, 0
    for sweep in range(4):
        fw = not fw
        u = rng.draws(2 * m.Nb * Lt)
        print(f"ft sweep L={L} Lt={Lt} ng={ng} sweep={sweep} forward={fw}", flush=True)
        mg = ch.sweep(fw, u, measure=True)
        acc_o, mo = orc.sweep(fw, u, measure=True)
        acc = ch.accepts()
        assert np.array_equal(acc, acc_o), sweep
        assert np.array_equal(ch.get_fields(), orc.get_fields()), sweep
        e = rel(ch.g_h, orc.g)
        print(f"  accepts {int(acc.sum())} rel G err {e:.3e} meas err {rel(mg, mo):.3e}", flush=True)
        assert e < 1e-9, sweep
        assert rel(mg, mo) < 1e-9, sweep
        assert same_lncw(ch.ln_cw, orc.ln_cw, m.Ns), (sweep, ch.ln_cw, orc.ln_cw)
        tot += int(acc.sum())
    assert tot > 0
    ch.close()


@pytest.mark.gpu
def test_gpu_ft_refusals():
    """Lt not a multiple of ngroup and the projector string readers are
    refused for a finite-temperature chain."""
    from pqmc_allinone_amd._lib import PQMCError
    from pqmc_allinone_amd.host import Chain
    with pytest.raises(PQMCError, match="Lt % ngroup"):
        Chain(build_model(ModelParams(L1=2, L2=2, Lt=12, ngroup=5, ne=4)), finite_temperature=True)
    m, x, _ = ft_pair(2, 12, 3)
    ch = ft_chain(m, x)
    ch.init_g(1)
    with pytest.raises(PQMCError, match="finite temperature"):
        ch.q_string(1)
    with pytest.raises(PQMCError, match="finite temperature"):
        ch.d_string()
    ch.close()


# ---- global moves at finite temperature (cal_new_weight_FT) -----------------


def test_ft_new_weight_matches_dense_and_leaves_the_chain():
    """cal_new_weight_FT of proposed fields is ln det(1 + B(beta,0)) of those
    fields; the strings and G of the chain are unchanged."""
    L, Lt, ng = 2, 20, 5
    m, x, orc = ft_pair(L, Lt, ng)
    orc.init(1)
    g0 = orc.g.copy()
    xn = x + 0.3 * np.random.default_rng(8).normal(size=x.shape)
    orc.set_fields(xn)
    orc.init_expkv()
    w = orc.cal_new_weight()
    check_lncw_dense(m, xn, w)
    assert np.array_equal(orc.g, g0)
    orc.init(1)
    assert abs(orc.ln_cw.real - w.real) < 1e-12 * abs(w.real)


@pytest.mark.parametrize("distance", [1.0, 0.05])
def test_ft_host_global_moves_match_oracle(distance):
    """The host's GlobalMoves on an oracle-backed finite-temperature chain
    against the oracle's update_global: same accepts, fields and ln_cw."""
    from pqmc_allinone_amd.host import GLOBAL_KINDS, GlobalMoves, Ran2
    from test_global_moves import KIND_ID, _OracleChain
    m, x, orc = ft_pair(4, 20, 5, seed=21)
    _, _, twin = ft_pair(4, 20, 5, seed=21)
    orc.init(1)
    twin.init(1)
    gm = GlobalMoves(_OracleChain(m, twin), Ran2(99), distance)
    ro = O.Ran2Stream(99)
    n_acc = 0
    for _ in range(4):
        flags_g = gm.run()
        flags_o = [orc.update_global(KIND_ID[k], ro, distance) for k in GLOBAL_KINDS]
        if any(flags_o):
            orc.init(1)
        assert flags_g == flags_o
        n_acc += sum(flags_o)
        assert np.array_equal(twin.get_fields(), orc.get_fields())
        assert abs(twin.ln_cw - orc.ln_cw) < 1e-9 * abs(orc.ln_cw)
    if distance < 0.1:
        assert n_acc > 0


@pytest.mark.gpu
@pytest.mark.parametrize("L,Lt,ng", [(2, 30, 5), (4, 20, 5), (6, 40, 10)])
def test_gpu_ft_global_weight_matches_oracle(L, Lt, ng):
    from pqmc_allinone_amd.host import GLOBAL_KINDS
    from test_global_moves import KIND_ID
    m, x, orc = ft_pair(L, Lt, ng)
    orc.init(1)
    ch = ft_chain(m, x)
    ch.init_g(1)
    g0 = ch.g_h
    for kind in GLOBAL_KINDS:
        xn, _ = orc.propose_global(KIND_ID[kind], O.Ran2Stream(11))
        orc.set_fields(xn)
        orc.init_expkv()
        wo = orc.cal_new_weight()
        print(f"ft global weight L={L} Lt={Lt} ng={ng} kind={kind}", flush=True)
        wg = ch.global_weight(xn)
        assert same_lncw(wg, wo, m.Ns), (kind, wg, wo)
    assert np.array_equal(ch.get_fields(), x)
    assert np.array_equal(ch.g_h, g0)
    ch.close()


@pytest.mark.gpu
def test_gpu_ft_sweeps_with_global_moves_match_oracle():
    """Four loops with the global update every second loop (a small move
    distance, so moves are accepted and init_g runs in the loop): identical
    global and local accepts, fields; G to 1e-9."""
    from pqmc_allinone_amd.host import GLOBAL_KINDS, GlobalMoves, Ran2
    from test_global_moves import KIND_ID
    L, Lt, ng, dist = 4, 40, 10, 0.05
    m, x, orc = ft_pair(L, Lt, ng, seed=77)
    orc.init(1)
    ch = ft_chain(m, x)
    ch.init_g(1)
    rg, ro = Ran2(2024), O.Ran2Stream(2024)
    gm = GlobalMoves(ch, rg, dist)
    forward = True
    for loop in range(1, 5):
        forward = not forward
        if not forward and (loop + 1) % 2 == 0:
            print(f"ft global loop {loop}", flush=True)
            flags_g = gm.run()
            flags_o = [orc.update_global(KIND_ID[k], ro, dist) for k in GLOBAL_KINDS]
            if any(flags_o):
                orc.init(1)
            assert flags_g == flags_o, loop
        u = rg.draws(2 * m.Nb * Lt)
        ro.draws(2 * m.Nb * Lt)
        print(f"ft sweep loop {loop} forward={forward}", flush=True)
        ch.sweep(forward, u)
        acc_o, _ = orc.sweep(forward, u)
        assert np.array_equal(ch.accepts(), acc_o), loop
        assert np.array_equal(ch.get_fields(), orc.get_fields()), loop
        assert rel(ch.g_h, orc.g) < 1e-9, loop
    assert gm.accepted > 0 and gm.accepted + gm.rejected == 6
    ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("L,Lt,ng,nsweeps", [(4, 1000, 10, 2), (8, 1000, 10, 1)])
def test_gpu_ft_large_beta_matches_oracle(L, Lt, ng, nsweeps):
    """beta = 100, where the scales of B square past the double range: the
    device's rescaled, pre-pivoted strings against the oracle's, at init and
    over sweeps (identical accepts and fields, G to 1e-9)."""
    m, x, orc = ft_pair(L, Lt, ng, seed=7)
    orc.init(1)
    ch = ft_chain(m, x)
    print(f"ft large beta L={L} Lt={Lt} init", flush=True)
    ch.init_g(1)
    e = rel(ch.g_h, orc.g)
    print(f"  rel G err {e:.3e} lncw {ch.ln_cw} / {orc.ln_cw}", flush=True)
    assert e < 1e-9
    assert same_lncw(ch.ln_cw, orc.ln_cw, m.Ns)
    rng = O.Ran2Stream(3)
    fw = True
    for sweep in range(nsweeps):
        fw = not fw
        u = rng.draws(2 * m.Nb * Lt)
        print(f"ft large beta L={L} sweep={sweep} forward={fw}", flush=True)
        ch.sweep(fw, u)
        acc_o, _ = orc.sweep(fw, u)
        assert np.array_equal(ch.accepts(), acc_o), sweep
        assert np.array_equal(ch.get_fields(), orc.get_fields()), sweep
        e = rel(ch.g_h, orc.g)
        print(f"  rel G err {e:.3e}", flush=True)
        assert e < 1e-9, sweep
    ch.close()
