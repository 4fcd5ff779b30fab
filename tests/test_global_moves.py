"""Global phonon moves: update_global (src/mod_update.f90:251-467) with the
full-determinant weight cal_new_weight_T0 (src/mod_evolution.f90:359-408).

CPU tests pin the oracle (dense products, brute-force actions, invariants)
and check the host's proposals against it bit for bit; GPU tests compare the
C ABI's weight and whole sweeps with global moves against the oracle.
"""
import numpy as np
import pytest

import bruteforce as BF
import oracle as O
from pqmc_allinone_amd.host import GLOBAL_KINDS, Ran2, propose_global
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

SP_TWO_PI = float(2 * np.float32(np.pi))
KIND_ID = {"kspace": O.GU_KSPACE, "rotate": O.GU_ROTATE, "kspace_time": O.GU_KSPACE_TIME}


def oracle_for(L, Lt, ngroup, ne, seed=5, trial_scale=None, lam=0.64, omega=1.0, biased=0.0):
    p = ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ngroup, ne=ne, lam=lam, omega=omega, biased=biased)
    m = build_model(p)
    x, sl = synthetic_inputs(m, seed)
    if trial_scale is not None:
        sl = sl @ trial_scale
    orc = O.Oracle(L, L, Lt, ngroup, ne, lam=lam, omega=omega, biased=biased)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init_expkv()
    return m, x, sl, orc


def total_action(m, x):
    """delt * sum_t sum_i [M/(2 dt^2) (x_t - x_{t-1})^2 + D/2 x_t^2 - biased x_t], periodic in t."""
    dt = m.p.dtau
    kin = 0.5 * m.M / dt ** 2 * ((x - np.roll(x, 1, axis=0)) ** 2).sum()
    pot = (0.5 * m.D * x ** 2 - m.p.biased * x).sum()
    return dt * (kin + pot)


@pytest.mark.parametrize("Lt,ngroup", [(12, 3), (12, 5)])
def test_oracle_new_weight_matches_dense_product(Lt, ngroup):
    """Re ln_cw = log|det(Q^H B Q)| + 2 sum log|D| for the trial Q D R; ngroup = 5
    leaves a partial last group without an LQ."""
    rng = np.random.default_rng(3)
    scale = np.diag(rng.uniform(0.3, 3.0, 4)) @ (np.eye(4) + 0.3 * rng.normal(size=(4, 4)))
    m, x, sl, orc = oracle_for(2, Lt, ngroup, 4, trial_scale=scale)
    w = orc.cal_new_weight()
    q, r = np.linalg.qr(sl)
    B = np.eye(m.Ns, dtype=complex)
    for t in range(1, Lt + 1):
        B = BF.slice_matrix(m, x, t) @ B
    ref = np.log(abs(np.linalg.det(q.conj().T @ B @ q))) + 2 * np.log(np.abs(np.diag(r))).sum()
    assert abs(sum(np.log(np.abs(np.diag(r))))) > 0.1  # the D terms are exercised
    assert abs(w.real - ref) < 1e-9 * max(1.0, abs(ref))


def test_oracle_new_weight_equals_fresh_init_weight():
    m, x, sl, orc = oracle_for(4, 20, 5, 16)
    orc.init(1)
    w0 = orc.ln_cw
    w = orc.cal_new_weight()
    assert abs(w.real - w0.real) < 1e-10 * abs(w0.real)
    d = (w.imag - w0.imag) / (2 * np.pi)
    assert abs(d - round(d)) < 1e-6


def test_oracle_new_weight_log_branches_follow_lapack():
    """Im ln_cw term by term on the reference's branches: the LQ of
    src/mod_matrixlib.f90:488-503 (ZGELQF) leaves its diagonal as
    (beta, -0) (pinned against LAPACK in test_oracle_physics), so a negative
    beta adds -i pi, and Im is reduced mod 2*pi in REAL(SP) only at the end.
    A +i pi branch would leave a residue of (2 pi - 2 pi_SP) per negative
    beta after that reduction."""
    L, Lt, ng = 4, 36, 5
    m, x, sl, orc = oracle_for(L, Lt, ng, L * L)
    q, d, _ = O.zqdr(sl)
    lmat = q.conj().T
    lnd = np.log(d).sum()
    nneg = 0
    for i, t in enumerate(range(Lt, 0, -1)):
        lmat = lmat @ BF.slice_matrix(m, x, t)
        if (i + 1) % ng == 0:
            qq, r = np.linalg.qr(lmat.conj().T)
            beta = np.diag(r).real
            nneg += int((beta < 0).sum())
            lnd += np.log(np.conj(beta + 0j)).sum()
            lmat = qq.conj().T
    lnd += np.log(d).sum() + np.log(np.linalg.det(lmat @ q))
    w = orc.cal_new_weight()
    assert nneg > 0
    assert abs(w.real - lnd.real) < 1e-10 * abs(lnd.real)
    dd = abs(w.imag - np.fmod(lnd.imag, SP_TWO_PI))
    assert min(dd, abs(dd - SP_TWO_PI)) < 1e-9


@pytest.mark.parametrize("kind", GLOBAL_KINDS)
def test_host_proposals_match_oracle(kind):
    """The host's generate_newfield_global restatement against the oracle's, bit for bit."""
    for L, Lt, biased in [(4, 20, 0.0), (6, 300, 0.3)]:
        m, x, sl, orc = oracle_for(L, Lt, 5, 2 * L * L // 2, omega=1.0, biased=biased)
        for seed in (1, 2, 3):
            xo, pho = orc.propose_global(KIND_ID[kind], O.Ran2Stream(seed))
            xh, phh = propose_global(m, x, kind, Ran2(seed))
            assert np.array_equal(xh, xo), (L, seed)
            assert phh == pho, (L, seed)


@pytest.mark.parametrize("kind", GLOBAL_KINDS)
def test_proposals_follow_the_bosonic_action(kind):
    """-delt*dE equals minus the change of the total bosonic action (kspace, kspace_time);
    rotate maps (x1, x2) of every cell by one orthogonal 2x2 matrix and reports no change
    (as the reference)."""
    m, x, sl, orc = oracle_for(6, 300, 5, 36, omega=1.0, biased=0.25)
    for seed in (4, 5):
        xn, ph = orc.propose_global(KIND_ID[kind], O.Ran2Stream(seed))
        if kind == "rotate":
            assert ph == 0.0
            n0 = x[:, 0::2] ** 2 + x[:, 1::2] ** 2
            n1 = xn[:, 0::2] ** 2 + xn[:, 1::2] ** 2
            assert np.abs(n1 - n0).max() < 1e-12 * n0.max()
            # one orthogonal 2x2 map (a rotation, or a rotation with the
            # second component flipped) for every cell and slice
            X = np.stack([x[:, 0::2].ravel(), x[:, 1::2].ravel()])
            Y = np.stack([xn[:, 0::2].ravel(), xn[:, 1::2].ravel()])
            R = Y @ X.T @ np.linalg.inv(X @ X.T)
            assert np.abs(R @ X - Y).max() < 1e-12 * np.abs(Y).max()
            assert np.abs(R @ R.T - np.eye(2)).max() < 1e-12
            assert abs(abs(R[0, 0]) - abs(R[1, 1])) < 1e-12 and abs(R[0, 0]) > np.cos(0.1 * np.pi) - 1e-12
            continue
        ref = -(total_action(m, xn) - total_action(m, x))
        assert abs(ph - ref) < 1e-9 * max(1.0, abs(ref))
        d = xn - x
        rows = np.nonzero(np.abs(d).max(axis=1) > 0)[0]
        if kind == "kspace":
            assert len(rows) == m.p.Lt
        else:
            assert rows[-1] == m.p.Lt - 1 and rows[0] <= m.p.Lt // 2
        # one staggered pattern shifted into every affected slice
        assert np.abs(d[rows] - d[rows[0]]).max() < 1e-12
        cells = d[rows[0]].reshape(-1, 2)
        cx = np.arange(m.Ncell) % m.p.L1
        cy = np.arange(m.Ncell) // m.p.L1
        sgn = np.where((cx + cy) % 2 == 0, 1.0, -1.0)
        amp = cells / sgn[:, None]
        assert np.abs(np.abs(amp) - np.abs(amp[0, 0])).max() < 1e-5 * abs(amp[0, 0])


class _OracleChain:
    """The four Chain calls GlobalMoves makes, served by a second oracle
    instance, so the host's accept logic runs without a device."""

    def __init__(self, m, orc):
        self.model, self.o = m, orc

    def get_fields(self):
        return self.o.get_fields()

    def set_fields(self, x):
        self.o.set_fields(x)

    def init_g(self, t):
        self.o.init(t)

    @property
    def ln_cw(self):
        return self.o.ln_cw

    def global_weight(self, x):
        keep = self.o.get_fields()
        self.o.set_fields(x)
        self.o.init_expkv()
        w = self.o.cal_new_weight()
        self.o.set_fields(keep)
        self.o.init_expkv()
        return w


@pytest.mark.parametrize("distance", [1.0, 0.05])
def test_host_global_moves_match_oracle(distance):
    """GlobalMoves.run against three oracle update_global calls: same accepts,
    fields and ln_cw over repeated blocks (a small distance makes accepts common)."""
    from pqmc_allinone_amd.host import GlobalMoves
    m, x, sl, orc = oracle_for(4, 40, 5, 16, seed=21)
    _, _, _, twin = oracle_for(4, 40, 5, 16, seed=21)
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
    assert gm.accepted == n_acc and gm.accepted + gm.rejected == 12
    if distance < 0.1:
        assert n_acc > 0


# ---------------------------------------------------------------- GPU


def make_pair(L, Lt, ngroup, ne, seed=1234):
    from pqmc_allinone_amd.host import Chain
    m, x, sl, orc = oracle_for(L, Lt, ngroup, ne, seed=seed)
    orc.init(1)
    ch = Chain(m)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    return m, x, orc, ch


@pytest.mark.gpu
@pytest.mark.parametrize("L,Lt,ngroup", [(4, 100, 10), (8, 40, 20), (4, 36, 5)])
def test_global_weight_matches_oracle(L, Lt, ngroup):
    m, x, orc, ch = make_pair(L, Lt, ngroup, L * L)
    for kind in GLOBAL_KINDS:
        xn, _ = orc.propose_global(KIND_ID[kind], O.Ran2Stream(11))
        orc.set_fields(xn)
        orc.init_expkv()
        wo = orc.cal_new_weight()
        wg = ch.global_weight(xn)
        assert abs(wg.real - wo.real) < 1e-10 * abs(wo.real), kind
        # the same branch of every log term: Im agrees after the reduction
        # mod 2*pi in REAL(SP), up to a wrap at the ends of the range
        d = abs(wg.imag - wo.imag)
        assert min(d, abs(d - SP_TWO_PI)) < 1e-8, kind
    # the chain itself is untouched
    assert np.array_equal(ch.get_fields(), x)


@pytest.mark.gpu
@pytest.mark.parametrize("distance", [1.0, 0.05])
def test_sweeps_with_global_moves_match_oracle(distance):
    """Four loops of src/Main_PQMC.f90:81-156 with the global update on
    (global_update_loop = 2): identical global and local accepts, fields and G.
    The small move distance makes accepted global moves (and init_g) certain."""
    from pqmc_allinone_amd.host import GlobalMoves
    L, Lt, ngroup = 4, 200, 10
    m, x, orc, ch = make_pair(L, Lt, ngroup, L * L, seed=77)
    rg, ro = Ran2(2024), O.Ran2Stream(2024)
    gm = GlobalMoves(ch, rg, distance)
    forward = True
    n_glob = 0
    for loop in range(1, 5):
        forward = not forward
        if not forward and (loop + 1) % 2 == 0:
            flags_g = gm.run()
            flags_o = [orc.update_global(KIND_ID[k], ro, distance) for k in GLOBAL_KINDS]
            if any(flags_o):
                orc.init(1)
            assert flags_g == flags_o, loop
            n_glob += 1
        u = rg.draws(2 * m.Nb * Lt)
        assert np.array_equal(u, ro.draws(2 * m.Nb * Lt))
        ch.sweep(forward, u)
        acc_o, _ = orc.sweep(forward, u)
        acc_g = ch.accepts()
        assert np.array_equal(acc_g, acc_o), loop
        assert np.array_equal(ch.get_fields(), orc.get_fields()), loop
        g, go = ch.g_h, orc.g
        assert np.abs(g - go).max() / np.abs(go).max() < 1e-10, loop
    assert n_glob == 2
    assert gm.accepted + gm.rejected == 6
    if distance < 0.1:
        assert gm.accepted > 0
