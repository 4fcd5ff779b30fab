"""The oracle against independent brute-force references (tests/bruteforce.py).

Pins, per function of the oracle:
* get_expKV_epsocz (closed form) vs scipy expm of -dtau H_b, incl. BC phases;
* left/right_evolve two-way vs dense similarity transforms;
* init_g_T0 / get_g_scratch_T0 / cal_g_and_det_T0 vs R (L R)^-1 L from dense
  B-string products, and the projector invariants g^2 = g, tr g = ne;
* update_phonon_field: the accept threshold equals the brute-force
  exp(dS + log|det(L B' R)/det(L B R)|) and the rank-4 update equals G from
  scratch with the new field;
* take_measurement: Wick contractions vs Fock-space many-body expectation
  values, k-space companions vs an explicit Fourier sum;
* a full sweep vs the dense-product G at the sweep's end.
"""
import numpy as np
import pytest

import bruteforce as BF
import oracle as O
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs


def setup(L=2, Lt=6, ngroup=2, ne=4, seed=5, lam=0.64, bc=(0.0, 0.0)):
    p = ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ngroup, ne=ne, lam=lam, bc_phase=bc)
    m = build_model(p)
    x, sl = synthetic_inputs(m, seed)
    orc = O.Oracle(L, L, Lt, ngroup, ne, lam=lam, bc_phase=bc)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init(1)
    return m, x, sl, orc


def trial_q(sl):
    q, _ = np.linalg.qr(sl)
    return q


@pytest.mark.parametrize("bc", [(0.0, 0.0), (1.0, 0.0), (0.5, 0.25)])
def test_expkv_closed_form_matches_expm(bc):
    p = ModelParams(L1=4, L2=4, Lt=2, ngroup=1, ne=16, bc_phase=bc)
    m = build_model(p)
    orc = O.Oracle(4, 4, 2, 1, 16, bc_phase=bc)
    rng = np.random.default_rng(0)
    for f in range(4):
        for b in range(m.npla):
            xv = float(rng.normal() * 2)
            e = orc.expkv(f + 1, b, xv, 0)
            ei = orc.expkv(f + 1, b, xv, 1)
            assert np.abs(e - BF.bond_exp(m, f, b, xv)).max() < 1e-13
            assert np.abs(ei - BF.bond_exp(m, f, b, xv, inv=True)).max() < 1e-13
            assert np.abs(e @ ei - np.eye(4)).max() < 1e-13
    # x = 0 reduces to the hopping exponential cosh/sinh of dtau
    e0 = orc.expkv(1, 0, 0.0, 0)
    assert abs(e0[0, 0] - np.cosh(0.1)) < 1e-15
    assert abs(abs(e0[0, 2]) - np.sinh(0.1)) < 1e-15


def test_model_tables_match_host_builder():
    for L in (2, 4, 6):
        p = ModelParams(L1=L, L2=L, Lt=2, ngroup=1, ne=L * L, bc_phase=(1.0, 0.5))
        m = build_model(p)
        orc = O.Oracle(L, L, 2, 1, L * L, bc_phase=(1.0, 0.5))
        s, b, c, ph = orc.tables()
        assert np.array_equal(s, m.sites) and np.array_equal(b, m.bfi) and np.array_equal(c, m.cross)
        assert np.allclose(ph, m.bcph, atol=1e-15)


def test_bond_families_partition_and_geometry():
    L = 6
    m = build_model(ModelParams(L1=L, L2=L, Lt=2, ngroup=1, ne=L * L))
    fields = set()
    for f in range(4):
        assert sorted(m.sites[f].ravel().tolist()) == list(range(m.Ns))
        for b in range(m.npla):
            s = m.sites[f, b]
            c0, c1 = s[0] // 2, s[2] // 2
            assert (s[0] % 2, s[1] % 2, s[2] % 2, s[3] % 2) == (0, 1, 0, 1)
            assert s[1] // 2 == c0 and s[3] // 2 == c1
            x0, y0, x1, y1 = c0 % L, c0 // L, c1 % L, c1 // L
            if f < 2:
                assert (x1, y1) == ((x0 + 1) % L, y0)
            else:
                assert (x1, y1) == (x0, (y0 + 1) % L)
            # checkerboard: families 1,3 start on (x - y) even cells, 2,4 on odd
            assert (x0 - y0) % 2 == f % 2
            fields.add(int(m.bfi[f, b]))
            assert m.bfi[f, b] == 2 * c0 + (0 if f < 2 else 1)
    assert fields == set(range(m.Nb))


@pytest.mark.parametrize("L,Lt,ngroup,ne", [(2, 6, 2, 4), (4, 12, 3, 16), (4, 10, 10, 8)])
def test_init_g_matches_dense_products(L, Lt, ngroup, ne):
    m, x, sl, orc = setup(L=L, Lt=Lt, ngroup=ngroup, ne=ne)
    P = trial_q(sl)
    g_ref = BF.g_after_slice(m, x, P, Lt, 4)
    g = orc.g
    assert np.abs(g - g_ref).max() < 1e-10
    assert np.abs(g @ g - g).max() < 1e-10
    assert abs(np.trace(g) - ne) < 1e-10
    # ln_cw real part: log|det(L R)| with L = P^H, R = B-string P, via dense products
    R = P.copy()
    for t in range(1, Lt + 1):
        R = BF.slice_matrix(m, x, t) @ R
    ref = np.log(abs(np.linalg.det(P.conj().T @ R)))
    # the reference adds log|D_slater| twice (blocks 1 and nblock) on top of log|det(L R)|
    d_sl = np.abs(np.diag(np.linalg.qr(sl)[1]))
    assert abs(orc.ln_cw.real - (ref + np.log(d_sl).sum())) < 1e-8 * max(1.0, abs(ref))


def test_qr_gauge_does_not_change_g():
    """G depends only on the column spaces: a random complex gauge of the trial state leaves it."""
    m, x, sl, orc = setup(L=4, Lt=8, ngroup=4, ne=16)
    g1 = orc.g
    rng = np.random.default_rng(1)
    Z = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    orc2 = O.Oracle(4, 4, 8, 4, 16)
    orc2.set_fields(x)
    orc2.set_trial(sl @ Z)
    orc2.init(1)
    assert np.abs(orc2.g - g1).max() < 1e-10


@pytest.mark.parametrize("forward", [True, False])
def test_two_way_wrap_is_similarity(forward):
    m, x, sl, orc = setup(L=4, Lt=6, ngroup=2, ne=8)
    g = orc.g
    for f in range(1, 5):
        E = BF.family_matrix(m, x, f - 1, 3)
        Ei = BF.family_matrix(m, x, f - 1, 3, inv=True)
        g = E @ g @ Ei if forward else Ei @ g @ E
        orc.wrap_g(3, f, forward)
    assert np.abs(orc.g - g).max() < 1e-11


def test_one_way_evolve_matches_dense():
    m, x, sl, orc = setup(L=4, Lt=6, ngroup=2, ne=8)
    rng = np.random.default_rng(2)
    a = rng.normal(size=(m.Ns, 5)) + 1j * rng.normal(size=(m.Ns, 5))
    for f in (1, 2, 3, 4):
        E = BF.family_matrix(m, x, f - 1, 4)
        assert np.abs(orc.left_evolve(4, a, f, 0) - E @ a).max() < 1e-12
        assert np.abs(orc.right_evolve(4, a.T, f, 0) - a.T @ E).max() < 1e-12


def test_update_threshold_and_rank4_update_match_bruteforce():
    """First bond of family 4 at the last slice, as the backward sweep meets it."""
    L, Lt, ngroup, ne = 4, 6, 2, 8
    m, x, sl, orc = setup(L=L, Lt=Lt, ngroup=ngroup, ne=ne, seed=9)
    P = trial_q(sl)
    t, f = Lt, 4
    orc.wrap_g(t, f, False)  # right_evolve(Lt, g_h, 4, two_way)
    g0 = BF.g_after_slice(m, x, P, t, 3)
    assert np.abs(orc.g - g0).max() < 1e-10
    b = 0
    field = int(m.bfi[f - 1, b])
    # brute-force weight ratio det(L B' R)/det(L B R)
    R = P.copy()
    for tt in range(1, t):
        R = BF.slice_matrix(m, x, tt) @ R
    for ff in range(f - 1):
        R = BF.family_matrix(m, x, ff, t) @ R
    Lm = P.conj().T.copy()
    for tt in range(Lt, t, -1):
        Lm = Lm @ BF.slice_matrix(m, x, tt)
    E = BF.family_matrix(m, x, f - 1, t)
    prob = None
    for u0 in (0.52, 0.48, 0.55, 0.45, 0.6, 0.4, 0.7, 0.3):
        u = np.full(2 * m.npla, 0.999, np.float32)
        u[0] = np.float32(u0)
        rs = np.float32(u[0] * np.float32(2) - np.float32(1))
        jump = m.jump_distance * float(rs)
        x2 = x.copy()
        x2[t - 1, field] += jump
        E2 = BF.family_matrix(m, x2, f - 1, t)
        ratio = np.linalg.det(Lm @ E2 @ R) / np.linalg.det(Lm @ E @ R)
        w = BF.bosonic_action_delta(m, x, field, t, jump) + np.log(abs(ratio))
        pr = min(1.0, np.exp(min(10.0, w)))
        if 1e-3 < pr < 0.999:
            prob = pr
            break
    assert prob is not None, "no unsaturated proposal found"
    # just above the threshold -> reject; just below -> accept
    for factor, expect in ((1 + 1e-6, 0), (1 - 1e-6, 1)):
        o2 = O.Oracle(L, L, Lt, ngroup, ne)
        o2.set_fields(x)
        o2.set_trial(sl)
        o2.init(1)
        o2.wrap_g(t, f, False)
        uu = u.copy()
        uu[1] = np.float32(prob * factor)
        acc = o2.update_phonon_field(f, t, uu)
        assert acc[0] == expect
        if expect:
            # rank-4 update == G from scratch with the new field
            gnew = BF.g_after_slice(m, o2.get_fields(), P, t, 3)
            assert np.abs(o2.g - gnew).max() < 1e-9
            assert o2.get_fields()[t - 1, field] == x[t - 1, field] + jump


def test_sweeps_match_dense_products_and_projector_invariants():
    L, Lt, ngroup, ne = 4, 8, 4, 8
    m, x, sl, orc = setup(L=L, Lt=Lt, ngroup=ngroup, ne=ne, seed=21)
    P = trial_q(sl)
    rng = O.Ran2Stream(77)
    forward = True
    for _ in range(3):
        forward = not forward
        acc, _ = orc.sweep(forward, rng.draws(2 * m.Nb * Lt))
        xf = orc.get_fields()
        if forward:  # ends after slice Lt: g(Lt+1)
            gref = BF.g_after_slice(m, xf, P, Lt, 4)
        else:  # ends before slice 1: g(1)
            gref = BF.g_after_slice(m, xf, P, 1, 0)
        assert np.abs(orc.g - gref).max() < 1e-9
        assert 0.05 < acc.mean() < 0.95
    g = orc.g
    assert np.abs(g @ g - g).max() < 1e-9 and abs(np.trace(g) - ne) < 1e-9


def _wick_value(ops, bra, ket, ls, rs, vl, vr):
    tot = 0j
    for i, (l1, l2) in enumerate(ls):
        for j, (r1, r2) in enumerate(rs):
            op = ops[l1].T @ ops[l2] @ ops[r1].T @ ops[r2]
            tot += vl[i] * vr[j] * BF.mixed_expectation(ops, bra, ket, op)
    return tot


def test_measurement_matches_fock_space():
    """L=2 (8 sites): each measured correlator vs the many-body mixed estimator."""
    L, Lt, ngroup, ne = 2, 4, 2, 4
    m, x, sl, orc = setup(L=L, Lt=Lt, ngroup=ngroup, ne=ne, seed=3)
    P = trial_q(sl)
    t = Lt // 2
    # g_h after init sits after slice Lt: R = B(Lt)...B(1) P, L = P^H
    R = P.copy()
    for tt in range(1, Lt + 1):
        R = BF.slice_matrix(m, x, tt) @ R
    Lm = P.conj().T
    out = orc.measure(t)
    ops = BF.fock_ops(m.Ns)
    ket = BF.slater_ket(ops, R)
    bra = BF.slater_bra(ops, Lm)
    Nc = m.Ncell
    off = {n: lo for n, lo, hi in __import__("pqmc_allinone_amd.host", fromlist=["x"]).observable_layout(Nc)}
    hop = 1.0

    def cells(c):
        x0, y0 = c % L, c // L
        return ((x0 + 1) % L) + L * y0, x0 + L * ((y0 + 1) % L)

    def spin_current(a, b):
        return [(2 * a, 2 * b + 1), (2 * a + 1, 2 * b), (2 * b, 2 * a + 1), (2 * b + 1, 2 * a)]

    vjx = -hop * np.array([-1j, -1j, 1j, 1j])
    vd = np.array([1, 0, 0, 1]) / 2
    vz = np.array([1, 0, 0, -1]) / 2
    for d in range(Nc):
        dx, dy = d % L, d // L
        den = jxx = sz = sc = 0j
        for c1 in range(Nc):
            x1, y1 = c1 % L, c1 // L
            c2 = (x1 + dx) % L + L * ((y1 + dy) % L)
            c1x, _ = cells(c1)
            c2x, _ = cells(c2)
            sm1 = [(2 * c1, 2 * c1), (2 * c1, 2 * c1 + 1), (2 * c1 + 1, 2 * c1), (2 * c1 + 1, 2 * c1 + 1)]
            sm2 = [(2 * c2, 2 * c2), (2 * c2, 2 * c2 + 1), (2 * c2 + 1, 2 * c2), (2 * c2 + 1, 2 * c2 + 1)]
            den += _wick_value(ops, bra, ket, sm1, sm2, vd, vd) / Nc
            sz += _wick_value(ops, bra, ket, sm1, sm2, vz, vz) / Nc
            jxx += _wick_value(ops, bra, ket, spin_current(c1, c1x), spin_current(c2, c2x), vjx, vjx) / Nc
            # SC_SC = <c1u^dag c1d^dag c2d c2u>
            op = ops[2 * c1].T @ ops[2 * c1 + 1].T @ ops[2 * c2 + 1] @ ops[2 * c2]
            sc += BF.mixed_expectation(ops, bra, ket, op) / Nc
        assert abs(out[off["den_den"] + d] - den) < 1e-10
        assert abs(out[off["Sz_Sz"] + d] - sz) < 1e-10
        assert abs(out[off["Jxx_Jxx"] + d] - jxx) < 1e-10
        assert abs(out[off["SC_SC"] + d] - sc) < 1e-10
    assert np.all(out[off["Jxx_Jyy"]: off["Jxx_Jyy"] + Nc] == 0)


def test_kspace_companions_are_fourier_sums():
    m, x, sl, orc = setup(L=4, Lt=6, ngroup=2, ne=8)
    out = orc.measure(3)
    from pqmc_allinone_amd.host import observable_layout
    lay = {n: (lo, hi) for n, lo, hi in observable_layout(m.Ncell)}
    L = 4
    kx = 2 * np.pi * np.arange(L) / L - np.pi
    for name in ("den_den", "Jxx_Jxx", "BFx_BFx", "cJxy_cJxy"):
        r = out[lay[name][0]: lay[name][1]]
        k = out[lay[name + "_k"][0]: lay[name + "_k"][1]]
        for kk in range(m.Ncell):
            ph = np.exp(1j * (kx[kk % L] * (np.arange(m.Ncell) % L) + kx[kk // L] * (np.arange(m.Ncell) // L)))
            assert abs(k[kk] - (ph * r).sum() / m.Ncell) < 1e-12


def test_boson_correlators_match_numpy():
    m, x, sl, orc = setup(L=4, Lt=6, ngroup=2, ne=8)
    t = 3
    out = orc.measure(t)
    from pqmc_allinone_amd.host import observable_layout
    lay = {n: lo for n, lo, hi in observable_layout(m.Ncell)}
    L = 4
    bx = x[t - 1, 0::2].reshape(L, L)  # [y, x]
    by = x[t - 1, 1::2].reshape(L, L)
    for d in range(m.Ncell):
        dx, dy = d % L, d // L
        sh = lambda a: np.roll(np.roll(a, -dy, axis=0), -dx, axis=1)
        assert abs(out[lay["BFx_BFx"] + d] - (bx * sh(bx)).mean()) < 1e-12
        assert abs(out[lay["BFx_BFy"] + d] - (bx * sh(by)).mean()) < 1e-12
        assert abs(out[lay["BFxy_BFxy"] + d] - (bx * by * sh(bx) * sh(by)).mean()) < 1e-12
    assert abs(out[lay["BF_X"]] - x[t - 1].sum() / m.Ncell) < 1e-12


def test_factorizations_follow_lapack():
    """The oracle's qdr and ldq (src/mod_matrixlib.f90:323-360) against
    LAPACK's ZGEQRF and ZGELQF: the same triangular factors, and the same
    sign of zero in the imaginary part of the diagonal, (beta, +0) for QR and
    (beta, -0) for LQ, which fixes the branch of log(d) for negative d."""
    import lapack_ref
    if not lapack_ref.available():
        pytest.skip("no LAPACK with zgeqrf/zgelqf found")
    rng = np.random.default_rng(11)
    nneg = 0
    for m, n in [(7, 4), (9, 9), (12, 5)]:
        a = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        fr, _ = lapack_ref.zgeqrf(a)
        q, d, r = O.zqdr(a)
        dr = np.diag(fr)[:n]
        assert np.allclose(d, dr, rtol=1e-13, atol=0)
        assert np.array_equal(np.signbit(d.imag), np.signbit(dr.imag))
        assert np.allclose(np.triu(fr[:n, :n]) / dr[:, None], r, rtol=0, atol=1e-13)
        b = a.T.copy()  # n x m, n <= m
        fl, _ = lapack_ref.zgelqf(b)
        qq, dl, l = O.zldq(b)
        dlap = np.diag(fl)[:n]
        assert np.allclose(dl, dlap, rtol=1e-13, atol=0)
        assert np.array_equal(np.signbit(dl.imag), np.signbit(dlap.imag))
        assert np.all(np.signbit(dlap.imag))
        assert np.allclose(np.tril(fl[:n, :n]) / dlap[None, :], l, rtol=0, atol=1e-13)
        assert np.allclose(np.log(dl), np.log(dlap), rtol=0, atol=1e-13)
        nneg += int((dl.real < 0).sum())
    assert nneg > 0  # some diagonal is negative, so the branch is exercised
