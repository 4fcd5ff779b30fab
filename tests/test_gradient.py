"""Field gradient of the fermionic log-weight, d ln_cw / dx for every phonon
field (no counterpart in the reference; the adjoint that hybrid Monte Carlo
needs).  The oracle's backward pass is pinned by central finite differences
of the exact full-determinant weight (cal_new_weight_T0,
src/mod_evolution.f90:359-408, and its finite-temperature counterpart), and
the bond-exponential derivative by finite differences of get_expKV_epsocz
(src/mod_phonon_field.f90:534-569).  GPU tests compare the C ABI's gradient
with the oracle's and check the chain is left where the oracle leaves it.
"""
import numpy as np
import pytest

import oracle as O
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

SP_TWO_PI = float(2 * np.float32(np.pi))
EPS = 1e-6

# (L, Lt, ngroup, ne, tr_weight, ncopy, bc_phase, ft)
CASES = [
    (2, 8, 2, 4, False, 1, (0.0, 0.0), False),
    (2, 9, 3, 4, True, 2, (0.3, 0.1), False),
    (4, 10, 5, 16, False, 1, (0.25, 0.0), False),
    (2, 8, 4, 4, False, 1, (0.2, 0.0), True),
]


def make(L, Lt, ngroup, ne, tr_weight, ncopy, bc, ft, seed=9):
    p = ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ngroup, ne=ne, bc_phase=bc, ncopy=ncopy, tr_weight=tr_weight)
    m = build_model(p)
    x, sl = synthetic_inputs(m, seed)
    orc = O.Oracle(L, L, Lt, ngroup, ne, bc_phase=bc, ncopy=ncopy, tr_weight=tr_weight, ft=ft)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init_expkv()
    return m, x, sl, orc


def weight_at(orc, x):
    orc.set_fields(x)
    orc.init_expkv()
    return orc.cal_new_weight()


def wrap_phase(d):
    """an Im difference of two ln_cw values brought into (-pi, pi]"""
    return (d + np.pi) % SP_TWO_PI - np.pi


def test_dexpkv_matches_finite_differences():
    _, _, _, orc = make(2, 4, 2, 4, False, 1, (0.3, 0.1), False)
    _, _, cross, _ = orc.tables()
    assert cross.any()
    for f in range(1, 5):
        for b in range(orc.npla):
            for xv in (-1.7, 0.0, 0.4, 2.3):
                fd = (orc.expkv(f, b, xv + EPS, 0) - orc.expkv(f, b, xv - EPS, 0)) / (2 * EPS)
                an = orc.dexpkv(f, b, xv)
                assert np.abs(fd - an).max() <= 1e-8, (f, b, xv)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"L{c[0]}Lt{c[1]}ng{c[2]}tr{int(c[4])}nc{c[5]}ft{int(c[7])}")
def test_oracle_gradient_matches_finite_differences(case):
    """Central differences of the exact weight at eps = 1e-6 on 24 seeded
    fields (and every field of the 2x2 Lt=8 case): relative error <= 1e-6
    above the differences' own rounding floor, 1e-9 |ln_cw|."""
    m, x, sl, orc = make(*case)
    orc.init(1)
    g = orc.gradient()
    assert g.shape == (m.p.Lt, m.Nb) and np.isfinite(g).all()
    if case[4]:
        assert np.abs(g.imag).max() == 0.0  # ln_cw + conj(ln_cw) is real
    rng = np.random.default_rng(17)
    idx = [(t, i) for t in range(m.p.Lt) for i in range(m.Nb)]
    if len(idx) > 64:
        idx = [idx[k] for k in rng.choice(len(idx), 24, replace=False)]
    floor = 1e-9 * max(1.0, abs(weight_at(orc, x)))
    for t, i in idx:
        xp = x.copy()
        xp[t, i] += EPS
        wp = weight_at(orc, xp)
        xp[t, i] -= 2 * EPS
        wm = weight_at(orc, xp)
        fd = complex((wp.real - wm.real) / (2 * EPS), wrap_phase(wp.imag - wm.imag) / (2 * EPS))
        assert abs(fd - g[t, i]) <= 1e-6 * abs(g[t, i]) + floor, (t, i, fd, g[t, i])


def test_oracle_gradient_leaves_the_chain_as_a_backward_sweep():
    """After the pass the chain is where a backward sweep that accepts nothing
    leaves it: its G equals a fresh init at the bottom and the next forward
    sweep runs from there."""
    m, x, sl, orc = make(2, 8, 2, 4, False, 1, (0.0, 0.0), False)
    orc.init(1)
    orc.gradient()
    g_after = orc.g
    m2, _, _, o2 = make(2, 8, 2, 4, False, 1, (0.0, 0.0), False)
    o2.init(1)
    for tc in range(1, m.p.Lt + 1):
        t = m.p.Lt - tc + 1
        for f in (4, 3, 2, 1):
            o2.wrap_g(t, f, False)
        if tc % m.p.ngroup == 0:
            o2.get_g_scratch(t, False)
    assert np.abs(g_after - o2.g).max() == 0.0
    assert np.abs(orc.q_string(1) - o2.q_string(1)).max() == 0.0


def test_gradient_of_the_bosonic_action_matches_finite_differences():
    from pqmc_allinone_amd.host import bosonic_action, bosonic_gradient
    p = ModelParams(L1=2, L2=2, Lt=7, ngroup=7, ne=4, biased=0.3)
    m = build_model(p)
    x, _ = synthetic_inputs(m, 4)
    g = bosonic_gradient(m, x)
    for t, i in [(0, 0), (3, 5), (6, 2)]:
        xp = x.copy()
        xp[t, i] += EPS
        sp = bosonic_action(m, xp)
        xp[t, i] -= 2 * EPS
        sm = bosonic_action(m, xp)
        assert abs((sp - sm) / (2 * EPS) - g[t, i]) <= 1e-6 * max(1.0, abs(g[t, i]))


def test_leapfrog_is_reversible_and_second_order():
    """On S = x^T A x / 2: reversing the momentum retraces the trajectory, and
    the energy error of a fixed-length trajectory falls 4x per halved step."""
    from pqmc_allinone_amd.host import leapfrog
    rng = np.random.default_rng(2)
    a = rng.normal(size=(6, 6))
    A = a @ a.T + 6 * np.eye(6)
    force = lambda x: -(A @ x.ravel()).reshape(x.shape)
    x0, p0 = rng.normal(size=(2, 3)), rng.normal(size=(2, 3))
    H = lambda x, p: 0.5 * x.ravel() @ A @ x.ravel() + 0.5 * (p * p).sum()
    x1, p1 = leapfrog(x0, p0, force, 0.05, 10)
    xb, pb = leapfrog(x1, -p1, force, 0.05, 10)
    assert np.abs(xb - x0).max() < 1e-13 and np.abs(pb + p0).max() < 1e-13
    d1 = abs(H(*leapfrog(x0, p0, force, 0.02, 20)) - H(x0, p0))
    d2 = abs(H(*leapfrog(x0, p0, force, 0.01, 40)) - H(x0, p0))
    assert 3.0 < d1 / d2 < 5.0


# ---------------------------------------------------------------- GPU

def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def gpu_chain(m, x, sl, ft):
    from pqmc_allinone_amd.host import Chain
    ch = Chain(m, finite_temperature=ft)
    ch.set_fields(x)
    if not ft:
        ch.set_trial(sl)
    return ch


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    (4, 20, 5, 16, False, 1, (0.0, 0.0), False),
    (4, 18, 3, 16, True, 2, (0.3, 0.1), False),
    (6, 40, 10, 36, False, 1, (0.25, 0.0), False),
    (4, 20, 5, 16, False, 1, (0.2, 0.0), True),
], ids=lambda c: f"L{c[0]}Lt{c[1]}ng{c[2]}tr{int(c[4])}ft{int(c[7])}")
def test_gpu_gradient_matches_oracle(case):
    """pqmc_gradient against the oracle's backward pass on the same inputs:
    gradient and the G it leaves within 1e-10 relative, and a forward sweep
    from there keeps the oracle's accept sequence."""
    import oracle as O2
    m, x, sl, orc = make(*case)
    print(f"gradient case {case}", flush=True)
    orc.init(1)
    ch = gpu_chain(m, x, sl, case[7])
    ch.init_g(1)
    g_ref = orc.gradient()
    g = ch.gradient()
    e = rel(g, g_ref)
    print(f"  rel grad err {e:.3e}  max|grad| {np.abs(g_ref).max():.3e}", flush=True)
    assert e < 1e-10
    assert rel(ch.g_h, orc.g) < 1e-10
    with pytest.raises(Exception, match="bottom"):
        ch.gradient()  # G now sits where a backward sweep ends
    u = O2.Ran2Stream(31).draws(2 * m.Nb * m.p.Lt)
    acc_o, _ = orc.sweep(True, u)
    ch.sweep(True, u)
    assert (ch.accepts() == acc_o).all()
    ch.close()


@pytest.mark.gpu
def test_gpu_gradient_matches_local_update_ratios_at_14x14_lt1000():
    """The 392,000-field chain (14x14, Lt=1000, ngroup 10), 64 seeded fields:
    the gradient against the derivative of the log of the local-update ratio
    R(d) = det(I + G_SS Delta(d)), Delta(d) = E(x)^-1 E(x + d) - I
    (src/mod_update.f90:97-104), taken by central differences at d = 1e-6
    with G where the local update of that bond would see it: after the
    two-way wrap of its family at its slice on a backward pass without
    updates, with G recomputed every 10 slices (the path the gradient pass
    takes).  R(d) - 1 is O(d), so the difference has no cancellation; the
    exponentials are scipy's expm of the bond Hamiltonian, independent of the
    closed form.  Relative error <= 1e-6."""
    import bruteforce as BF
    p = ModelParams(L1=14, L2=14, Lt=1000, ngroup=10, ne=196)
    m = build_model(p)
    x, sl = synthetic_inputs(m, 7)
    ch = gpu_chain(m, x, sl, False)
    print("gradient 14x14 Lt=1000: pass", flush=True)
    ch.init_g(1)
    grad = ch.gradient()
    ch.init_g(1)
    rng = np.random.default_rng(64)
    targets = {}
    for _ in range(64):
        t, f, b = int(rng.integers(1, 1001)), int(rng.integers(4)), int(rng.integers(m.npla))
        targets.setdefault((t, f), []).append(b)
    d = 1e-6
    worst = 0.0
    print("walk down the slices", flush=True)
    for tc in range(1, 1001):
        t = 1001 - tc
        for f in (3, 2, 1, 0):
            ch.right_evolve(t, f + 1)
            if (t, f) not in targets:
                continue
            G = ch.g_h
            for b in targets[(t, f)]:
                S = m.sites[f, b]
                fld = m.bfi[f, b]
                xv = x[t - 1, fld]
                einv = BF.bond_exp(m, f, b, xv, inv=True)
                lr = []
                for s in (d, -d):
                    delta = einv @ BF.bond_exp(m, f, b, xv + s) - np.eye(4)
                    lr.append(np.log(np.linalg.det(np.eye(4) + G[np.ix_(S, S)] @ delta)))
                ref = (lr[0] - lr[1]) / (2 * d)
                got = grad[t - 1, fld]
                e = abs(got - ref) / abs(ref)
                worst = max(worst, e)
                assert e <= 1e-6, (t, f, b, got, ref)
        if tc % 10 == 0:
            ch.get_g_scratch(t, False)
    print(f"  worst relative error {worst:.2e} over 64 fields", flush=True)
    ch.close()


@pytest.mark.gpu
def test_gpu_gradients_of_16_chains_at_14x14_lt1000_match_serial():
    """Sixteen 392,000-field chains (14x14, Lt=1000) computing their gradients
    at once, one host thread and one stream each, give bit for bit what each
    gives alone."""
    from concurrent.futures import ThreadPoolExecutor
    p = ModelParams(L1=14, L2=14, Lt=1000, ngroup=10, ne=196)
    m = build_model(p)
    x, sl = synthetic_inputs(m, 7)
    chains, serial = [], []
    for k in range(16):
        ch = gpu_chain(m, x + 1e-3 * k, sl, False)
        ch.init_g(1)
        serial.append(ch.gradient())
        ch.init_g(1)
        chains.append(ch)
    print("concurrent gradients of 16 chains", flush=True)
    with ThreadPoolExecutor(16) as pool:
        conc = list(pool.map(lambda c: c.gradient(), chains))
    for k, (a, b) in enumerate(zip(serial, conc)):
        assert np.array_equal(a, b), k
    for ch in chains:
        ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("ft", [False, True], ids=["projector", "finite_temperature"])
def test_gpu_gradient_matches_finite_differences_at_14x14_lt1000(ft):
    """The 392,000-field chain (14x14, Lt=1000, ngroup 10): central differences
    of pqmc_global_weight at eps = 1e-4 on 8 seeded fields agree with the
    gradient to 1e-6 relative plus the rounding floor of the differences.
    |ln_cw| is about 3.2e4 here, so each weight carries a rounding error of
    some 1e-15 |ln_cw| (measured: at most 7e-16) and the difference quotient
    an absolute error of about 1e-14 |ln_cw| / eps = 3e-6 at this eps: a
    coarse check of the whole-weight derivative, while the local-update ratio
    test above checks 64 fields to 1e-6 relative without that cancellation."""
    p = ModelParams(L1=14, L2=14, Lt=1000, ngroup=10, ne=196)
    m = build_model(p)
    x, sl = synthetic_inputs(m, 7)
    ch = gpu_chain(m, x, sl, ft)
    print(f"gradient 14x14 Lt=1000 ft={ft}: init", flush=True)
    ch.init_g(1)
    g = ch.gradient().real
    assert g.shape == (1000, m.Nb) and np.isfinite(g).all()
    w0 = ch.global_weight(x)
    floor = 1e-14 * abs(w0.real) / 1e-4
    rng = np.random.default_rng(5)
    eps = 1e-4
    for k in range(8):
        t, i = int(rng.integers(1000)), int(rng.integers(m.Nb))
        xp = x.copy()
        xp[t, i] += eps
        wp = ch.global_weight(xp).real
        xp[t, i] -= 2 * eps
        wm = ch.global_weight(xp).real
        fd = (wp - wm) / (2 * eps)
        print(f"  field ({t},{i}) fd {fd:.9e} grad {g[t, i]:.9e}", flush=True)
        assert abs(fd - g[t, i]) <= 1e-6 * abs(g[t, i]) + floor
    ch.close()


@pytest.mark.gpu
def test_gpu_hybrid_mc_trajectory():
    """With the device gradient: a reversed trajectory returns to its start,
    the energy error falls about 4x per halved step, and HMC moves with the
    exact-determinant test are accepted and leave a chain whose G matches a
    fresh init of the oracle at the new fields, and a local backward sweep
    from there keeps the oracle's accept sequence."""
    from pqmc_allinone_amd.host import HybridMC
    m, x, sl, orc = make(4, 20, 5, 16, False, 1, (0.0, 0.0), False)
    ch = gpu_chain(m, x, sl, False)
    ch.init_g(1)
    hmc = HybridMC(ch, step=0.05, nsteps=10, seed=3)
    rng = np.random.default_rng(8)
    p0 = rng.standard_normal(x.shape)
    print("hmc reversibility", flush=True)
    x1, p1 = hmc.trajectory(x, p0)
    xb, pb = hmc.trajectory(x1, -p1)
    assert np.abs(xb - x).max() < 1e-9 and np.abs(pb + p0).max() < 1e-9
    print("hmc step scaling", flush=True)
    w0 = ch.global_weight(x)
    dh = []
    for step, n in ((0.02, 10), (0.01, 20)):
        hmc.step, hmc.nsteps = step, n
        xe, pe = hmc.trajectory(x, p0)
        dh.append(abs(hmc.hamiltonian(xe, pe, ch.global_weight(xe)) - hmc.hamiltonian(x, p0, w0)))
    print(f"  |dH| {dh}", flush=True)
    assert 2.5 < dh[0] / dh[1] < 6.0
    hmc.step, hmc.nsteps = 0.05, 10
    ch.set_fields(x)
    ch.init_g(1)
    print("hmc updates", flush=True)
    for _ in range(4):
        hmc.update()
    print(f"  accepted {hmc.accepted} rejected {hmc.rejected} last dH {hmc.last_dH:.3e}", flush=True)
    assert hmc.accepted >= 1
    xf = ch.get_fields()
    orc.set_fields(xf)
    orc.init(1)
    assert rel(ch.g_h, orc.g) < 1e-10
    # mixed with the local sweeps: the chain is ready for a backward sweep
    import oracle as O2
    u = O2.Ran2Stream(77).draws(2 * m.Nb * m.p.Lt)
    acc_o, _ = orc.sweep(False, u)
    ch.sweep(False, u)
    assert (ch.accepts() == acc_o).all() and acc_o.any()
    assert rel(ch.g_h, orc.g) < 1e-10
    ch.close()


@pytest.mark.gpu
def test_gpu_autograd_function():
    """fermion_log_weight: the value is the exact weight, backward gives the
    oracle's gradient."""
    import torch
    from pqmc_allinone_amd.autograd import fermion_log_weight
    m, x, sl, orc = make(4, 20, 5, 16, False, 1, (0.0, 0.0), False)
    orc.init(1)
    g_ref = orc.gradient().real
    ch = gpu_chain(m, x, sl, False)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    print("autograd forward", flush=True)
    w = fermion_log_weight(xt, ch)
    assert abs(w.item() - ch.global_weight(x).real) <= 1e-9 * max(1.0, abs(w.item()))
    w.backward()
    assert rel(xt.grad.numpy(), g_ref) < 1e-10
    ch.close()


def test_bosonic_mass_is_the_action_hessian():
    """BosonicMass: K p equals the bosonic action's Hessian times p (the
    action is quadratic), sampled momenta have covariance K, and with the
    bosonic force alone, from an equilibrium x, a Fourier-accelerated
    trajectory keeps the energy far better than a unit-mass one of the same
    step."""
    from pqmc_allinone_amd.host import BosonicMass, bosonic_action, bosonic_gradient, leapfrog
    p = ModelParams(L1=2, L2=2, Lt=40, ngroup=8, ne=4, biased=0.2)
    m = build_model(p)
    mass = BosonicMass(m)
    rng = np.random.default_rng(6)
    v = rng.normal(size=(40, m.Nb))
    z = np.zeros_like(v)
    hv = bosonic_gradient(m, v) - bosonic_gradient(m, z)
    assert np.abs(mass.inverse(hv) - v).max() < 1e-12
    ps = np.stack([mass.sample(rng, (40, 1))[:, 0] for _ in range(4000)])
    K = np.linalg.inv(np.array([mass.inverse(e[:, None])[:, 0] for e in np.eye(40)]))
    assert np.abs(np.cov(ps.T) - K).max() < 0.15 * np.abs(K).max()
    force = lambda x: -bosonic_gradient(m, x)
    x0 = mass._apply(rng.standard_normal((40, m.Nb)), 1 / np.sqrt(mass.khat))  # x ~ exp(-S_bose)
    p_acc, p_unit = mass.sample(rng, x0.shape), rng.normal(size=x0.shape)
    x1, p1 = leapfrog(x0, p_acc, force, 0.05, 10, mass.inverse)
    d_acc = abs(bosonic_action(m, x1) + mass.kinetic(p1) - bosonic_action(m, x0) - mass.kinetic(p_acc))
    x2, p2 = leapfrog(x0, p_unit, force, 0.05, 10)
    d_unit = abs(bosonic_action(m, x2) + 0.5 * (p2 ** 2).sum() - bosonic_action(m, x0) - 0.5 * (p_unit ** 2).sum())
    print(f"|dH| Fourier-accelerated {d_acc:.3e}, unit mass {d_unit:.3e}")
    assert d_acc < 0.05 and d_unit > 10 * d_acc


@pytest.mark.gpu
def test_gpu_gradients_of_concurrent_chains_match_serial():
    """Four chains of different fields computing their gradients at once, one
    host thread and one stream each, give bit for bit the gradients each
    computes alone."""
    from concurrent.futures import ThreadPoolExecutor
    p = ModelParams(L1=6, L2=6, Lt=40, ngroup=10, ne=36)
    m = build_model(p)
    chains, serial = [], []
    for k in range(4):
        x, sl = synthetic_inputs(m, 20 + k)
        ch = gpu_chain(m, x, sl, False)
        ch.init_g(1)
        serial.append(ch.gradient())
        ch.init_g(1)
        chains.append(ch)
    print("concurrent gradients of 4 chains", flush=True)
    with ThreadPoolExecutor(4) as pool:
        conc = list(pool.map(lambda c: c.gradient(), chains))
    for a, b in zip(serial, conc):
        assert np.array_equal(a, b)
    for ch in chains:
        ch.close()
