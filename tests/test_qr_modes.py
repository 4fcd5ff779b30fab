"""The stabilization's Householder QR in its two device forms (pqmc_set_qr_mode):
one launch per factorization (a workgroup per panel, panels handed on by
flags; the default, which every other GPU test runs) and one launch per panel
and per block-reflector application.  Both follow zgeqr2 / zlarfg / zlarft;
they differ only in summation order, so G, the Q and D strings and ln_cw
agree to rounding after init_g, after stabilizations in both directions and
after sweeps with the same draws, which must keep identical accepts.  The
lattices cover one panel (4x4: Ns=32, ne=16), several (8x8, 14x14: 13 panels
of 16 columns, the production shape) and 8-column panels over 800 rows
(20x20), projector and finite temperature (Ns x Ns strings)."""
import numpy as np
import pytest

import oracle as O
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def pair(L, Lt, ng, ft=False, seed=5):
    from pqmc_allinone_amd.host import Chain
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=L * L))
    x, sl = synthetic_inputs(m, seed)
    out = []
    for mode in ("flow", "panels"):
        ch = Chain(m, finite_temperature=ft)
        ch.set_qr_mode(mode)
        ch.set_fields(x)
        if not ft:
            ch.set_trial(sl)
        ch.init_g(1)
        out.append(ch)
    return m, out[0], out[1]


def same_state(m, a, b, ft, tol=1e-11):
    assert rel(a.g_h, b.g_h) < tol
    la, lb = a.ln_cw, b.ln_cw
    assert abs(la.real - lb.real) < tol * max(1.0, abs(lb.real))
    # the phase: each log term of a real D entry is 0 or +-i pi, a rounded
    # zero choosing the branch, and the final mod 2 pi_SP of cal_g_and_det
    # turns a 2 pi difference of the unreduced sums into 2 pi - 2 pi_SP each
    step = 2 * np.pi - 2 * float(np.float32(np.pi))
    d = (la.imag - lb.imag + np.pi) % (2 * np.pi) - np.pi
    k = round(d / step)
    assert abs(d - k * step) < 1e-9 and abs(k) <= m.Ns * (m.nblock + 1), (la, lb)
    if not ft:
        for blk in (1, 2, m.nblock // 2 + 1, m.nblock):
            assert rel(a.q_string(blk), b.q_string(blk)) < tol, blk
        assert rel(a.d_string(), b.d_string()) < tol


@pytest.mark.parametrize("L,Lt,ng,ft", [(4, 40, 10, False), (8, 40, 10, False), (14, 40, 10, False),
                                        (20, 20, 10, False), (4, 40, 10, True), (8, 20, 10, True)])
def test_qr_modes_agree(L, Lt, ng, ft):
    m, a, b = pair(L, Lt, ng, ft)
    print(f"{L}x{L} Lt={Lt} ft={ft}: init_g", flush=True)
    same_state(m, a, b, ft)
    # stabilizations as a backward sweep makes them; the projector also at a
    # forward one (its left strings are all set by init_g; at finite
    # temperature they are the ones the backward sweep leaves behind)
    steps = [(Lt - ng + 1, False), (Lt - 2 * ng + 1, False)] + ([] if ft else [(ng + 1, True)])
    for t, fw in steps:
        ea = a.get_g_scratch(t, fw)
        eb = b.get_g_scratch(t, fw)
        assert abs(ea - eb) < 1e-9 * max(1.0, abs(eb)), (t, fw, ea, eb)
        same_state(m, a, b, ft)
    a.close()
    b.close()


@pytest.mark.parametrize("L,Lt,ntc,ft", [(4, 100, 100, False), (14, 40, 20, False), (4, 40, 40, True),
                                         (8, 20, 20, True)])
def test_qr_modes_same_accepts(L, Lt, ntc, ft):
    m, a, b = pair(L, Lt, 10, ft)
    rng = O.Ran2Stream(77)
    fw = True
    for s in range(2):
        fw = not fw
        u = rng.draws(2 * m.Nb * Lt)
        a.sweep(fw, u, ntc=ntc)
        b.sweep(fw, u, ntc=ntc)
        n = 4 * m.npla * ntc
        assert np.array_equal(a.accepts()[:n], b.accepts()[:n]), s
        assert rel(a.g_h, b.g_h) < 1e-10, s
        print(f"{L}x{L} sweep {s}: drift {a.drift()} vs {b.drift()}", flush=True)
    assert a.counters()["sync_timeouts"] == 0
    a.close()
    b.close()
