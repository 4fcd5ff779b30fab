"""pqmc_gradient_noalloc (the field gradient with the library's fixed device
array as its accumulator, no device allocation) through the C ABI against
pqmc_gradient on the same chain: two calls in a row, the second on a smaller
chain, each bit for bit; a chain beyond the array's 1,048,576 fields is
refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from pqmc_allinone_amd import _lib
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

pytestmark = pytest.mark.gpu


def _chain(L, Lt, ng, seed):
    from pqmc_allinone_amd.host import Chain
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=L * L))
    x, sl = synthetic_inputs(m, seed)
    ch = Chain(m)
    ch.set_fields(x)
    ch.set_trial(sl)
    return m, ch


def _grad(L_, ch, m, fn):
    ch.init_g(1)
    g = np.empty((m.p.Lt, m.Nb), np.complex128)
    rc = getattr(L_, fn)(ch.h, g.ctypes.data)
    _lib.check(rc)
    return g, ch.g_h


def test_noalloc_gradient_matches_pqmc_gradient():
    L_ = _lib.lib()
    for L, Lt, ng, seed in ((6, 60, 10, 21), (4, 20, 5, 22)):  # the second call is the smaller input
        print(f"noalloc gradient {L}x{L} Lt={Lt}", flush=True)
        m, ch = _chain(L, Lt, ng, seed)
        ref, g_ref = _grad(L_, ch, m, "pqmc_gradient")
        got, g_got = _grad(L_, ch, m, "pqmc_gradient_noalloc")
        assert np.array_equal(got, ref)
        assert np.array_equal(g_got, g_ref)
        ch.close()


def test_noalloc_gradient_refuses_more_fields_than_the_array():
    L_ = _lib.lib()
    m, ch = _chain(20, 1400, 10, 3)  # 800 x 1400 = 1,120,000 fields
    g = np.empty((m.p.Lt, m.Nb), np.complex128)
    rc = L_.pqmc_gradient_noalloc(ch.h, g.ctypes.data)
    assert rc == _lib.PQMC_EINVAL
    assert b"exceed" in L_.pqmc_last_error()
    ch.close()
