"""Child process of test_concurrency.py: identical chains advanced together
must agree bit for bit with one chain advanced alone.  Prints one JSON line.
Run as a separate process so that GPU_MAX_HW_QUEUES (read when the HIP
runtime starts) can give every chain of the batch its own hardware queue."""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from pqmc_allinone_amd.host import Chain, Ran2, sweep_batch  # noqa: E402
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs  # noqa: E402

K, L, Lt, nsw, bph = (int(v) for v in sys.argv[1:6])
m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=10, ne=L * L))
x, sl = synthetic_inputs(m, 7)
nd = 2 * m.Nb * Lt


def fresh():
    ch = Chain(m, helper_bonds=bph)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    return ch


draws = []
rng = Ran2(7)
for _ in range(nsw):
    draws.append(rng.draws(nd))
alone = fresh()
ref = []
fw = True
for s in range(nsw):
    fw = not fw
    alone.sweep(fw, draws[s], measure=True)
    ref.append((alone.accepts().copy(), alone.g_h))
alone.close()
chains = [fresh() for _ in range(K)]
out = {"sweeps": [], "err_fast": None}
fw = True
for s in range(nsw):
    fw = not fw
    sweep_batch(chains, [fw] * K, [draws[s]] * K, measure=True)
    out["sweeps"].append({"acc_equal": [bool(np.array_equal(c.accepts(), ref[s][0])) for c in chains],
                          "g_maxdiff": [float(np.abs(c.g_h - ref[s][1]).max()) for c in chains]})
out["err_fast"] = [c.counters()["err_fast_max"] for c in chains]
for c in chains:
    c.close()
print(json.dumps(out), flush=True)
