"""Child process of test_mixed_batch.py: the mixed batch of 128 independent
chains (32 each at L = 8, 10, 12, 14; Lt cycling through 400..1000; seeds
1000 + i) advanced by one pqmc_sweep_many sweep.  Prints one JSON line.  A
separate process, so that GPU_MAX_HW_QUEUES (read when the HIP runtime
starts) gives each concurrent chain a hardware queue of its own."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402
from pqmc_allinone_amd.host import Chain, Ran2, sweep_many  # noqa: E402
from pqmc_allinone_amd.model import ModelParams, build_model, mixed_chain_specs, synthetic_inputs  # noqa: E402

n, threads, bph = (int(v) for v in sys.argv[1:4])
sample = [int(v) for v in sys.argv[4].split(",")]
specs = mixed_chain_specs(n)
models = {}


def inputs(i):
    L, Lt, seed = specs[i]
    if (L, Lt) not in models:
        models[(L, Lt)] = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=10, ne=L * L))
    m = models[(L, Lt)]
    x, sl = synthetic_inputs(m, seed)
    return m, x, sl, Ran2(seed).draws(2 * m.Nb * Lt)


def chain(m, x, sl):
    ch = Chain(m, helper_bonds=bph)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    return ch


t0 = time.perf_counter()
chains, draws = [], []
for i in range(n):
    m, x, sl, u = inputs(i)
    chains.append(chain(m, x, sl))
    draws.append(u)
t_init = time.perf_counter() - t0
print(f"init of {n} chains: {t_init:.1f} s", file=sys.stderr, flush=True)
t0 = time.perf_counter()
meas = sweep_many(chains, [False] * n, draws, measure=True, max_concurrent=threads)
t_sweep = time.perf_counter() - t0
print(f"one sweep of {n} chains: {t_sweep:.1f} s", file=sys.stderr, flush=True)
out = {"n": n, "init_s": t_init, "sweep_s": t_sweep,
       "lt_sum": int(sum(s[1] for s in specs)),
       "drift_g": [c.drift()[0] for c in chains], "drift_lncw": [c.drift()[1] for c in chains],
       "err_fast_max": [c.counters()["err_fast_max"] for c in chains],
       "sync_timeouts": int(sum(c.counters()["sync_timeouts"] for c in chains)), "alone": [], "oracle": None,
       "large_drift": []}
for i, c in enumerate(chains):
    h = c.drift_history()
    if h[:, 0].max() > 1e-8:
        k = int(np.argmax(h[:, 0]))
        out["large_drift"].append({"i": i, "L": specs[i][0], "Lt": specs[i][1], "stab": k, "n": int(len(h)),
                                   "drift": float(h[k, 0]), "above_1e-8": int((h[:, 0] > 1e-8).sum())})
# a chain with a large drift is swept again alone
for d in out["large_drift"][:3]:
    if d["i"] not in sample:
        sample.append(d["i"])
for i in sample:
    m, x, sl, u = inputs(i)
    a = chain(m, x, sl)
    ma = a.sweep(False, u, measure=True)
    out["alone"].append({"i": i, "L": specs[i][0], "Lt": specs[i][1], "drift_alone": a.drift()[0],
                         "acc_equal": bool(np.array_equal(a.accepts(), chains[i].accepts())),
                         "g_maxdiff": float(np.abs(a.g_h - chains[i].g_h).max()),
                         "meas_maxdiff": float(np.abs(ma - meas[i]).max())})
    a.close()
# the smallest lattice's first chain against the oracle
i = 0
m, x, sl, u = inputs(i)
L, Lt, _ = specs[i]
orc = O.Oracle(L, L, Lt, 10, L * L)
orc.set_fields(x)
orc.set_trial(sl)
orc.init(1)
acc_o, mo = orc.sweep(False, u, measure=True)
g_o = orc.g
out["oracle"] = {"i": i, "L": L, "Lt": Lt, "acc_equal": bool(np.array_equal(acc_o, chains[i].accepts())),
                 "g_rel": float(np.abs(chains[i].g_h - g_o).max() / np.abs(g_o).max()),
                 "meas_rel": float(np.abs(meas[i] - mo).max() / np.abs(mo).max())}
for c in chains:
    c.close()
print(json.dumps(out), flush=True)
