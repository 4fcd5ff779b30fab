"""The mixed batch: 128 independent chains of uneven lattice and Lt (32
each at L = 8, 10, 12, 14; Lt cycling through 400, 600, 800, 1000; seeds
1000 + i) advanced by one sweep of pqmc_sweep_many, 16 at a time, four bonds
per column-owning workgroup.  Every stabilization of every chain stays
within the drift bound; sampled chains agree bit for bit with the same chain
swept alone; the smallest lattice's chain agrees with the oracle."""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.gpu
def test_mixed_batch_of_128_chains():
    env = dict(os.environ, GPU_MAX_HW_QUEUES="20")
    print("child: 128 mixed chains, one sweep_many sweep, 16 threads, 4 bonds per helper", flush=True)
    r = subprocess.run([sys.executable, os.path.join(HERE, "_mixed_batch.py"), "128", "16", "4", "0,45,77,127"],
                       capture_output=True, text=True, timeout=280, env=env)
    print(r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print("chains with a G drift above 1e-8:", out["large_drift"], flush=True)
    print("swept alone:", out["alone"], flush=True)
    assert out["sync_timeouts"] == 0
    # the reference algorithm itself drifts by up to 3e-6 on these chains
    # (the oracle at chain 79's 30th stabilization, after an accepted ratio of
    # 1.9e-4); every stabilization of every chain stays within 1e-4
    assert max(out["drift_g"]) <= 1e-4 and max(out["err_fast_max"]) <= 1e-4
    for a in out["alone"]:
        assert a["acc_equal"] and a["g_maxdiff"] == 0.0 and a["meas_maxdiff"] == 0.0, a
    o = out["oracle"]
    assert o["acc_equal"] and o["g_rel"] < 1e-10 and o["meas_rel"] < 1e-10, o
