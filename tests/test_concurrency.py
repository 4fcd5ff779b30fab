"""Chains sharing the device.  A family pass on a backward sweep wraps every
column with the bond exponentials from before the pass's updates; when other
work holds compute units, some of its workgroups start late, and before the
fix the leader could already have replaced an accepted bond's exponentials
by then (a silent, deterministic error).  Five 14x14 chains with two bonds per
helper workgroup, or nine with four, keep the device's compute units nearly
all busy; each must agree bit for bit with the same chain advanced alone,
with no drift above the initial bound."""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.gpu
@pytest.mark.parametrize("nchains,bph", [(5, 2), (9, 4)])
def test_concurrent_chains_match_a_lone_chain(nchains, bph):
    """bph: bonds per column-owning workgroup (a 14x14 pass keeps 50 compute
    units at 2, 26 at 4)."""
    env = dict(os.environ, GPU_MAX_HW_QUEUES=str(nchains + 3))
    print(f"child: {nchains} chains 14x14 Lt=200, 3 sweeps, {bph} bonds per helper", flush=True)
    r = subprocess.run([sys.executable, os.path.join(HERE, "_concurrent_chains.py"), str(nchains), "14", "200", "3",
                        str(bph)], capture_output=True, text=True, timeout=110, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    for s, sw in enumerate(out["sweeps"]):
        assert all(sw["acc_equal"]), (s, sw)
        assert max(sw["g_maxdiff"]) == 0.0, (s, sw)
    assert max(out["err_fast"]) <= 1e-6
