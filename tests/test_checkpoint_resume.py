"""Measurement bins and checkpoint/resume on the device, through both hosts.

- The Fortran driver's bin files against bin means of the oracle's
  measurements on the same chain (1e-10 relative).
- A run checkpointed after loop 4 and resumed in a new process (Fortran) or a
  new chain (Python) against the same run without a stop: identical local and
  global accepts, fields and counters; G, ln_cw and bins to 1e-10.  The run
  has global moves (global_update_loop = 2) and bins after a warm-up.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from pqmc_allinone_amd import fieldio, host, measio
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

L, LT, NG, NE, SEED = 4, 40, 10, 16, 4321
GKINDS = (O.GU_KSPACE, O.GU_ROTATE, O.GU_KSPACE_TIME)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def setup():
    m = build_model(ModelParams(L1=L, L2=L, Lt=LT, ngroup=NG, ne=NE))
    x, sl = synthetic_inputs(m, SEED)
    return m, x, sl


def folder(p):
    (p / "temp").mkdir(parents=True)
    (p / "data").mkdir()
    return str(p) + "/"


def run_driver(tmp_path, name, m, x, sl, nsweeps, outdir, **kw):
    fin, fout = tmp_path / f"{name}.in", tmp_path / f"{name}.out"
    host.write_driver_input(str(fin), m, x, sl, nsweeps=nsweeps, seed=SEED, **kw)
    print(f"driver {name}: nsweeps={nsweeps} {kw}", flush=True)
    r = subprocess.run([host.FORTRAN_DRIVER, str(fin), str(fout), outdir], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return host.read_driver_output(str(fout), m, nsweeps)


needs_driver = pytest.mark.skipif(not os.path.exists(host.FORTRAN_DRIVER),
                                  reason="Fortran driver not built (amdflang missing)")


@pytest.mark.gpu
@needs_driver
def test_fortran_bins_match_oracle(tmp_path):
    m, x, sl = setup()
    nsweeps, warmup, nbins, nmeas = 5, 1, 2, 2
    out = folder(tmp_path / "run")
    got = run_driver(tmp_path, "bins", m, x, sl, nsweeps, out, nbins=nbins, nmeas_per_bin=nmeas, warmup=warmup)

    orc = O.Oracle(L, L, LT, NG, NE)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init(1)
    rng = O.Ran2Stream(SEED)
    forward, meas = True, []
    for loop in range(1, nsweeps + 1):
        forward = not forward
        _, mo = orc.sweep(forward, rng.draws(2 * m.Nb * LT), measure=loop > warmup)
        if loop > warmup:
            meas.append(mo)
    ref = np.array(meas).real.reshape(nbins, nmeas, -1).mean(axis=1)
    tw = m.total_width
    ob = measio.read_bin_means(str(tmp_path / "run" / "data" / "out_core0.bin"), tw)
    tb = measio.read_bin_means(str(tmp_path / "run" / "temp" / "tempout_core0.bin"), tw)
    assert np.array_equal(ob, tb)
    assert rel(ob, ref) < 1e-10
    assert got["n_meas"] == nsweeps - warmup
    count = open(tmp_path / "run" / "temp" / "tempout_core_count.dat").read().splitlines()
    assert count == ["C   0   |    B   1", "C   0   |    B   2"]
    py = tmp_path / "py"
    measio.MeasurementSystem(m.Ncell, LT, nbins, nmeas, output_addr=str(py) + "/")
    assert (tmp_path / "run" / "data" / "reading_guide.nml").read_bytes() == \
        (py / "data" / "reading_guide.nml").read_bytes()


@pytest.mark.gpu
@needs_driver
def test_fortran_resume_matches_uninterrupted_run(tmp_path):
    m, x, sl = setup()
    kw = dict(global_update_loop=2, global_distance=0.05, nbins=2, nmeas_per_bin=2, warmup=2)
    a = run_driver(tmp_path, "a", m, x, sl, 6, folder(tmp_path / "a"), **kw)
    bdir = folder(tmp_path / "b")
    run_driver(tmp_path, "b1", m, x, sl, 4, bdir, checkpoint_loop=4, **kw)
    b = run_driver(tmp_path, "b2", m, x, sl, 6, bdir, resume=True, **kw)
    per = LT * 4 * m.npla
    assert np.array_equal(a["accepts"][4 * per:], b["accepts"][4 * per:])
    assert np.array_equal(a["fields"], b["fields"])
    assert rel(b["g"], a["g"]) < 1e-10
    for k in ("positive_accept", "negative_accept", "global_accept", "global_reject", "n_meas"):
        assert a[k] == b[k], k
    assert a["global_accept"] > 0
    assert rel(b["meas_sum"], a["meas_sum"]) < 1e-10
    tw = m.total_width
    for rp in ("data/out_core0.bin", "temp/tempout_core0.bin"):
        fa = measio.read_bin_means(str(tmp_path / "a" / rp), tw)
        fb = measio.read_bin_means(str(tmp_path / "b" / rp), tw)
        assert fa.shape == fb.shape == (2, tw)
        assert rel(fb, fa) < 1e-10, rp
    assert (tmp_path / "a" / "temp" / "tempout_core_count.dat").read_text() == \
        (tmp_path / "b" / "temp" / "tempout_core_count.dat").read_text()


def python_loops(m, ch, rng, gm, ms, loops, forward, accs):
    """Loops of src/Main_PQMC.f90:81-156 with global moves every second loop
    and bins after a warm-up of 2 loops, on a Python Chain."""
    for loop in loops:
        forward = not forward
        if not forward and (loop + 1) % 2 == 0:
            gm.run()
        sample = ch.sweep(forward, rng.draws(2 * m.Nb * LT), measure=loop > 2)
        if loop > 2:
            ms.begin_measure(forward, LT // 2, sample)
        accs[loop] = ch.accepts()
    return forward


@pytest.mark.gpu
def test_python_resume_matches_uninterrupted_run(tmp_path):
    m, x, sl = setup()

    def fresh():
        ch = host.Chain(m)
        ch.set_fields(x)
        ch.set_trial(sl)
        return ch

    ch = fresh()
    ch.init_g(1)
    rng = host.Ran2(SEED)
    gm = host.GlobalMoves(ch, rng, 0.05)
    ms = measio.MeasurementSystem(m.Ncell, LT, 2, 2)
    acc_a = {}
    python_loops(m, ch, rng, gm, ms, range(1, 7), True, acc_a)
    ref = {"g": ch.g_h, "fields": ch.get_fields(), "mean": ms.mean.copy(), "counters": ch.counters(),
           "global": (gm.accepted, gm.rejected), "lncw": ch.ln_cw}
    ch.close()

    ch = fresh()
    ch.init_g(1)
    rng = host.Ran2(SEED)
    gm = host.GlobalMoves(ch, rng, 0.05)
    ms = measio.MeasurementSystem(m.Ncell, LT, 2, 2)
    acc_b = {}
    fwd = python_loops(m, ch, rng, gm, ms, range(1, 5), True, acc_b)
    cnt = ch.counters()
    st = rng.state
    ck = fieldio.Checkpoint((L, L, LT, NG, NE, m.Nb), 4, fwd, (st.i, st.j, st.k, st.m, st.n),
                            cnt["positive_accept"], cnt["negative_accept"], gm.accepted, gm.rejected,
                            cnt["err_fast_max"], ch.get_fields(), m.total_width, bins=fieldio.bins_state(ms))
    path = str(tmp_path / "ck.bin")
    fieldio.save_checkpoint(path, ck)
    ch.close()

    c = fieldio.load_checkpoint(path)
    assert c.forward and c.loop == 4
    ch = host.Chain(m)
    ch.set_trial(sl)
    ch.set_fields(c.fields)
    ch.init_g(1)
    ch.set_counters(c.positive_accept, c.negative_accept, c.err_fast_max)
    rng = host.Ran2(0)
    for k, v in zip("ijkmn", c.ran2):
        setattr(rng.state, k, v)
    gm = host.GlobalMoves(ch, rng, 0.05)
    gm.accepted, gm.rejected = c.global_accept, c.global_reject
    ms = measio.MeasurementSystem(m.Ncell, LT, 2, 2)
    fieldio.restore_bins(ms, c.bins)
    python_loops(m, ch, rng, gm, ms, range(5, 7), c.forward, acc_b)
    for loop in (5, 6):
        assert np.array_equal(acc_a[loop], acc_b[loop]), loop
    assert np.array_equal(ch.get_fields(), ref["fields"])
    assert rel(ch.g_h, ref["g"]) < 1e-10
    assert rel(ms.mean, ref["mean"]) < 1e-10
    cb = ch.counters()
    for k in ("positive_accept", "negative_accept"):
        assert cb[k] == ref["counters"][k], k
    assert (gm.accepted, gm.rejected) == ref["global"] and gm.accepted > 0
    assert abs(ch.ln_cw.real - ref["lncw"].real) < 1e-9 * abs(ref["lncw"].real)
    ch.close()
