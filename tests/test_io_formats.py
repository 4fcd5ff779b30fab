"""Measurement bins, the reference's output files, the text field format,
post-processing statistics and checkpoints (host code, no device).

The Fortran drop-in (pqmc_iocheck) and the Python host write the same bytes
for the same data; bin means, post-processing statistics and the field text
format are pinned against numpy and against the formats' definitions:
src/mod_meas.f90:729-1017, src/outputnew.f90:94-232, src/Main_PQMC.f90:355-368.
"""
import os
import subprocess

import numpy as np
import pytest

from pqmc_allinone_amd import fieldio, measio
from pqmc_allinone_amd.host import FORTRAN_DRIVER, Ran2, observable_layout

IOCHECK = os.path.join(os.path.dirname(FORTRAN_DRIVER), "pqmc_iocheck")


def window_samples(rng, ntime, tw, nmeas_total, meas_number, partial=0):
    """(forward, time, sample) records: nmeas_total full windows around ntime/2
    with alternating directions, then `partial` slices of one more window."""
    recs = []
    forward = True
    for k in range(nmeas_total + (1 if partial else 0)):
        forward = not forward
        times = list(range(ntime // 2 - meas_number, ntime // 2 + meas_number + 1))
        if not forward:
            times = times[::-1]
        if k == nmeas_total:
            times = times[:partial]
        for t in times:
            s = rng.normal(size=tw) + 1j * rng.normal(size=tw)
            recs.append((forward, t, s))
    return recs


def python_files(d, L, Lt, nbins, nmeas, mnum, recs):
    ms = measio.MeasurementSystem(L * L, Lt, nbins, nmeas, mnum, output_addr=str(d) + "/")
    for fwd, t, s in recs:
        ms.begin_measure(fwd, t, s)
    ms.data_analyse()
    return ms


@pytest.mark.parametrize("mnum", [0, 1])
def test_bin_means_match_numpy(tmp_path, mnum):
    L, Lt, nbins, nmeas = 2, 12, 3, 4
    tw = 3 + 38 * L * L
    recs = window_samples(np.random.default_rng(5 + mnum), Lt, tw, nbins * nmeas, mnum)
    ms = python_files(tmp_path, L, Lt, nbins, nmeas, mnum, recs)
    w = 2 * mnum + 1
    per_meas = np.array([sum(r[2] for r in recs[i * w:(i + 1) * w]).real / w for i in range(nbins * nmeas)])
    ref = per_meas.reshape(nbins, nmeas, tw).mean(axis=1)
    tb = measio.read_bin_means(str(tmp_path / "temp" / "tempout_core0.bin"), tw)
    ob = measio.read_bin_means(str(tmp_path / "data" / "out_core0.bin"), tw)
    assert tb.shape == (nbins, tw) and ob.shape == (nbins, tw)
    assert np.abs(tb - ref).max() < 1e-14 * max(1.0, np.abs(ref).max())
    assert np.array_equal(tb, ob)
    lines = open(tmp_path / "temp" / "tempout_core_count.dat").read().splitlines()
    assert lines == [f"C   0   |    B   {b}" for b in range(1, nbins + 1)]
    with pytest.raises(measio.MeasurementError):
        ms.begin_measure(True, Lt // 2, np.zeros(tw))


def test_reading_guide_describes_the_layout(tmp_path):
    L = 4
    python_files(tmp_path, L, 20, 2, 1, 0, [])
    basics, obs = measio.read_reading_guide(str(tmp_path / "data" / "reading_guide.nml"))
    assert basics["total_width"] == 3 + 38 * L * L and basics["n_obs"] == 3 + 2 * 19 == len(obs)
    assert basics["n_bins"] == 2 and basics["n_meas_per_bin"] == 1 and basics["field_width"] == L * L
    assert basics["size_of_dp"] == 8 and basics["output_format"] == "f18.8"
    lay = observable_layout(L * L)
    for o, (name, lo, hi) in zip(obs, lay):
        assert (o["name"], o["offset_lo"], o["offset_hi"], o["width"]) == (name, lo + 1, hi, hi - lo)
        assert o["kind"] == (1 if hi - lo == 1 else 2)
    assert obs[-1]["offset_hi"] == basics["total_width"]


def test_postpro_statistics_match_numpy(tmp_path):
    """4 cores in 2 MPI blocks: mean over cores and bins of a block, the error
    of independent bins sqrt(var/N) and the error of core means sqrt(var_c/C)."""
    L, Lt, nbins, nmeas, ncores, nblock = 2, 8, 5, 3, 4, 2
    tw = 3 + 38 * L * L
    rng = np.random.default_rng(11)
    for myid in range(ncores):
        ms = measio.MeasurementSystem(L * L, Lt, nbins, nmeas, 0, output_addr=str(tmp_path) + "/", myid=myid,
                                      n_cores=ncores, mpi_nblock=nblock)
        for fwd, t, s in window_samples(rng, Lt, tw, nbins * nmeas, 0):
            ms.begin_measure(fwd, t, s + 3.0 * myid)
        ms.data_analyse()
    res = measio.postpro(str(tmp_path / "data"))
    allb = np.stack([measio.read_bin_means(str(tmp_path / "data" / f"out_core{i}.bin"), tw) for i in range(ncores)])
    one = ncores // nblock
    for name, lo, hi in observable_layout(L * L):
        r = res[name]
        for ib in range(nblock):
            d = allb[ib * one:(ib + 1) * one, :, lo:hi]           # (cores, bins, width)
            flat = d.reshape(-1, hi - lo)
            assert np.allclose(r["out_mean"][ib], flat.mean(axis=0), rtol=1e-13, atol=1e-13)
            assert np.allclose(r["out_err"][ib], flat.std(axis=0, ddof=1) / np.sqrt(flat.shape[0]), rtol=1e-12)
            cm = d.mean(axis=1)
            assert np.allclose(r["large_err"][ib], cm.std(axis=0, ddof=1) / np.sqrt(one), rtol=1e-12)
    target = tmp_path / "result.txt"
    target.write_text("run parameters\n#--observables--#\nold table that is replaced\n")
    measio.write_results(str(target), res, block=2)
    lines = target.read_text().splitlines()
    assert lines[:2] == ["run parameters", "#--observables--#"] and "old table" not in target.read_text()
    rows = [ln for ln in lines if ln[:15].strip() == "Sy_Sy_k"]
    assert len(rows) == L * L
    vals = [float(rows[0][21 + 18 * i:21 + 18 * (i + 1)]) for i in range(3)]
    assert np.allclose(vals, [res["Sy_Sy_k"][k][1, 0] for k in ("out_mean", "out_err", "large_err")], atol=1e-8)


def test_ph_field_text_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(7, 32)) * 3
    x[0, :3] = [0.0, -1e-9, 123456.25]
    p = tmp_path / "ph_field_core1.dat"
    fieldio.write_ph_field(str(p), x)
    lines = p.read_text().splitlines()
    assert len(lines) == 7 and all(len(ln) == 18 * 32 for ln in lines)
    assert lines[0][:54] == "          0.000000         -0.000000     123456.250000"
    y = fieldio.read_ph_field(str(p), 7, 32)
    assert np.abs(y - x).max() <= 5e-7 + 1e-12


def test_checkpoint_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    tw = 3 + 38 * 4
    ms = measio.MeasurementSystem(4, 10, 2, 3, 0)
    for fwd, t, s in window_samples(rng, 10, tw, 4, 0):
        ms.begin_measure(fwd, t, s)
    c = fieldio.Checkpoint((2, 2, 10, 5, 4, 8), 4, True, (1, -2, 3, 4, 5), 10, 20, 1, 2, 3.5e-9,
                           rng.normal(size=(10, 8)), tw, 4, rng.normal(size=tw) + 0j, fieldio.bins_state(ms))
    p = str(tmp_path / "c.bin")
    fieldio.save_checkpoint(p, c)
    d = fieldio.load_checkpoint(p)
    assert (d.dims, d.loop, d.forward, d.ran2) == (c.dims, c.loop, c.forward, c.ran2)
    assert (d.positive_accept, d.negative_accept, d.global_accept, d.global_reject) == (10, 20, 1, 2)
    assert d.err_fast_max == c.err_fast_max and np.array_equal(d.fields, c.fields)
    assert np.array_equal(d.meas_sum, c.meas_sum) and d.n_meas == 4
    ms2 = measio.MeasurementSystem(4, 10, 2, 3, 0)
    fieldio.restore_bins(ms2, d.bins)
    assert (ms2.cur_bin, ms2.cur_meas) == (2, 2)
    assert np.array_equal(ms2.mean, ms.mean) and np.array_equal(ms2.bin_sum, ms.bin_sum)


@pytest.mark.skipif(not os.path.exists(IOCHECK), reason="Fortran host not built (amdflang missing)")
@pytest.mark.parametrize("mnum", [0, 1])
def test_fortran_files_equal_python_files(tmp_path, mnum):
    """Bins, reading guide, count file, field text and checkpoint: the
    Fortran drop-in's bytes equal the Python host's."""
    L1, L2, Lt, ng, ne, nbins, nmeas, seed = 4, 4, 20, 5, 16, 3, 2, 77
    ncell, Nb = L1 * L2, 2 * L1 * L2
    tw = 3 + 38 * ncell
    rng = np.random.default_rng(9 + mnum)
    x = rng.normal(size=(Lt, Nb))
    recs = window_samples(rng, Lt, tw, 5, mnum, partial=mnum + 1)
    fdir, pdir = tmp_path / "f", tmp_path / "p"
    for d in (fdir, pdir):
        (d / "temp").mkdir(parents=True)
        (d / "data").mkdir()
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        np.array([L1, L2, Lt, ng, ne, nbins, nmeas, mnum, len(recs), seed], np.int32).tofile(f)
        x.tofile(f)
        for fwd, t, s in recs:
            np.array([int(fwd), t], np.int32).tofile(f)
            s.astype(np.complex128).tofile(f)
    r = subprocess.run([IOCHECK, str(fin), str(fdir) + "/"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr

    ms = python_files(pdir, L1, Lt, nbins, nmeas, mnum, recs)
    for rel in ("temp/tempout_core0.bin", "temp/tempout_core_count.dat", "data/out_core0.bin",
                "data/reading_guide.nml"):
        assert (fdir / rel).read_bytes() == (pdir / rel).read_bytes(), rel

    fieldio.write_ph_field(str(pdir / "ph_field_core1.dat"), x)
    assert (fdir / "ph_field_core1.dat").read_bytes() == (pdir / "ph_field_core1.dat").read_bytes()
    back = np.fromfile(fdir / "ph_field_read.bin", np.float64).reshape(Lt, Nb)
    assert np.array_equal(back, fieldio.read_ph_field(str(pdir / "ph_field_core1.dat"), Lt, Nb))

    rg = Ran2(seed)
    rg.draws(7)
    st = rg.state
    c = fieldio.Checkpoint((L1, L2, Lt, ng, ne, Nb), 6, True, (st.i, st.j, st.k, st.m, st.n), 11, 22, 3, 4,
                           1.5e-7, x, tw, len(recs), sum(s for _, _, s in recs), fieldio.bins_state(ms))
    fieldio.save_checkpoint(str(pdir / "ckpt.bin"), c)
    assert (fdir / "ckpt.bin").read_bytes() == (pdir / "ckpt.bin").read_bytes()
