"""The Fortran host (pqmc_allinone_amd/fortran: iso_c_binding module, drop-in
update_phonon_field / left_evolve / right_evolve / get_g_scratch /
take_measurement, and a driver running the reference's sweep loop one
subroutine call at a time) against the oracle on the same inputs and the
same ran2 stream: identical accept/reject sequences and fields, G and the
summed measurement to 1e-10 relative."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from pqmc_allinone_amd import host
from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

TOL = 1e-10


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.skipif(not os.path.exists(host.FORTRAN_DRIVER), reason="Fortran driver not built (amdflang missing)")
def test_driver_input_layout_round_trip(tmp_path):
    """The input writer produces exactly the byte count the driver reads."""
    m = build_model(ModelParams(L1=4, L2=4, Lt=20, ngroup=10, ne=16))
    x, sl = synthetic_inputs(m, 1234)
    path = tmp_path / "in.bin"
    host.write_driver_input(str(path), m, x, sl, nsweeps=2, seed=1234)
    npla, Ns, Nb, Lt, ne = m.npla, m.Ns, m.Nb, m.p.Lt, m.p.ne
    expect = (8 * 4 + 4 * npla * 4 * 4 + 2 * 4 * npla * 4 + 4 * npla * 16 * 16 + (4 + 4 + 16 + 16 + 2) * 16
              + 7 * 8 + 2 * 4 + Nb * Lt * 8 + Ns * ne * 16)
    assert path.stat().st_size == expect


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(host.FORTRAN_DRIVER), reason="Fortran driver not built (amdflang missing)")
@pytest.mark.parametrize("nb", [0, 3])
def test_fortran_driver_matches_oracle(tmp_path, nb):
    L, Lt, ng, ne, seed, nsweeps = 4, 40, 10, 16, 1234, 2
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne))
    x, sl = synthetic_inputs(m, seed)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    host.write_driver_input(str(fin), m, x, sl, nsweeps=nsweeps, seed=seed, block_bonds=nb)
    r = subprocess.run([host.FORTRAN_DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = host.read_driver_output(str(fout), m, nsweeps)

    orc = O.Oracle(L, L, Lt, ng, ne)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init(1)
    rng = O.Ran2Stream(seed)
    forward = True
    accs, meas = [], np.zeros(m.total_width, np.complex128)
    for _ in range(nsweeps):
        forward = not forward
        a, mo = orc.sweep(forward, rng.draws(2 * m.Nb * Lt), measure=True)
        accs.append(a)
        meas += mo
    acc_o = np.concatenate(accs)
    assert np.array_equal(got["accepts"], acc_o), f"first mismatch at {np.argmax(got['accepts'] != acc_o)}"
    assert np.array_equal(got["fields"], orc.get_fields())
    assert rel(got["g"], orc.g) < TOL
    assert got["n_meas"] == nsweeps
    assert rel(got["meas_sum"], meas) < TOL
    pa, na, _ = orc.counters()
    assert (got["positive_accept"], got["negative_accept"]) == (pa, na)
