"""The Fortran host (pqmc_allinone_amd/fortran: iso_c_binding module, drop-in
update_phonon_field / left_evolve / right_evolve / get_g_scratch /
take_measurement / update_global, and a driver running the reference's sweep
loop one subroutine call at a time) against the oracle on the same inputs and
the same ran2 stream: identical accept/reject sequences (local and global) and
fields, G and the summed measurement to 1e-10 relative."""
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
              + 7 * 8 + 2 * 4 + Nb * Lt * 8 + Ns * ne * 16 + 4 + 8 + 8 * 4)
    assert path.stat().st_size == expect


PROPOSE = os.path.join(os.path.dirname(host.FORTRAN_DRIVER), "pqmc_propose")


@pytest.mark.skipif(not os.path.exists(PROPOSE), reason="Fortran host not built (amdflang missing)")
@pytest.mark.parametrize("kind", [1, 2, 3])
def test_fortran_global_proposals_match_oracle(tmp_path, kind):
    """The drop-in's generate_newfield_global (host code, no device) against
    the oracle's: proposed fields and -delt*dE bit for bit."""
    for L, Lt, biased, dist, seed in [(4, 20, 0.0, 1.0, 3), (6, 300, 0.3, 0.7, 8), (4, 7, 0.1, 2.0, 5)]:
        p = ModelParams(L1=L, L2=L, Lt=Lt, ngroup=5, ne=L * L, biased=biased)
        m = build_model(p)
        x, sl = synthetic_inputs(m, seed)
        fin, fout = tmp_path / "p_in.bin", tmp_path / "p_out.bin"
        with open(fin, "wb") as f:
            np.array([L, L, Lt, kind, seed], np.int32).tofile(f)
            np.array([m.M, m.D, biased, p.dtau, dist * m.char_length], np.float64).tofile(f)
            np.ascontiguousarray(x, np.float64).tofile(f)
        r = subprocess.run([PROPOSE, str(fin), str(fout)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        out = np.fromfile(fout, np.float64)
        xf, phf = out[:-1].reshape(Lt, m.Nb), out[-1]
        orc = O.Oracle(L, L, Lt, 5, L * L, biased=biased)
        orc.set_fields(x)
        orc.set_trial(sl)
        xo, pho = orc.propose_global(kind - 1, O.Ran2Stream(seed), dist)
        assert np.array_equal(xf, xo), (L, Lt, kind)
        assert phf == pho, (L, Lt, kind)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(host.FORTRAN_DRIVER), reason="Fortran driver not built (amdflang missing)")
@pytest.mark.parametrize("nb,glob,ft", [(0, False, False), (3, False, False), (0, True, False), (0, False, True),
                                        (0, True, True)])
def test_fortran_driver_matches_oracle(tmp_path, nb, glob, ft):
    """glob: global_update_loop = 2 with a small move distance, so that global
    moves are accepted and init_g(1) runs inside the loop.  ft: proj =
    .false., the finite-temperature Green's function."""
    L, Lt, ng, ne, seed = 4, 40, 10, 16, 1234
    nsweeps = 4 if glob or ft else 2
    dist = 0.05
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne))
    x, sl = synthetic_inputs(m, seed)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    host.write_driver_input(str(fin), m, x, sl, nsweeps=nsweeps, seed=seed, block_bonds=nb,
                            global_update_loop=2 if glob else 0, global_distance=dist, proj=not ft)
    r = subprocess.run([host.FORTRAN_DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = host.read_driver_output(str(fout), m, nsweeps)

    orc = O.Oracle(L, L, Lt, ng, ne, ft=ft)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init(1)
    rng = O.Ran2Stream(seed)
    forward = True
    accs, meas = [], np.zeros(m.total_width, np.complex128)
    for loop in range(1, nsweeps + 1):
        forward = not forward
        if glob and not forward and (loop + 1) % 2 == 0:
            flags = [orc.update_global(k, rng, dist) for k in (O.GU_KSPACE, O.GU_ROTATE, O.GU_KSPACE_TIME)]
            if any(flags):
                orc.init(1)
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
    ga, gr = orc.global_counters()
    assert (got["global_accept"], got["global_reject"]) == (ga, gr)
    if glob:
        assert ga > 0 and ga + gr == 6


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(host.FORTRAN_DRIVER), reason="Fortran driver not built (amdflang missing)")
def test_fortran_driver_gradient_matches_python_host(tmp_path):
    """The driver's pqmc_gradient call (iso_c_binding, complex(c_double_complex)
    array) returns what the Python host gets for the same final fields."""
    from pqmc_allinone_amd.host import Chain
    L, Lt, ng, ne, seed = 4, 40, 10, 16, 1234
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne))
    x, sl = synthetic_inputs(m, seed)
    fin, fout, fgrad = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "grad.bin"
    host.write_driver_input(str(fin), m, x, sl, nsweeps=2, seed=seed)
    r = subprocess.run([host.FORTRAN_DRIVER, str(fin), str(fout), str(tmp_path) + "/", str(fgrad)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = host.read_driver_output(str(fout), m, 2)
    g_f = np.fromfile(fgrad, np.complex128).reshape(Lt, m.Nb)
    ch = Chain(m)
    ch.set_fields(got["fields"])
    ch.set_trial(sl)
    ch.init_g(1)
    g_p = ch.gradient()
    ch.close()
    assert np.isfinite(g_f).all() and np.abs(g_f).max() > 0
    assert np.abs(g_f - g_p).max() <= 1e-12 * np.abs(g_p).max()

