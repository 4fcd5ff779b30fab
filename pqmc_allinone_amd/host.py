"""Host-side mirror of the reference's sweep interfaces over the C ABI.

The names and argument meanings follow the Fortran modules the sweep loop of
src/Main_PQMC.f90:81-156 calls:

===========================  ============================================
reference                    here
===========================  ============================================
init_MC / init_g(1)          ``Chain.init_g(start_time)``
update_phonon_field(ppf,t)   ``Chain.update_phonon_field(pf_id, time, u)``
left_evolve(t,g_h,..,.true.) ``Chain.left_evolve(time, pf_id)``
right_evolve(t,g_h,..)       ``Chain.right_evolve(time, pf_id)``
get_g_scratch(t,fwd,loop)    ``Chain.get_g_scratch(time, forward)``
Meas_sys%begin_measure       ``measio.MeasurementSystem.begin_measure``
update_global(kind)          ``GlobalMoves.update(kind, x, ln_cw)``
loop body (one sweep)        ``Chain.sweep(forward, draws, measure)``
ran2_s / rands / ran_sysm    ``Ran2`` (src/mod_ranstate.f90:171-236)
proj = .false. (_FT family)  ``Chain(model, finite_temperature=True)``
cal_new_weight_T0 / _FT      ``Chain.global_weight(x)``
(adjoint, no counterpart)    ``Chain.gradient()``, ``HybridMC``
one chain per MPI rank       ``sweep_batch`` / ``sweep_many`` (chains together)
===========================  ============================================

All state lives on the GPU; ``g_h``, fields and B-string checkpoints are
copied to the host only when read.  Errors raise ``PQMCError`` where the
reference would ``stop``.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _lib
from .model import Model

# observable registry of input/model/EPSOCZ.nml:262-283 (field kinds get a _k twin)
SCALARS = ("BF_KE", "BF_PE", "BF_X")
FIELDS = ("BFx_BFx", "BFx_BFy", "BFy_BFy", "BFxy_BFxy", "Jxx_Jxx", "Jxy_Jxy", "Jyx_Jyx", "Jyy_Jyy",
          "Jxx_Jyx", "Jxy_Jyy", "Jxx_Jyy", "Jxy_Jyx", "cJxy_cJxy", "cJyx_cJyx", "den_den", "SC_SC",
          "Sz_Sz", "Sx_Sx", "Sy_Sy")


def observable_layout(ncell: int):
    """(name, lo, hi) offsets as ms_finalize_layout (src/mod_meas.f90:691-707), 0-based."""
    out, cur = [], 0
    for n in SCALARS:
        out.append((n, cur, cur + 1))
        cur += 1
    for n in FIELDS:
        out.append((n, cur, cur + ncell))
        cur += ncell
        out.append((n + "_k", cur, cur + ncell))
        cur += ncell
    return out


def expkv_epsoc(dtau: float, kcoe, vcoe, k12, v12, fam, x, bcph=None, cross=None, inv: bool = False):
    """The EPSOC model's 2x2 bond exponentials (get_expKV_epsoc,
    src/mod_phonon_field.f90:572-604) on the device: kcoe, vcoe, k12, v12
    per family (complex); fam, x (and the boundary phase bcph where cross is
    set) per (bond, slice) pair.  Returns (n, 2, 2) complex."""
    L = _lib.lib()
    kc, vc, kk, vv = (np.ascontiguousarray(np.asarray(a, np.complex128)) for a in (kcoe, vcoe, k12, v12))
    fam = np.ascontiguousarray(fam, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    n = fam.size
    if x.size != n:
        raise ValueError("fam and x need one entry per pair")
    cr = ph = None
    if cross is not None:
        cr = np.ascontiguousarray(cross, np.int32)
        ph = np.ascontiguousarray(np.asarray(bcph, np.complex128))
        if cr.size != n or ph.size != n:
            raise ValueError("cross and bcph need one entry per pair")
    out = np.zeros((n, 2, 2), np.complex128)
    work = np.zeros(4 * n, np.complex128)
    _lib.check(L.pqmc_expkv_epsoc(float(dtau), int(kc.size), _lib.ptr(kc), _lib.ptr(vc), _lib.ptr(kk), _lib.ptr(vv),
                                  int(n), _lib.ptr(fam), _lib.ptr(x), _lib.ptr(cr) if cr is not None else None,
                                  _lib.ptr(ph) if ph is not None else None, int(bool(inv)), _lib.ptr(work)))
    out[:] = work.reshape(n, 2, 2).transpose(0, 2, 1)  # column-major 2x2 per pair
    return out


class Ran2:
    """ran2 of src/mod_ranstate.f90:171-191, seeded as init_rng_by_seed (:267-283)."""

    def __init__(self, seed: int):
        self.state = _lib.Ran2()
        _lib.lib().pqmc_ran2_seed(C.byref(self.state), int(seed))

    def draws(self, n: int) -> np.ndarray:
        out = np.empty(int(n), np.float32)
        _lib.lib().pqmc_ran2_fill(C.byref(self.state), int(n), _lib.ptr(out))
        return out

    def rands(self) -> float:
        return float(self.draws(1)[0])

    def ran_sysm(self) -> float:
        r = self.draws(1)[0]
        return float(np.float32(r * np.float32(2) - np.float32(1)))


# REAL(SP) PI of src/mod_nrtype.f90:42
SP_PI = np.float32(np.pi)
# update_global kinds in the order the sweep loop calls them (src/Main_PQMC.f90:88-94)
GLOBAL_KINDS = ("kspace", "rotate", "kspace_time")


def _irands(rng: "Ran2", n: int) -> int:
    """irands (src/mod_ranstate.f90:221-226): int(r * n) in REAL(SP)."""
    return int(np.float32(rng.draws(1)[0]) * np.float32(n))


def _seqsum(start: float, terms: np.ndarray) -> float:
    """start + t1 + t2 + ... added left to right, as the reference's loops do."""
    return float(np.add.accumulate(np.concatenate(([start], np.ravel(terms))))[-1])


def propose_global(model: Model, x: np.ndarray, kind: str, rng: "Ran2", distance: float = 1.0):
    """generate_newfield_global (src/mod_update.f90:292-439) for ``kind``.

    Returns the proposed fields (Lt x Nb, row t = boson_field(:, t)) and
    -delt*dE, the bosonic part of the weight change.  Draws come from ``rng``
    in the reference's order.  Arithmetic keeps the reference's REAL(SP)
    steps (PI, ran_sysm, irands) and its summation order.  For kspace_time
    the window is cut at ntime, where the reference would index past it
    (start_time + M/dtau^2 > ntime).
    """
    p = model.p
    Lt, Nb = p.Lt, model.Nb
    xo = np.ascontiguousarray(x, np.float64).reshape(Lt, Nb)
    xn = xo.copy()
    D, M, dt = model.D, model.M, p.dtau
    if kind in ("kspace", "kspace_time"):
        incell = [_irands(rng, 2) * 2.0 - 1.0 for _ in range(2)]
        # generate_newfield_space (:441-467) with K_vec = 2*PI*[0.5d0, 0.5d0], phi = 0
        gud = distance * model.char_length
        dist = 2 * gud / model.Ncell
        amplitude = dist * float(rng.draws(1)[0])
        K = float(np.float32(2) * SP_PI) * 0.5
        bj = np.empty(Nb)
        for c in range(model.Ncell):
            cx, cy = c % p.L1, c // p.L1  # rpos = dpos (0-based) for lat_vec = identity
            cs = math.cos((K * cx + K * cy) + 0.0)
            for i in range(2):
                bj[2 * c + i] = incell[i] * amplitude * cs
        t0, t1 = 1, Lt
        if kind == "kspace_time":
            t0 = _irands(rng, Lt // 2) + 1
            t1 = min(max(t0 + int(1.0 * M / (dt * dt)), Lt), Lt)
        xn[t0 - 1:t1] = xo[t0 - 1:t1] + bj
        a, b = xn[t0 - 1:t1], xo[t0 - 1:t1]
        terms = np.stack([0.5 * D * (a * a - b * b), -p.biased * (a - b)], axis=-1)
        de = _seqsum(0.0, terms)
        if kind == "kspace_time":
            kin = 0.5 * (M / (dt * dt))
            tb, ta = (t0 - 2) % Lt, t1 % Lt  # modi(t0 - 1, ntime), modi(t1 + 1, ntime), 0-based rows
            e0 = kin * ((xn[t0 - 1] - xn[tb]) ** 2 - (xo[t0 - 1] - xo[tb]) ** 2)
            e1 = kin * ((xn[t1 - 1] - xn[ta]) ** 2 - (xo[t1 - 1] - xo[ta]) ** 2)
            de = _seqsum(de, np.stack([e0, e1], axis=-1))
        return xn, -de * dt
    if kind == "rotate":
        ang = float(np.float32(rng.ran_sysm())) * 0.1 * float(SP_PI)
        r11, r12, r21, r22 = math.cos(ang), -math.sin(ang), math.sin(ang), math.cos(ang)
        if rng.ran_sysm() < 0.0:
            r12, r22 = -r12, -r22
        a, b = xo[:, 0::2], xo[:, 1::2]
        xn[:, 0::2] = r11 * a + r12 * b
        xn[:, 1::2] = r21 * a + r22 * b
        return xn, 0.0
    raise ValueError(f"unknown global update kind {kind!r}")


class Chain:
    """One projector-DQMC Markov chain resident on the GPU."""

    def __init__(self, model: Model, block_bonds: int = 0, helper_bonds: int = 0,
                 finite_temperature: bool = False):
        """finite_temperature: G = (1 + B(beta,0))^-1 (proj = .false. of
        src/input.f90) instead of the projector G; needs no trial matrix."""
        self.model = model
        p = model.p
        self._L = _lib.lib()
        self.finite_temperature = bool(finite_temperature)
        cfg = _lib.Config(p.L1, p.L2, p.Lt, p.ngroup, p.ne, int(block_bonds), int(helper_bonds),
                          int(self.finite_temperature))
        h = C.c_void_p()
        _lib.check(self._L.pqmc_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = [np.ascontiguousarray(model.sites, np.int32), np.ascontiguousarray(model.bfi, np.int32),
                      np.ascontiguousarray(model.cross, np.int32),
                      np.ascontiguousarray(np.transpose(model.bcph, (0, 1, 3, 2)), np.complex128),
                      np.ascontiguousarray(model.kcoe), np.ascontiguousarray(model.vcoe),
                      np.ascontiguousarray(np.transpose(model.k12, (0, 2, 1))),
                      np.ascontiguousarray(np.transpose(model.v12, (0, 2, 1))),
                      np.ascontiguousarray(model.bc)]
        k = self._keep
        mt = _lib.ModelTables(*[_lib.ptr(a) for a in k], p.dtau, model.M, model.D, p.biased,
                              model.jump_distance, p.hop, p.lam, int(p.ncopy), int(bool(p.tr_weight)))
        _lib.check(self._L.pqmc_set_model(self.h, C.byref(mt)))
        dims = np.zeros(6, np.int32)
        _lib.check(self._L.pqmc_get_dims(self.h, _lib.ptr(dims)))
        self.Ns, self.Ncell, self.Nb, self.npla, self.nblock, self.total_width = (int(v) for v in dims)

    def close(self):
        if getattr(self, "h", None):
            self._L.pqmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- inputs -------------------------------------------------------
    def set_fields(self, x: np.ndarray):
        x = np.ascontiguousarray(x, np.float64)
        if x.size != self.Nb * self.model.p.Lt:
            raise ValueError("fields must hold Nb*Lt values")
        _lib.check(self._L.pqmc_set_fields(self.h, _lib.ptr(x)))

    def get_fields(self) -> np.ndarray:
        x = np.empty((self.model.p.Lt, self.Nb), np.float64)
        _lib.check(self._L.pqmc_get_fields(self.h, _lib.ptr(x)))
        return x

    def set_trial(self, slater: np.ndarray):
        s = np.asfortranarray(slater, np.complex128)
        if s.shape != (self.Ns, self.model.p.ne):
            raise ValueError("slater must be Ns x ne")
        _lib.check(self._L.pqmc_set_trial(self.h, _lib.ptr(s)))

    # ---- reference interfaces ----------------------------------------
    def init_g(self, start_time: int = 1):
        _lib.check(self._L.pqmc_init_g(self.h, int(start_time)))

    def update_phonon_field(self, pf_id: int, time: int, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, np.float32)
        if u.size != 2 * self.npla:
            raise ValueError("need 2*npla draws")
        acc = np.zeros(self.npla, np.uint8)
        _lib.check(self._L.pqmc_update_family(self.h, int(pf_id), int(time), _lib.ptr(u), _lib.ptr(acc)))
        return acc

    def left_evolve(self, time: int, pf_id: int):
        _lib.check(self._L.pqmc_evolve_g(self.h, int(time), int(pf_id), 1))

    def right_evolve(self, time: int, pf_id: int):
        _lib.check(self._L.pqmc_evolve_g(self.h, int(time), int(pf_id), 0))

    def evolve_matrix(self, time: int, pf_id: int, mat: np.ndarray, left: bool, two_way: bool) -> np.ndarray:
        """left_evolve(time, mat(Ns,n)) / right_evolve(time, mat(n,Ns)) on a host matrix."""
        m = np.asfortranarray(mat, np.complex128).copy(order="F")
        n = m.shape[1] if left else m.shape[0]
        _lib.check(self._L.pqmc_evolve_matrix(self.h, int(time), int(pf_id), int(bool(left)), int(bool(two_way)),
                                              _lib.ptr(m), int(n)))
        return m

    def get_g_scratch(self, time: int, forward: bool) -> float:
        err = C.c_double(0.0)
        _lib.check(self._L.pqmc_g_scratch(self.h, int(time), int(bool(forward)), C.addressof(err)))
        return err.value

    def take_measurement(self, time: int) -> np.ndarray:
        out = np.empty(self.total_width, np.complex128)
        _lib.check(self._L.pqmc_measure(self.h, int(time), _lib.ptr(out)))
        return out

    def sweep(self, forward: bool, u: np.ndarray, measure: bool = False, ntc: int = 0):
        """One iteration of the reference's `do loop` body; returns the measurement or None."""
        u = np.ascontiguousarray(u, np.float32)
        if u.size != 2 * self.Nb * self.model.p.Lt:
            raise ValueError("need 2*Nb*Lt draws")
        out = np.empty(self.total_width, np.complex128) if measure else None
        _lib.check(self._L.pqmc_sweep(self.h, int(bool(forward)), _lib.ptr(u), int(ntc), int(bool(measure)),
                                      _lib.ptr(out) if out is not None else None))
        return out

    # ---- state -------------------------------------------------------
    @property
    def g_h(self) -> np.ndarray:
        g = np.empty((self.Ns, self.Ns), np.complex128, order="F")
        _lib.check(self._L.pqmc_get_g(self.h, _lib.ptr(g)))
        return g

    @g_h.setter
    def g_h(self, g: np.ndarray):
        g = np.asfortranarray(g, np.complex128)
        _lib.check(self._L.pqmc_set_g(self.h, _lib.ptr(g)))

    @property
    def ln_cw(self) -> complex:
        v = np.zeros(2)
        _lib.check(self._L.pqmc_get_lncw(self.h, _lib.ptr(v)))
        return complex(v[0], v[1])

    def global_weight(self, x: np.ndarray) -> complex:
        """cal_new_weight_T0 (src/mod_evolution.f90:359-408) for the fields x."""
        x = np.ascontiguousarray(x, np.float64)
        if x.size != self.Nb * self.model.p.Lt:
            raise ValueError("fields must hold Nb*Lt values")
        v = np.zeros(2)
        _lib.check(self._L.pqmc_global_weight(self.h, _lib.ptr(x), _lib.ptr(v)))
        return complex(v[0], v[1])

    def gradient(self, noalloc: bool = False) -> np.ndarray:
        """d ln_cw / dx for every field, complex, shape (Lt, Nb): a backward
        pass with no updates (pqmc_gradient).  Call where a backward sweep
        starts (after init_g(1) or a forward sweep); the chain ends where a
        backward sweep that accepts nothing leaves it.  noalloc: the
        library's fixed device array as the accumulator
        (pqmc_gradient_noalloc), no device allocation."""
        g = np.empty((self.model.p.Lt, self.Nb), np.complex128)
        fn = self._L.pqmc_gradient_noalloc if noalloc else self._L.pqmc_gradient
        _lib.check(fn(self.h, _lib.ptr(g)))
        return g

    def q_string(self, block: int) -> np.ndarray:
        q = np.empty((self.Ns, self.model.p.ne), np.complex128, order="F")
        _lib.check(self._L.pqmc_get_qstring(self.h, int(block), _lib.ptr(q)))
        return q

    def d_string(self) -> np.ndarray:
        d = np.empty((self.model.p.ne, self.nblock), np.complex128, order="F")
        _lib.check(self._L.pqmc_get_dstring(self.h, _lib.ptr(d)))
        return d

    def accepts(self, n: int | None = None) -> np.ndarray:
        n = self.Nb * self.model.p.Lt if n is None else int(n)
        a = np.empty(n, np.uint8)
        _lib.check(self._L.pqmc_get_accepts(self.h, _lib.ptr(a), n))
        return a

    def counters(self):
        o = np.zeros(3, np.int64)
        efm = C.c_double(0.0)
        _lib.check(self._L.pqmc_get_counters(self.h, _lib.ptr(o), C.addressof(efm)))
        return {"positive_accept": int(o[0]), "negative_accept": int(o[1]), "sync_timeouts": int(o[2]),
                "err_fast_max": efm.value}

    def drift(self):
        """Largest G and Re ln_cw drift over the stabilizations of the latest
        sweep, gradient pass or get_g_scratch (max|G_fast - G_scratch|,
        |Re(ln_cw - ln_cw_fast)|, src/mod_evolution.f90:179-193)."""
        o = np.zeros(2)
        _lib.check(self._L.pqmc_get_drift(self.h, _lib.ptr(o)))
        return float(o[0]), float(o[1])

    PRECISIONS = {"fp64": 0, "fp32x3": 1, "bf16x3": 2}

    def set_precision(self, mode: str):
        """B-string products between QR checkpoints in "fp64" (default),
        "fp32x3" (FP32 MFMA, three-term split operands) or "bf16x3" (BF16
        MFMA); everything else stays FP64.  Takes effect from the next
        init_g / get_g_scratch / sweep."""
        _lib.check(self._L.pqmc_set_precision(self.h, self.PRECISIONS[mode]))

    QR_MODES = {"flow": 0, "panels": 1}

    def set_qr_mode(self, mode: str):
        """Householder QR of the stabilizations as one launch ("flow",
        default: a workgroup per panel, panels handed on by flags) or panel
        by panel ("panels": a panel launch and a block-reflector launch per
        16 columns).  Same conventions and outputs; the results differ only
        by rounding."""
        _lib.check(self._L.pqmc_set_qr_mode(self.h, self.QR_MODES[mode]))

    def drift_history(self) -> np.ndarray:
        """(n, 2): G and Re ln_cw drift of each stabilization of the latest
        sweep, gradient pass or get_g_scratch, in order."""
        cap = self.model.p.Lt // self.model.p.ngroup + 1
        o = np.zeros((cap, 2))
        n = C.c_int32(0)
        _lib.check(self._L.pqmc_get_drift_history(self.h, _lib.ptr(o), cap, C.addressof(n)))
        return o[: min(n.value, cap)].copy()

    def set_counters(self, positive_accept: int, negative_accept: int, err_fast_max: float):
        """Restore the accept counters and err_fast_max (resume from a checkpoint)."""
        v = np.array([positive_accept, negative_accept], np.int64)
        _lib.check(self._L.pqmc_set_counters(self.h, _lib.ptr(v), float(err_fast_max)))

    def synchronize(self):
        _lib.check(self._L.pqmc_synchronize(self.h))

    def timing(self, kind: int, enable: bool):
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        _lib.check(self._L.pqmc_timing(self.h, int(bool(enable)), int(kind), C.addressof(ms), C.addressof(n)))
        return ms.value, n.value


# ---- Fortran host driver (fortran/pqmc_driver) file formats ---------------

FORTRAN_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "pqmc_driver")


def write_driver_input(path: str, model: Model, fields: np.ndarray, slater: np.ndarray, nsweeps: int, seed: int,
                       block_bonds: int = 0, global_update_loop: int = 0, global_distance: float = 1.0,
                       nbins: int = 0, nmeas_per_bin: int = 0, meas_number: int = 0, warmup: int = 0,
                       meas_interval: int = 1, checkpoint_loop: int = 0, resume: bool = False,
                       proj: bool = True):
    """Input file of fortran/pqmc_driver (layout in its header comment).
    global_distance is the input's global_update_distance, before its scaling
    by char_length (src/input.f90:184).  proj = False runs the finite-
    temperature Green's function (the input's proj flag)."""
    p = model.p
    with open(path, "wb") as f:
        np.array([p.L1, p.L2, p.Lt, p.ngroup, p.ne, block_bonds, nsweeps, seed], np.int32).tofile(f)
        np.ascontiguousarray(model.sites, np.int32).tofile(f)
        np.ascontiguousarray(model.bfi, np.int32).tofile(f)
        np.ascontiguousarray(model.cross, np.int32).tofile(f)
        np.ascontiguousarray(np.transpose(model.bcph, (0, 1, 3, 2)), np.complex128).tofile(f)
        np.ascontiguousarray(model.kcoe, np.complex128).tofile(f)
        np.ascontiguousarray(model.vcoe, np.complex128).tofile(f)
        np.ascontiguousarray(np.transpose(model.k12, (0, 2, 1)), np.complex128).tofile(f)
        np.ascontiguousarray(np.transpose(model.v12, (0, 2, 1)), np.complex128).tofile(f)
        np.ascontiguousarray(model.bc, np.complex128).tofile(f)
        np.array([p.dtau, model.M, model.D, p.biased, model.jump_distance, p.hop, p.lam], np.float64).tofile(f)
        np.array([p.ncopy, int(bool(p.tr_weight))], np.int32).tofile(f)
        np.ascontiguousarray(fields, np.float64).tofile(f)
        np.asfortranarray(slater, np.complex128).ravel(order="F").tofile(f)
        np.array([global_update_loop], np.int32).tofile(f)
        np.array([global_distance * model.char_length], np.float64).tofile(f)
        np.array([nbins, nmeas_per_bin, meas_number, warmup, meas_interval, checkpoint_loop, int(bool(resume)),
                  int(bool(proj))], np.int32).tofile(f)


def read_driver_output(path: str, model: Model, nsweeps: int) -> dict:
    p = model.p
    n_acc = nsweeps * p.Lt * 4 * model.npla
    with open(path, "rb") as f:
        acc = np.fromfile(f, np.int8, n_acc).astype(np.uint8)
        g = np.fromfile(f, np.complex128, model.Ns * model.Ns).reshape((model.Ns, model.Ns), order="F")
        fields = np.fromfile(f, np.float64, model.Nb * p.Lt).reshape((p.Lt, model.Nb))
        meas = np.fromfile(f, np.complex128, model.total_width)
        n_meas = int(np.fromfile(f, np.int32, 1)[0])
        cnt = np.fromfile(f, np.int64, 4)
    return {"accepts": acc, "g": g, "fields": fields, "meas_sum": meas, "n_meas": n_meas,
            "positive_accept": int(cnt[0]), "negative_accept": int(cnt[1]),
            "global_accept": int(cnt[2]), "global_reject": int(cnt[3])}


def sweep_batch(chains, forwards, draws, measure: bool = False, many: bool = False, max_concurrent: int = 0):
    """One sweep of every chain, at the same time (pqmc_sweep_batch): one
    host thread per chain, each on its chain's stream.  Returns the
    measurements (one total_width vector per chain) when measure is set.
    many: any number of chains (pqmc_sweep_many), each started as soon as
    its family passes fit the device beside the sweeps already running,
    with up to max_concurrent host threads (0: 16)."""
    n = len(chains)
    if n < 1 or len(forwards) != n or len(draws) != n:
        raise ValueError("one direction and one draw array per chain")
    us = []
    for ch, d in zip(chains, draws):
        u = np.ascontiguousarray(d, np.float32)
        if u.size != 2 * ch.Nb * ch.model.p.Lt:
            raise ValueError("need 2*Nb*Lt draws per chain")
        us.append(u)
    hs = (C.c_void_p * n)(*[ch.h.value for ch in chains])
    fw = (C.c_int * n)(*[int(bool(f)) for f in forwards])
    up = (C.c_void_p * n)(*[u.ctypes.data for u in us])
    outs = [np.empty(ch.total_width, np.complex128) for ch in chains] if measure else None
    op = (C.c_void_p * n)(*[o.ctypes.data for o in outs]) if measure else None
    if many:
        _lib.check(_lib.lib().pqmc_sweep_many(hs, n, fw, up, int(bool(measure)), op, int(max_concurrent)))
    else:
        _lib.check(_lib.lib().pqmc_sweep_batch(hs, n, fw, up, int(bool(measure)), op))
    return outs


def sweep_many(chains, forwards, draws, measure: bool = False, max_concurrent: int = 0):
    """One sweep of every chain of any number of chains of any sizes
    (pqmc_sweep_many)."""
    return sweep_batch(chains, forwards, draws, measure, many=True, max_concurrent=max_concurrent)


class GlobalMoves:
    """The global phonon moves of the sweep loop (src/Main_PQMC.f90:87-95).

    ``update(kind)`` is update_global (src/mod_update.f90:251-290): propose,
    weigh the new fields on the GPU (cal_new_weight_T0), and accept with
    cal_prob when rands() < prob.  ``run()`` is the block the sweep loop runs
    before a backward sweep: kspace, rotate and kspace_time, then init_g(1)
    if any of them was accepted.
    """

    def __init__(self, chain: Chain, rng: Ran2, distance: float = 1.0):
        self.chain, self.rng, self.distance = chain, rng, distance
        self.accepted = 0
        self.rejected = 0

    def update(self, kind: str, x: np.ndarray, lncw: complex):
        x_new, ph = propose_global(self.chain.model, x, kind, self.rng, self.distance)
        w = self.chain.global_weight(x_new)
        prob = min(1.0, math.exp(min(10.0, (w + ph - lncw).real)))
        if float(self.rng.draws(1)[0]) < prob:
            self.accepted += 1
            return True, x_new, w
        self.rejected += 1
        return False, x, lncw

    def run(self):
        x = self.chain.get_fields()
        lncw = self.chain.ln_cw
        flags = []
        for kind in GLOBAL_KINDS:
            acc, x, lncw = self.update(kind, x, lncw)
            flags.append(acc)
        if any(flags):
            self.chain.set_fields(x)
            self.chain.init_g(1)
        return flags


def autotune_pass(model: Model, x: np.ndarray, slater: np.ndarray | None = None,
                  block_bonds=(2, 4, 6, 8), helper_bonds=(1, 2, 4), slices: int = 20,
                  finite_temperature: bool = False, seed: int = 1):
    """Pick the family pass's shape for one lattice by timing it (no
    reference counterpart).  For every (block_bonds, helper_bonds) pair that
    the device accepts, a chain on these fields runs the first `slices`
    slices of a backward sweep from init_g(1) twice, the first as a warm-up;
    the second is timed.  Every shape gives the same accept sequence, so the
    choice only moves time.  Returns ((block_bonds, helper_bonds), {pair:
    seconds})."""
    import time
    u = Ran2(seed).draws(2 * model.Nb * model.p.Lt)
    times = {}
    for hb in helper_bonds:
        for bb in block_bonds:
            try:
                ch = Chain(model, block_bonds=bb, helper_bonds=hb, finite_temperature=finite_temperature)
            except _lib.PQMCError:
                continue
            try:
                ch.set_fields(x)
                if not finite_temperature:
                    ch.set_trial(slater)
                for rep in range(2):
                    ch.init_g(1)
                    ch.synchronize()
                    t0 = time.perf_counter()
                    ch.sweep(False, u, ntc=slices)
                    ch.synchronize()
                    dt = time.perf_counter() - t0
                times[(bb, hb)] = dt
            finally:
                ch.close()
    if not times:
        raise ValueError("no family-pass shape fits the device for this lattice")
    return min(times, key=times.get), times


def bosonic_action(model: Model, x: np.ndarray) -> float:
    """The phonon action whose differences cal_energy_difference
    (src/mod_update.f90:203-228) takes: delt sum_t sum_i [M/(2 delt^2)
    (x_t - x_(t-1))^2 + D/2 x_t^2 - biased x_t], periodic in t; x is (Lt, Nb)."""
    dt = model.p.dtau
    kin = 0.5 * model.M / dt ** 2 * ((x - np.roll(x, 1, axis=0)) ** 2).sum()
    return float(dt * (kin + (0.5 * model.D * x ** 2 - model.p.biased * x).sum()))


def bosonic_gradient(model: Model, x: np.ndarray) -> np.ndarray:
    """d bosonic_action / dx, (Lt, Nb)."""
    dt = model.p.dtau
    lap = 2 * x - np.roll(x, 1, axis=0) - np.roll(x, -1, axis=0)
    return dt * (model.M / dt ** 2 * lap + model.D * x - model.p.biased)


def leapfrog(x: np.ndarray, p: np.ndarray, force, step: float, nsteps: int, minv=None):
    """nsteps leapfrog steps of dx/ds = minv(p), dp/ds = force(x) (minv:
    the inverse mass matrix applied to p, identity if None); reversible and
    volume-preserving, so with an exact Metropolis test it leaves the target
    distribution unchanged."""
    minv = (lambda q: q) if minv is None else minv
    x = x.copy()
    p = p + 0.5 * step * force(x)
    for k in range(nsteps):
        x = x + step * minv(p)
        p = p + (step if k + 1 < nsteps else 0.5 * step) * force(x)
    return x, p


class BosonicMass:
    """Fourier-accelerated mass matrix: the Hessian of bosonic_action,
    K = delt (M/delt^2 (2 - shift - shift^-1) + D) per field, circulant in
    time, so that every free phonon mode moves at unit frequency and the
    leapfrog error no longer grows with M/delt^2.  Its time-Fourier
    eigenvalues are delt (M/delt^2 (2 - 2 cos(2 pi k/Lt)) + D); D > 0 keeps
    them positive."""

    def __init__(self, model: Model):
        dt, Lt = model.p.dtau, model.p.Lt
        if not model.D > 0:
            raise ValueError("the bosonic mass matrix needs D > 0")
        k = np.arange(Lt // 2 + 1)
        self.khat = dt * (model.M / dt ** 2 * (2 - 2 * np.cos(2 * np.pi * k / Lt)) + model.D)
        self.Lt = Lt

    def _apply(self, v, w):
        return np.fft.irfft(w[:, None] * np.fft.rfft(v, axis=0), n=self.Lt, axis=0)

    def inverse(self, p):
        return self._apply(p, 1.0 / self.khat)

    def sample(self, rng, shape):
        """p ~ N(0, K)"""
        return self._apply(rng.standard_normal(shape), np.sqrt(self.khat))

    def kinetic(self, p):
        return 0.5 * float((p * self.inverse(p)).sum())


class HybridMC:
    """Hybrid-Monte-Carlo global phonon move (no counterpart in the
    reference; it sits beside update_global, src/mod_update.f90:251-290).
    The action is S(x) = bosonic_action(x) - Re ln_cw(x); the force
    -dS/dx = -d bosonic/dx + Re(Chain.gradient()) drives a leapfrog
    trajectory from Gaussian momenta, and the endpoint is accepted with the
    exact full-determinant weight (Chain.global_weight) in the Metropolis
    test min(1, exp(-dH)).  mass = "bosonic" (the default when D > 0) uses
    the Fourier-accelerated BosonicMass, "unit" the identity.  The chain is
    left after init_g(1), ready for a backward sweep, as after the sweep
    loop's global moves (src/Main_PQMC.f90:87-95).
    """

    def __init__(self, chain: Chain, step: float = 0.05, nsteps: int = 10, seed: int = 0,
                 mass: str | None = None):
        self.chain, self.step, self.nsteps = chain, float(step), int(nsteps)
        self.rng = np.random.default_rng(seed)
        if mass is None:
            mass = "bosonic" if chain.model.D > 0 else "unit"
        if mass not in ("bosonic", "unit"):
            raise ValueError("mass must be 'bosonic' or 'unit'")
        self.mass = BosonicMass(chain.model) if mass == "bosonic" else None
        self.accepted = 0
        self.rejected = 0
        self.last_dH = 0.0

    def force(self, x: np.ndarray) -> np.ndarray:
        c = self.chain
        c.set_fields(x)
        c.init_g(1)
        return c.gradient().real - bosonic_gradient(c.model, x)

    def kinetic(self, p: np.ndarray) -> float:
        return self.mass.kinetic(p) if self.mass else 0.5 * float((p * p).sum())

    def momenta(self, shape) -> np.ndarray:
        return self.mass.sample(self.rng, shape) if self.mass else self.rng.standard_normal(shape)

    def hamiltonian(self, x: np.ndarray, p: np.ndarray, lncw: complex) -> float:
        return bosonic_action(self.chain.model, x) - lncw.real + self.kinetic(p)

    def trajectory(self, x: np.ndarray, p: np.ndarray):
        """(x, p) after one leapfrog trajectory; the chain's fields are left at x_end"""
        return leapfrog(x, p, self.force, self.step, self.nsteps, self.mass.inverse if self.mass else None)

    def update(self) -> bool:
        c = self.chain
        x0 = c.get_fields()
        w0 = c.global_weight(x0)
        p0 = self.momenta(x0.shape)
        x1, p1 = self.trajectory(x0, p0)
        w1 = c.global_weight(x1)
        dH = self.hamiltonian(x1, p1, w1) - self.hamiltonian(x0, p0, w0)
        self.last_dH = dH
        acc = bool(self.rng.random() < math.exp(min(0.0, -dH)))
        c.set_fields(x1 if acc else x0)
        c.init_g(1)
        if acc:
            self.accepted += 1
        else:
            self.rejected += 1
        return acc
