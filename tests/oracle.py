"""ctypes access to the oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")


class Params(C.Structure):
    _fields_ = [("L1", C.c_int), ("L2", C.c_int), ("Lt", C.c_int), ("ngroup", C.c_int), ("ne", C.c_int),
                ("dtau", C.c_double), ("lam", C.c_double), ("omega", C.c_double), ("jump", C.c_double),
                ("hop", C.c_double), ("bc_phase1", C.c_double), ("bc_phase2", C.c_double),
                ("biased", C.c_double), ("ncopy", C.c_int), ("tr_weight", C.c_int)]


class Dims(C.Structure):
    _fields_ = [("Ns", C.c_int), ("Ncell", C.c_int), ("Nb", C.c_int), ("nfam", C.c_int), ("npla", C.c_int),
                ("nblock", C.c_int), ("total_width", C.c_int), ("M", C.c_double), ("D", C.c_double),
                ("char_length", C.c_double), ("jump_distance", C.c_double)]


class Ran2State(C.Structure):
    _fields_ = [(n, C.c_int32) for n in "ijkmn"]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params)]
        for n in ["orc_destroy", "orc_init_expkv"]:
            getattr(L, n).argtypes = [C.c_void_p]
        L.orc_get_dims.argtypes = [C.c_void_p, C.POINTER(Dims)]
        L.orc_get_model.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.orc_get_coeffs.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        for n in ["orc_set_fields", "orc_get_fields", "orc_set_trial", "orc_get_g", "orc_set_g",
                  "orc_get_dstring", "orc_get_lncw"]:
            getattr(L, n).argtypes = [C.c_void_p, C.c_void_p]
        L.orc_get_qstring.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_get_expkv.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p]
        L.orc_expkv_epsoc.argtypes = [C.c_double] + [C.c_void_p] * 5 + [C.c_double, C.c_int, C.c_void_p]
        L.orc_init_g.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_ft.argtypes = [C.c_void_p, C.c_int]
        L.orc_left_evolve.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_right_evolve.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_wrap_g.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_update_phonon_field.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_get_g_scratch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_measure.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_sweep.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_counters.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_ran2_seed.argtypes = [C.POINTER(Ran2State), C.c_int32]
        L.orc_ran2_fill.argtypes = [C.POINTER(Ran2State), C.c_int64, C.c_void_p]
        L.orc_zqdr.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_zldq.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_zinverse.argtypes = [C.c_int, C.c_void_p]
        L.orc_zdet.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.orc_cal_new_weight.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_gradient.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_get_dexpkv.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p]
        L.orc_update_global.argtypes = [C.c_void_p, C.c_int, C.POINTER(GlobalParams), C.POINTER(Ran2State)]
        L.orc_update_global.restype = C.c_int
        L.orc_get_global_counters.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_propose_global.argtypes = [C.c_void_p, C.c_int, C.POINTER(GlobalParams), C.POINTER(Ran2State),
                                         C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def P(a):
    return a.ctypes.data


def ran2(seed: int, n: int) -> np.ndarray:
    s = Ran2State()
    lib().orc_ran2_seed(C.byref(s), int(seed))
    out = np.empty(int(n), np.float32)
    lib().orc_ran2_fill(C.byref(s), int(n), P(out))
    return out


class GlobalParams(C.Structure):
    _fields_ = [("distance", C.c_double)]


# kinds of update_global in the order the sweep loop calls them (src/Main_PQMC.f90:88-94)
GU_KSPACE, GU_ROTATE, GU_KSPACE_TIME = 0, 1, 2


class Ran2Stream:
    def __init__(self, seed: int):
        self.s = Ran2State()
        lib().orc_ran2_seed(C.byref(self.s), int(seed))

    def draws(self, n: int) -> np.ndarray:
        out = np.empty(int(n), np.float32)
        lib().orc_ran2_fill(C.byref(self.s), int(n), P(out))
        return out


class Oracle:
    """One oracle chain (the reference's module state on the CPU)."""

    def __init__(self, L1, L2, Lt, ngroup, ne, dtau=0.1, lam=0.64, omega=1.0, jump=1.0, hop=1.0,
                 bc_phase=(0.0, 0.0), biased=0.0, ncopy=1, tr_weight=False, ft=False):
        self.p = Params(L1, L2, Lt, ngroup, ne, dtau, lam, omega, jump, hop, bc_phase[0], bc_phase[1], biased,
                        ncopy, int(tr_weight))
        self.m = lib().orc_create(C.byref(self.p))
        d = Dims()
        lib().orc_get_dims(self.m, C.byref(d))
        self.d = d
        self.Ns, self.Nb, self.npla, self.nblock = d.Ns, d.Nb, d.npla, d.nblock
        self.Lt, self.ne, self.ngroup = Lt, ne, ngroup
        self.total_width = d.total_width
        if ft:
            lib().orc_set_ft(self.m, 1)

    def __del__(self):
        try:
            lib().orc_destroy(self.m)
        except Exception:
            pass

    def tables(self):
        sites = np.zeros((4, self.npla, 4), np.int32)
        bfi = np.zeros((4, self.npla), np.int32)
        cross = np.zeros((4, self.npla), np.int32)
        bcph = np.zeros((4, self.npla, 4, 4), np.complex128)  # [f,b,j,i] = BC_phases(i,j)
        lib().orc_get_model(self.m, P(sites), P(bfi), P(cross), P(bcph))
        return sites, bfi, cross, np.transpose(bcph, (0, 1, 3, 2))

    def set_fields(self, x):
        x = np.ascontiguousarray(x, np.float64)
        lib().orc_set_fields(self.m, P(x))

    def get_fields(self):
        x = np.empty((self.Lt, self.Nb))
        lib().orc_get_fields(self.m, P(x))
        return x

    def set_trial(self, slater):
        s = np.asfortranarray(slater, np.complex128)
        lib().orc_set_trial(self.m, P(s))

    def init(self, start_time=1):
        lib().orc_init_expkv(self.m)
        lib().orc_init_g(self.m, int(start_time))

    def init_expkv(self):
        lib().orc_init_expkv(self.m)

    def expkv(self, pf_id, pla, x, inv):
        e = np.zeros((4, 4), np.complex128, order="F")
        lib().orc_get_expkv(self.m, int(pf_id), int(pla), float(x), int(inv), P(e))
        return e

    @property
    def g(self):
        g = np.empty((self.Ns, self.Ns), np.complex128, order="F")
        lib().orc_get_g(self.m, P(g))
        return g

    @g.setter
    def g(self, v):
        v = np.asfortranarray(v, np.complex128)
        lib().orc_set_g(self.m, P(v))

    def q_string(self, block):
        q = np.empty((self.Ns, self.ne), np.complex128, order="F")
        lib().orc_get_qstring(self.m, int(block), P(q))
        return q

    def d_string(self):
        d = np.empty((self.ne, self.nblock), np.complex128, order="F")
        lib().orc_get_dstring(self.m, P(d))
        return d

    @property
    def ln_cw(self):
        v = np.zeros(2)
        lib().orc_get_lncw(self.m, P(v))
        return complex(v[0], v[1])

    def left_evolve(self, time, mat, pf_id, two_way):
        m = np.asfortranarray(mat, np.complex128).copy(order="F")
        lib().orc_left_evolve(self.m, int(time), P(m), int(m.shape[1]), int(pf_id), int(two_way))
        return m

    def right_evolve(self, time, mat, pf_id, two_way):
        m = np.asfortranarray(mat, np.complex128).copy(order="F")
        lib().orc_right_evolve(self.m, int(time), P(m), int(m.shape[0]), int(pf_id), int(two_way))
        return m

    def wrap_g(self, time, pf_id, forward):
        lib().orc_wrap_g(self.m, int(time), int(pf_id), int(bool(forward)))

    def update_phonon_field(self, pf_id, time, u):
        u = np.ascontiguousarray(u, np.float32)
        acc = np.zeros(self.npla, np.uint8)
        lib().orc_update_phonon_field(self.m, int(pf_id), int(time), P(u), P(acc))
        return acc

    def get_g_scratch(self, time, forward):
        e = C.c_double(0.0)
        lib().orc_get_g_scratch(self.m, int(time), int(bool(forward)), C.addressof(e))
        return e.value

    def measure(self, time):
        out = np.zeros(self.total_width, np.complex128)
        lib().orc_measure(self.m, int(time), P(out))
        return out

    def sweep(self, forward, u, ntc=0, measure=False):
        u = np.ascontiguousarray(u, np.float32)
        acc = np.zeros(self.Nb * self.Lt, np.uint8)
        out = np.zeros(self.total_width, np.complex128) if measure else None
        nm = C.c_int(0)
        lib().orc_sweep(self.m, int(bool(forward)), P(u), int(ntc), P(acc), P(out) if measure else None,
                        C.addressof(nm))
        return acc, out

    def cal_new_weight(self):
        out = np.zeros(2)
        lib().orc_cal_new_weight(self.m, P(out))
        return complex(out[0], out[1])

    def gradient(self):
        """complex d ln_cw / dx, shape (Lt, Nb)"""
        g = np.zeros((self.Lt, self.Nb), np.complex128)
        lib().orc_gradient(self.m, P(g))
        return g

    def dexpkv(self, pf_id, pla, x):
        e = np.zeros((4, 4), np.complex128, order="F")
        lib().orc_get_dexpkv(self.m, int(pf_id), int(pla), float(x), P(e))
        return e

    def update_global(self, kind, rng, distance=1.0):
        gp = GlobalParams(float(distance))
        return bool(lib().orc_update_global(self.m, int(kind), C.byref(gp), C.byref(rng.s)))

    def propose_global(self, kind, rng, distance=1.0):
        gp = GlobalParams(float(distance))
        x = np.zeros((self.Lt, self.Nb))
        ph = C.c_double(0)
        lib().orc_propose_global(self.m, int(kind), C.byref(gp), C.byref(rng.s), P(x), C.addressof(ph))
        return x, ph.value

    def global_counters(self):
        a, r = C.c_long(0), C.c_long(0)
        lib().orc_get_global_counters(self.m, C.addressof(a), C.addressof(r))
        return a.value, r.value

    def counters(self):
        pa, na, efm = C.c_long(0), C.c_long(0), C.c_double(0)
        lib().orc_get_counters(self.m, C.addressof(pa), C.addressof(na), C.addressof(efm))
        return pa.value, na.value, efm.value


def zqdr(a):
    a = np.asfortranarray(a, np.complex128).copy(order="F")
    m, n = a.shape
    r = np.zeros((n, n), np.complex128, order="F")
    d = np.zeros(n, np.complex128)
    lib().orc_zqdr(m, n, P(a), P(r), P(d))
    return a, d, r


def zldq(a):
    a = np.asfortranarray(a, np.complex128).copy(order="F")
    m, n = a.shape
    l = np.zeros((m, m), np.complex128, order="F")
    d = np.zeros(m, np.complex128)
    lib().orc_zldq(m, n, P(a), P(l), P(d))
    return a, d, l


def zinverse(a):
    a = np.asfortranarray(a, np.complex128).copy(order="F")
    lib().orc_zinverse(a.shape[0], P(a))
    return a


def zdet(a):
    a = np.asfortranarray(a, np.complex128)
    out = np.zeros(2)
    lib().orc_zdet(a.shape[0], P(a), P(out))
    return complex(out[0], out[1])


def expkv_epsoc(dtau, kcoe, vcoe, k12, v12, x, ph=None, inv=False):
    """The oracle's get_expKV_epsoc (src/mod_phonon_field.f90:572-604): one
    2x2 bond exponential; ph is the BC_phases(1,2) of a crossing bond."""
    a = [np.array([complex(v)], np.complex128) for v in (kcoe, vcoe, k12, v12)]
    p = np.array([complex(ph)], np.complex128) if ph is not None else None
    out = np.zeros(4, np.complex128)
    lib().orc_expkv_epsoc(float(dtau), *[P(v) for v in a], P(p) if p is not None else None, float(x), int(bool(inv)),
                          P(out))
    return out.reshape(2, 2).T  # column-major
