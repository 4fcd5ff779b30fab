"""Independent brute-force references for tiny lattices (numpy/scipy only).

These pin the oracle without restating its algorithm: dense matrix
exponentials instead of the closed form, explicit dense B-string products
instead of checkerboard sweeps with QR checkpoints, from-scratch G instead
of rank-4 updates, and a Fock-space many-body expectation value instead of
Wick contractions.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.linalg as sla

# Kmatrix and Vmatrix of input/model/EPSOCZ.nml:139-237 (full 4x4)
K4 = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], np.complex128)
V4_X = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], np.complex128)
V4_Y = np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]], np.complex128)


def bond_h(model, f, b, x):
    """H_b = (K_coe K + V_coe x V) .* BC_phases (get_expKV_general, mod_phonon_field.f90:507-523)."""
    V = V4_X if f < 2 else V4_Y
    h = model.kcoe[f] * K4 + model.vcoe[f] * x * V
    if model.cross[f, b]:
        h = h * model.bcph[f, b]
    return h


def bond_exp(model, f, b, x, inv=False):
    return sla.expm((1.0 if inv else -1.0) * model.p.dtau * bond_h(model, f, b, x))


def family_matrix(model, x, f, t, inv=False):
    """Dense E_f(t) (Ns x Ns) from the fields x[t-1, :]."""
    E = np.eye(model.Ns, dtype=np.complex128)
    for b in range(model.npla):
        S = model.sites[f, b]
        xv = x[t - 1, model.bfi[f, b]]
        E[np.ix_(S, S)] = bond_exp(model, f, b, xv, inv)
    return E


def slice_matrix(model, x, t):
    """B(t) = E_4 E_3 E_2 E_1 (left_evolve applies family 1 first)."""
    B = np.eye(model.Ns, dtype=np.complex128)
    for f in range(4):
        B = family_matrix(model, x, f, t) @ B
    return B


def g_from_product(R, L):
    """g_h = R (L R)^-1 L."""
    return R @ np.linalg.solve(L @ R, L)


def g_after_slice(model, x, P, tpos, fam_done=4):
    """g_h with slices 1..tpos-1 complete and families 1..fam_done of slice tpos applied on the right side."""
    R = P.copy()
    for t in range(1, tpos):
        R = slice_matrix(model, x, t) @ R
    for f in range(fam_done):
        R = family_matrix(model, x, f, tpos) @ R
    L = P.conj().T.copy()
    for t in range(model.p.Lt, tpos, -1):
        L = L @ slice_matrix(model, x, t)
    for f in range(3, fam_done - 1, -1):
        L = L @ family_matrix(model, x, f, tpos)
    return g_from_product(R, L)


def bosonic_action_delta(model, x, field, t, jump):
    """-delt (E_2 - E_1) with periodic time neighbours (cal_energy_difference)."""
    Lt = model.p.Lt
    be = x[(t - 2) % Lt, field]
    af = x[t % Lt, field]
    now = x[t - 1, field]
    nw = now + jump
    M, D, dt = model.M, model.D, model.p.dtau

    def E(v):
        return 0.5 * M * ((be - v) ** 2 + (af - v) ** 2) / dt ** 2 + 0.5 * D * v ** 2 - model.p.biased * v

    return -dt * (E(nw) - E(now))


# ---- Fock space (Jordan-Wigner) ----------------------------------------

def fock_ops(ns):
    """Annihilation operators c_i as dense 2^ns matrices."""
    dim = 1 << ns
    ops = []
    for i in range(ns):
        c = np.zeros((dim, dim))
        for s in range(dim):
            if s >> i & 1:
                sign = (-1) ** bin(s & ((1 << i) - 1)).count("1")
                c[s ^ (1 << i), s] = sign
        ops.append(c)
    return ops


def slater_ket(ops, R):
    """prod_k (sum_i R_ik c_i^dag) |0>."""
    dim = ops[0].shape[0]
    v = np.zeros(dim, np.complex128)
    v[0] = 1.0
    for k in range(R.shape[1]):
        cd = sum(R[i, k] * ops[i].T for i in range(R.shape[0]))
        v = cd @ v
    return v


def slater_bra(ops, L):
    """<0| prod_k (sum_i L_ki c_i), as a row vector, for <Psi_L| with L = (ne x Ns)."""
    # <Psi_L| = (|Psi_L>)^dag with |Psi_L> = prod_k (sum_i conj(L_ki) c_i^dag)|0>
    return slater_ket(ops, L.conj().T).conj()


def mixed_expectation(ops, bra, ket, op):
    return (bra @ (op @ ket)) / (bra @ ket)


def all_site_subsets(ns, ne):
    return list(itertools.combinations(range(ns), ne))


class FockSector:
    """The fixed-particle-number sector of ns modes with n particles.

    Basis state `mask` is c^dag_{i1} c^dag_{i2} ... |0> with i1 < i2 < ...
    (the convention of `fock_ops`), so c_i and c^dag_i carry the sign
    (-1)^(number of occupied modes below i).  Slater determinants have the
    amplitudes det(R[occ, :]) (ket) and det(L[:, occ]) (bra), each up to one
    overall constant that cancels in a mixed expectation value."""

    def __init__(self, ns: int, n: int):
        self.ns, self.n = ns, n
        masks = np.array([m for m in range(1 << ns) if bin(m).count("1") == n], np.int64)
        self.masks = masks
        self.index = -np.ones(1 << ns, np.int64)
        self.index[masks] = np.arange(masks.size)
        bits = (masks[:, None] >> np.arange(ns)[None, :]) & 1
        self.occ = np.array([np.nonzero(b)[0] for b in bits], np.int64)  # (nstates, n), ascending
        self._hop = {}

    @staticmethod
    def _sign(mask, i):
        below = mask & ((1 << i) - 1)
        cnt = np.zeros_like(below)
        for k in range(i):
            cnt += (below >> k) & 1
        return 1 - 2 * (cnt & 1)

    def _hop_tables(self, a, b):
        """c^dag_a c_b on this sector: source and destination state indices, signs."""
        key = (a, b)
        if key not in self._hop:
            m = self.masks
            src = np.nonzero((m >> b) & 1)[0]
            mb = m[src] & ~np.int64(1 << b)
            s1 = self._sign(m[src], b)
            if a != b:
                ok = ((mb >> a) & 1) == 0
                src, mb, s1 = src[ok], mb[ok], s1[ok]
            s2 = self._sign(mb, a)
            dst = self.index[mb | np.int64(1 << a)]
            self._hop[key] = (src, dst, (s1 * s2).astype(np.float64))
        return self._hop[key]

    def hop(self, a, b, v):
        """(c^dag_a c_b) v for a sector vector v (or a stack v[..., nstates])."""
        src, dst, sg = self._hop_tables(a, b)
        out = np.zeros_like(v)
        out[..., dst] = sg * v[..., src]
        return out

    def slater_ket(self, R):
        return np.linalg.det(R[self.occ, :])

    def slater_bra(self, L):
        return np.linalg.det(np.transpose(L[:, self.occ], (1, 0, 2)))

    def pair_table(self, bra, ket):
        """T[a*ns + b, c*ns + d] = <L| c^dag_a c_b c^dag_c c_d |R> / <L|R>, in
        Fock space, for every a, b, c, d.  <L| c^dag_a c_b is the vector
        (c^dag_b c_a)^T applied to <L| as a column (the operators are real)."""
        ns = self.ns
        K = np.empty((ns * ns, ket.size), np.complex128)
        B = np.empty((ns * ns, bra.size), np.complex128)
        for a in range(ns):
            for b in range(ns):
                K[a * ns + b] = self.hop(a, b, ket)
                B[a * ns + b] = self.hop(b, a, bra)
        return (B @ K.T) / (bra @ ket)

    def one_body(self, bra, ket):
        """<L| c^dag_a c_b |R> / <L|R> as [a, b]."""
        ns = self.ns
        out = np.empty((ns, ns), np.complex128)
        for a in range(ns):
            for b in range(ns):
                out[a, b] = bra @ self.hop(a, b, ket)
        return out / (bra @ ket)

    def pair_annihilation(self, bra, ket, i, j, k, l):
        """<L| c^dag_i c^dag_j c_k c_l |R> / <L|R> through the sectors n-1 and n-2."""
        down1 = FockSector(self.ns, self.n - 1)
        down2 = FockSector(self.ns, self.n - 2)

        def annihilate(sec_from, sec_to, i, v):
            m = sec_from.masks
            src = np.nonzero((m >> i) & 1)[0]
            dst = sec_to.index[m[src] & ~np.int64(1 << i)]
            out = np.zeros(sec_to.masks.size, np.complex128)
            out[dst] = sec_from._sign(m[src], i) * v[src]
            return out

        r = annihilate(down1, down2, k, annihilate(self, down1, l, ket))
        # <L| c^dag_i c^dag_j = (c_j c_i applied to <L| as a column)
        lv = annihilate(down1, down2, j, annihilate(self, down1, i, bra))
        return (lv @ r) / (bra @ ket)


def inv_extended(M):
    """Inverse by Gauss-Jordan with partial pivoting in numpy's extended
    precision (clongdouble): a reference for ill-conditioned 4x4 blocks."""
    n = M.shape[0]
    a = np.concatenate([M.astype(np.clongdouble), np.eye(n, dtype=np.clongdouble)], 1)
    for i in range(n):
        p = i + int(np.argmax(np.abs(a[i:, i])))
        a[[i, p]] = a[[p, i]]
        a[i] /= a[i, i]
        for k in range(n):
            if k != i:
                a[k] -= a[k, i] * a[i]
    return a[:, n:]


def sequential_updates_extended(G, deltas):
    """G after the rank-4 updates G += G[:, S] X (I - G)[S, :],
    X = Delta (I + G[S, S] Delta)^-1, applied in order (update_phonon_field,
    src/mod_update.f90:97-190), in extended precision.  deltas: list of
    (sites S, Delta)."""
    ns = G.shape[0]
    g = G.astype(np.clongdouble)
    eye = np.eye(ns, dtype=np.clongdouble)
    for S, D in deltas:
        Dl = D.astype(np.clongdouble)
        M = np.eye(4, dtype=np.clongdouble) + g[np.ix_(S, S)] @ Dl
        g = g + g[:, S] @ (Dl @ inv_extended(M)) @ (eye - g)[S, :]
    return g.astype(np.complex128)
