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
