"""EPSOCZ model set-up on the host: lattice, bond families, fields, trial state.

This restates the host-side set-up of the reference (it is not on the GPU
path) so the C ABI receives the same tables the reference builds:

* lattice indexing ``cell = x + L1*y``, ``site = 2*cell + subsite``
  (src/mod_lattice.f90:106-162, 229-280);
* the four checkerboard bond families of input/model/EPSOCZ.nml:120-238,
  built as set_typed_pf does (src/mod_phonon_field.f90:249-421): plaquette
  order, site lists, boundary phases, bond -> field map;
* derived parameters of init / init_pf_para (src/input.f90:167-216).

Synthetic inputs (documented, not reference logic): fields are i.i.d.
N(0,1) from ``numpy.random.default_rng(seed)`` in boson_field(Nb, Lt)
order, and the trial state is the lowest ``ne`` eigenvectors of the
non-interacting hopping matrix plus a 1e-3 uniform random hopping
perturbation drawn from the same generator after the fields.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# EPSOCZ.nml:120-238 -- (offset, pla_int_subsites rows (dx, dy, subsite), bf_index, Vmatrix(1:2,3:4))
_FAMILIES = (
    ((0, 0), ((0, 0, 1), (0, 0, 2), (1, 0, 1), (1, 0, 2)), 1, ((0, -1), (1, 0))),
    ((1, 0), ((0, 0, 1), (0, 0, 2), (1, 0, 1), (1, 0, 2)), 1, ((0, -1), (1, 0))),
    ((0, 0), ((0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 1, 2)), 2, ((0, 1j), (1j, 0))),
    ((0, 1), ((0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 1, 2)), 2, ((0, 1j), (1j, 0))),
)
# pla_tsl_dvec_uc = (2,0),(1,1) for every family: cells with (dx - dy) even
_TSL_INV = np.array([[0.5, -0.5], [0.0, 1.0]])


@dataclasses.dataclass
class ModelParams:
    L1: int
    L2: int
    Lt: int
    ngroup: int
    ne: int
    dtau: float = 0.1
    lam: float = 0.64
    omega: float = 1.0
    jump: float = 1.0
    hop: float = 1.0
    bc_phase: tuple = (0.0, 0.0)
    biased: float = 0.0
    ncopy: int = 1
    tr_weight: bool = False


@dataclasses.dataclass
class Model:
    p: ModelParams
    Ncell: int
    Ns: int
    Nb: int
    npla: int
    nblock: int
    total_width: int
    M: float
    D: float
    char_length: float
    jump_distance: float
    sites: np.ndarray  # int32 [4, npla, 4]
    bfi: np.ndarray    # int32 [4, npla]
    cross: np.ndarray  # int32 [4, npla]
    bcph: np.ndarray   # complex128 [4, npla, 4, 4] (BC_phases(i, j) at [.., i, j])
    kcoe: np.ndarray   # complex128 [4]
    vcoe: np.ndarray   # complex128 [4]
    k12: np.ndarray    # complex128 [4, 2, 2] ([f, a, b] = Kmatrix(1+a, 3+b))
    v12: np.ndarray    # complex128 [4, 2, 2]
    bc: np.ndarray     # complex128 [2]


def build_model(p: ModelParams) -> Model:
    """Tables of set_typed_pf + init + init_pf_para for the EPSOCZ model."""
    if p.L1 < 2 or p.L2 < 2 or (p.L1 % 2) or (p.L2 % 2):
        raise ValueError("the EPSOCZ checkerboard needs even L1, L2 >= 2")
    L = (p.L1, p.L2)
    Ncell = p.L1 * p.L2
    Ns = 2 * Ncell
    npla = Ns // 4
    # init (src/input.f90:167-205)
    D = 1.0
    M = D * p.omega ** (-2)
    if M < 0.001:
        M = 0.0
        char_length = math.sqrt(1.0 / (D * p.lam))
    else:
        char_length = 1.0 / (M * p.omega)
    jump_distance = p.jump * (0.1 / p.dtau) * char_length
    bc = np.exp(1j * np.array(p.bc_phase, dtype=float) * math.acos(-1.0))
    sites = np.zeros((4, npla, 4), np.int32)
    bfi = np.zeros((4, npla), np.int32)
    cross = np.zeros((4, npla), np.int32)
    bcph = np.ones((4, npla, 4, 4), np.complex128)
    for f, (off, sub, bfx, _) in enumerate(_FAMILIES):
        cnt = 0
        for c in range(Ncell):
            dp = np.array([c % p.L1, c // p.L1], float) - off
            for j in range(2):
                if dp[j] < -0.0001:
                    dp[j] += L[j]
                if dp[j] >= L[j] - 0.0001:
                    dp[j] -= L[j]
            xp = _TSL_INV @ dp
            if np.any(np.abs(np.rint(xp) - xp) > 1e-6):
                continue
            sites[f, cnt, 0] = 2 * c + (sub[0][2] - 1)
            cnt += 1
        if cnt != npla:
            raise RuntimeError("plaquette count mismatch")
        for b in range(npla):
            c0 = sites[f, b, 0] // 2
            pos = (c0 % p.L1, c0 // p.L1)
            for i in range(1, 4):
                cd = [pos[j] + sub[i][j] for j in range(2)]
                for j in range(2):
                    if cd[j] > L[j] - 0.0001:
                        cd[j] -= L[j]
                    if cd[j] < -0.0001:
                        cd[j] += L[j]
                cell = (cd[0] % p.L1) + p.L1 * (cd[1] % p.L2)
                sites[f, b, i] = 2 * cell + (sub[i][2] - 1)
            cr = 0
            for i in range(4):
                ci = sites[f, b, i] // 2
                idp = (ci % p.L1, ci // p.L1)
                for j in range(4):
                    v = 1.0 + 0j
                    for k in range(2):
                        jd = idp[k] - sub[i][k] + sub[j][k]
                        if jd >= L[k] - 0.0001:
                            cr = 1
                            v *= bc[k]
                        if jd < -0.0001:
                            cr = 1
                            v *= np.conj(bc[k])
                    bcph[f, b, i, j] = v
            cross[f, b] = cr
            bfi[f, b] = (bfx - 1) + 2 * c0
    kcoe = np.full(4, -p.hop, np.complex128)
    vcoe = np.full(4, p.lam, np.complex128)
    k12 = np.tile(np.eye(2, dtype=np.complex128), (4, 1, 1))
    v12 = np.array([np.array(fam[3], np.complex128) for fam in _FAMILIES])
    return Model(p=p, Ncell=Ncell, Ns=Ns, Nb=2 * Ncell, npla=npla, nblock=p.Lt // p.ngroup + 2,
                 total_width=3 + 19 * 2 * Ncell, M=M, D=D, char_length=char_length,
                 jump_distance=jump_distance, sites=sites, bfi=bfi, cross=cross, bcph=bcph,
                 kcoe=kcoe, vcoe=vcoe, k12=k12, v12=v12, bc=bc)


def hopping_matrix(m: Model) -> np.ndarray:
    """Non-interacting H = sum over bonds of K_coe * Kmatrix (with BC phases)."""
    H = np.zeros((m.Ns, m.Ns), np.complex128)
    for f in range(4):
        for b in range(m.npla):
            S = m.sites[f, b]
            for a in range(2):
                for c in range(2):
                    v = m.kcoe[f] * m.k12[f, a, c]
                    if m.cross[f, b]:
                        v *= m.bcph[f, b, a, 2 + c]
                    if v != 0:
                        H[S[a], S[2 + c]] += v
                        H[S[2 + c], S[a]] += np.conj(v)
    return H


def synthetic_inputs(m: Model, seed: int):
    """Fields boson_field(Nb, Lt) ~ N(0,1) and the trial Slater matrix (Ns x ne)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((m.p.Lt, m.Nb))  # row t holds boson_field(:, t)
    H = hopping_matrix(m)
    pert = rng.uniform(-1.0, 1.0, size=(4, m.npla, 2))
    for f in range(4):
        for b in range(m.npla):
            S = m.sites[f, b]
            for a in range(2):
                H[S[a], S[2 + a]] += 1e-3 * pert[f, b, a]
                H[S[2 + a], S[a]] += 1e-3 * pert[f, b, a]
    w, v = np.linalg.eigh(H)
    slater = np.ascontiguousarray(v[:, : m.p.ne].astype(np.complex128))
    return np.ascontiguousarray(x), np.asfortranarray(slater)


def sweep_draw_count(m: Model) -> int:
    """ran2 draws per sweep: two per bond per slice (src/mod_update.f90:84,117)."""
    return 2 * m.Nb * m.p.Lt


# The mixed batch of independent chains: 32 chains at each of L = 8, 10, 12,
# 14 (Ns = 2 L^2, ne = L^2), Lt cycling through 400, 600, 800, 1000 within
# each lattice, fields, trial perturbation and draws from seed 1000 + i.
MIXED_LATTICES = (8, 10, 12, 14)
MIXED_LTS = (400, 600, 800, 1000)


def mixed_chain_specs(n: int = 128, seed0: int = 1000):
    """(L, Lt, seed) of chain i of the mixed batch, i = 0..n-1."""
    per = max(1, n // len(MIXED_LATTICES))
    return [(MIXED_LATTICES[min(i // per, len(MIXED_LATTICES) - 1)], MIXED_LTS[i % len(MIXED_LTS)], seed0 + i)
            for i in range(n)]
