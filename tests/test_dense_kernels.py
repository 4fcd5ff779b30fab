"""The stabilization's dense kernels on their own, against numpy.

`pqmc_zgemm` (C = alpha op(A) op(B) + beta C, the ZGEMM calls of
cal_g_and_det_T0, src/mod_evolution.f90:344-350) in all four op forms, and
`pqmc_trsm_upper` (Z <- R^-1 Z, the solve of G = QR (L R)^-1 QL^H) on both of
its paths: R's panels in registers (n <= 272) and through LDS (larger n).
Sizes are the 14x14 stabilization's (ne = 196, Ns = 392) plus ragged ones.
Both run through `pqmc_dense_check` on host arrays (column-major complex).
"""
import numpy as np
import pytest

from pqmc_allinone_amd import _lib

pytestmark = pytest.mark.gpu


def _rand(rng, m, n):
    return rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))


def _f(a):
    """Column-major complex128 copy."""
    return np.asfortranarray(a, np.complex128)


@pytest.mark.parametrize("opa,opb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(196, 392, 196), (392, 392, 196), (37, 53, 71)])
def test_zgemm_matches_numpy(opa, opb, M, N, K):
    L = _lib.lib()
    rng = np.random.default_rng(1000 * M + 10 * N + 2 * opa + opb)
    A = _f(_rand(rng, K, M) if opa else _rand(rng, M, K))
    B = _f(_rand(rng, N, K) if opb else _rand(rng, K, N))
    C0 = _f(_rand(rng, M, N))
    alpha, beta = 0.7 - 0.2j, -0.3 + 0.5j
    ab = np.array([alpha.real, alpha.imag, beta.real, beta.imag])
    got = C0.copy(order="F")
    _lib.check(L.pqmc_dense_check(0, opa, opb, M, N, K, ab.ctypes.data, A.ctypes.data, A.shape[0], B.ctypes.data,
                                  B.shape[0], got.ctypes.data, M))
    opA = A.conj().T if opa else A
    opB = B.conj().T if opb else B
    want = alpha * (opA @ opB) + beta * C0
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 1e-13, err


@pytest.mark.parametrize("n,ncols", [(196, 392), (196, 5), (100, 37), (300, 64)])
def test_trsm_upper_matches_numpy(n, ncols):
    L = _lib.lib()
    rng = np.random.default_rng(7 * n + ncols)
    # upper triangular, diagonal well away from zero (as R of a QR with its
    # scales divided out); condition number 8-31 at these sizes
    R = np.triu(_rand(rng, n, n)) * 0.1
    R[np.diag_indices(n)] = (1.0 + np.abs(rng.standard_normal(n))) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    R = _f(R)
    Z = _f(_rand(rng, n, ncols))
    got = Z.copy(order="F")
    _lib.check(L.pqmc_dense_check(1, 0, 0, n, ncols, 0, None, R.ctypes.data, n, None, 0, got.ctypes.data, n))
    want = np.linalg.solve(R, Z)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 1e-12, err
    # the residual R X - Z, independent of numpy's solver
    res = np.abs(R @ got - Z).max() / np.abs(Z).max()
    assert res < 1e-12, res
