"""Pins the oracle's LAPACK-semantics LU and solves (dgetrf_/zgetrf_/dgetrs_/
zgetrs_ as linked by radau2a.c:273,277,404,417,479,498; generic/lapack_dfns.h)
against what the mathematics and LAPACK's documented semantics fix:

* P A = L U reconstruction with unit-lower L and 1-based ipiv (row interchange
  sequence), for real and complex matrices;
* partial pivoting picks the FIRST maximum of |a| (idamax) for real and of
  |Re| + |Im| (izamax / dcabs1) -- not the modulus -- for complex;
* info = index of the first exactly-zero pivot;
* solves agree with numpy.linalg.solve; brute force on 1x1 / 2x2.
"""
import numpy as np
import pytest


def reconstruct_real(lu, ipiv):
    n = lu.shape[0]
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    return L, U


def apply_pivots(A, ipiv):
    A = A.copy()
    for j, p in enumerate(ipiv):
        p -= 1
        if p != j:
            A[[j, p]] = A[[p, j]]
    return A


@pytest.mark.parametrize("n", [1, 2, 5, 14, 54])
def test_dgetrf_reconstruction(oracle_lib, n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) * 10 ** rng.uniform(-3, 3, size=(n, 1))
    lu, ipiv, info = oracle_lib.dgetrf(A)
    assert info == 0
    L, U = reconstruct_real(lu, ipiv)
    PA = apply_pivots(A, ipiv)
    assert np.abs(PA - L @ U).max() <= 1e-13 * np.abs(A).max()
    # partial pivoting bounds the multipliers
    assert np.abs(np.tril(lu, -1)).max() <= 1.0
    b = rng.normal(size=n)
    x = oracle_lib.dgetrs(lu, ipiv, b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("n", [1, 3, 14, 54])
def test_zgetrf_reconstruction(oracle_lib, n):
    rng = np.random.default_rng(100 + n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    lu, ipiv, info = oracle_lib.zgetrf(A)
    assert info == 0
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    PA = apply_pivots(A, ipiv)
    assert np.abs(PA - L @ U).max() <= 1e-13 * np.abs(A).max()
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    x = oracle_lib.zgetrs(lu, ipiv, b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_real_pivot_first_maximum(oracle_lib):
    # column 0 has |-3| tying with |3|: LAPACK idamax takes the first (row 1)
    A = np.array([[1.0, 2.0, 0.0], [-3.0, 1.0, 1.0], [3.0, 0.5, 2.0]])
    lu, ipiv, info = oracle_lib.dgetrf(A)
    assert ipiv[0] == 2


def test_complex_pivot_uses_abs1(oracle_lib):
    # |1.0+1.0i| = 1.414 > |1.35| but |Re|+|Im| = 2.0 > 1.35 as well; build a
    # case where the two rules disagree: 0.9+0.9i (abs1 1.8, modulus 1.27)
    # versus 1.5 (abs1 1.5, modulus 1.5).  izamax must pick row 0.
    A = np.array([[0.9 + 0.9j, 1.0], [1.5 + 0.0j, 2.0]])
    lu, ipiv, info = oracle_lib.zgetrf(A)
    assert ipiv[0] == 1
    assert np.abs(np.abs(A[0, 0]) - 1.2728) < 1e-3 and np.abs(A[1, 0]) == 1.5


def test_singular_info(oracle_lib):
    A = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])
    # column 1 becomes zero below the diagonal after step 0 when rows 0/1 are dependent
    lu, ipiv, info = oracle_lib.dgetrf(np.array([[0.0, 1.0], [0.0, 2.0]]))
    assert info == 1
    lu, ipiv, info = oracle_lib.dgetrf(np.array([[1.0, 2.0], [2.0, 4.0]]))
    assert info == 2
    lu, ipiv, info = oracle_lib.zgetrf(np.array([[0.0, 1.0], [0.0, 1.0j]]))
    assert info == 1
    lu, ipiv, info = oracle_lib.dgetrf(A)
    assert info in (0, 3)


def test_brute_force_2x2(oracle_lib):
    a, b, c, d = 2.0, -1.0, 4.0, 3.0
    lu, ipiv, info = oracle_lib.dgetrf(np.array([[a, b], [c, d]]))
    # pivot row 2 (|4| > |2|): L21 = a/c, U = [[c, d], [0, b - a/c*d]]
    assert list(ipiv) == [2, 2]
    assert np.allclose(lu, [[c, d], [a / c, b - a / c * d]], rtol=0, atol=1e-15)
