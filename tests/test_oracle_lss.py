"""The oracle's restatement of imats::linearsolve (vcp/imats.hpp:292-362),
pinned by mathematics: exact answers where the inverse is exact (the
tridiagonal (-1, 2, -1) matrix of order 2^p - 1 has a dyadic inverse), the
exact rational solutions of vertex systems inside the enclosure, a width
bound for point systems, the reference's verification failure, and the
nearest-rounded k-ascending FMA chain behind mx = R mid(b) against an exact
rational FMA chain."""
from fractions import Fraction as F

import numpy as np
import pytest

import oracle_ffi as O


def tridiag(n):
    return 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)


def tridiag_inverse(n):
    # (T^-1)_ij = min(i, j) (n + 1 - max(i, j)) / (n + 1), 1-based
    return np.array([[min(i, j) * (n + 1 - max(i, j)) / (n + 1) for j in range(1, n + 1)] for i in range(1, n + 1)])


def point_pairs(M):
    P = O.pairs(M.shape[0], M.shape[1])
    P[..., 0] = M
    P[..., 1] = M
    return P


def exact_solve(A, b):
    """Gaussian elimination in rationals (A, b: lists of Fractions)"""
    n = len(A)
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        p = max(range(c, n), key=lambda r: abs(M[r][c]))
        M[c], M[p] = M[p], M[c]
        for r in range(c + 1, n):
            f = M[r][c] / M[c][c]
            if f:
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    x = [F(0)] * n
    for r in range(n - 1, -1, -1):
        x[r] = (M[r][n] - sum(M[r][j] * x[j] for j in range(r + 1, n))) / M[r][r]
    return x


def random_system(rng, n, nrhs, rad_exp):
    A = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    b = rng.uniform(-1, 1, (n, nrhs))
    IA = O.pairs(n, n)
    IA[..., 0] = A - np.ldexp(np.abs(A), rad_exp)
    IA[..., 1] = A + np.ldexp(np.abs(A), rad_exp)
    Ib = O.pairs(n, nrhs)
    Ib[..., 0] = b - np.ldexp(np.abs(b), rad_exp)
    Ib[..., 1] = b + np.ldexp(np.abs(b), rad_exp)
    return IA, Ib, np.linalg.inv(A)


@pytest.mark.parametrize("n", [3, 7, 15])
@pytest.mark.parametrize("nrhs", [1, 2])
def test_lss_exact_inverse_gives_the_exact_point_solution(n, nrhs):
    R = tridiag_inverse(n)
    assert np.array_equal(R @ tridiag(n), np.eye(n))   # dyadic: exact in double
    b = np.arange(1, n * nrhs + 1, dtype=np.float64).reshape(n, nrhs)
    X, g, ok = O.lss(point_pairs(tridiag(n)), point_pairs(b), R)
    assert ok and g == 0.0
    for j in range(nrhs):
        x = exact_solve([[F(v) for v in row] for row in tridiag(n)], [F(v) for v in b[:, j]])
        assert [F(v) for v in X[:, j, 0]] == x and [F(v) for v in X[:, j, 1]] == x


@pytest.mark.parametrize("n,nrhs,seed", [(4, 1, 1), (5, 2, 2), (6, 1, 3)])
def test_lss_contains_the_solutions_of_vertex_systems(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    IA, Ib, R = random_system(rng, n, nrhs, -20)
    X, g, ok = O.lss(IA, Ib, R)
    assert ok and 0.0 < g < 1.0
    assert (X[..., 0] <= X[..., 1]).all()
    for trial in range(12):
        pick = lambda I, shape: np.where(rng.integers(0, 2, shape) == 1, I[..., 1], I[..., 0])
        As = pick(IA, (n, n)) if trial else 0.5 * IA[..., 0] + 0.5 * IA[..., 1]
        bs = pick(Ib, (n, nrhs))
        for j in range(nrhs):
            x = exact_solve([[F(v) for v in row] for row in As], [F(v) for v in bs[:, j]])
            for i in range(n):
                assert F(X[i, j, 0]) <= x[i] <= F(X[i, j, 1])


def test_lss_point_system_is_tight():
    rng = np.random.default_rng(4)
    n = 12
    A = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    b = rng.uniform(-1, 1, (n, 1))
    X, g, ok = O.lss(point_pairs(A), point_pairs(b), np.linalg.inv(A))
    assert ok
    x = exact_solve([[F(v) for v in row] for row in A], [F(v) for v in b[:, 0]])
    for i in range(n):
        assert F(X[i, 0, 0]) <= x[i] <= F(X[i, 0, 1])
        assert X[i, 0, 1] - X[i, 0, 0] <= 64 * n * 2.0 ** -52 * max(abs(float(x[i])), 1e-300)


def test_lss_verification_failure_when_R_is_not_an_inverse():
    n = 5
    IA = point_pairs(tridiag(n))
    Ib = point_pairs(np.ones((n, 1)))
    X, g, ok = O.lss(IA, Ib, np.zeros((n, n)))     # R A - I = -I: norm exactly 1
    assert g == 1.0 and not ok
    X, g, ok = O.lss(IA, Ib, 2.5 * np.linalg.inv(tridiag(n)))   # R A - I ~ 1.5 I
    assert g >= 1.0 and not ok


def test_lss_rejects_scalar_systems():
    with pytest.raises(ValueError):
        O.lss(point_pairs(np.ones((1, 1))), point_pairs(np.ones((1, 1))), np.ones((1, 1)))


def test_nearest_chain_is_the_correctly_rounded_fma_chain():
    """orc_rgemm in nearest mode (the product mx = R mid(b)) against the exact
    rational chain acc = RN(a b + acc), k ascending from 0"""
    rng = np.random.default_rng(5)
    m, n, k = 6, 3, 29
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * 2.0 ** rng.integers(-30, 30, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * 2.0 ** rng.integers(-30, 30, (k, n)))
    C = O.rgemm(A, B, 0)
    for i in range(m):
        for j in range(n):
            acc = 0.0
            for t in range(k):
                acc = float(F(A[i, t]) * F(B[t, j]) + F(acc))   # float(Fraction) rounds to nearest even
            assert C[i, j] == acc
