"""The oracle's dd residual R = B - A X (pddblas::linearsolve's refinement
step, vcp/pddblas.hpp:459-466), pinned by exact rational arithmetic: within
2^-100 (|B| + |A||X|) of the exact residual, exactly zero where the product
is exact (integer data), and the subtraction's sign and low parts matter."""
from fractions import Fraction as F

import numpy as np
import pytest

import oracle_ffi as O


def dd_rand(rng, r, c):
    hi = rng.uniform(-1, 1, (r, c)) * 2.0 ** rng.integers(-8, 8, (r, c))
    lo = hi * 2.0 ** -53 * rng.uniform(-1, 1, (r, c))
    s = hi + lo
    X = O.pairs(r, c)
    X[..., 0] = s
    X[..., 1] = lo - (s - hi)
    return X


def val(p):
    return F(p[0]) + F(p[1])


@pytest.mark.parametrize("m,n,k,seed", [(3, 2, 5, 1), (7, 3, 17, 2), (5, 5, 64, 3)])
def test_dd_residual_close_to_exact(m, n, k, seed):
    rng = np.random.default_rng(seed)
    A, X, B = dd_rand(rng, m, k), dd_rand(rng, k, n), dd_rand(rng, m, n)
    R = O.dd_residual(A, X, B)
    for i in range(m):
        for j in range(n):
            exact = val(B[i, j]) - sum(val(A[i, t]) * val(X[t, j]) for t in range(k))
            scale = abs(val(B[i, j])) + sum(abs(val(A[i, t]) * val(X[t, j])) for t in range(k))
            assert abs(val(R[i, j]) - exact) <= F(2) ** -100 * scale
            assert abs(R[i, j, 1]) <= 2.0 ** -52 * abs(R[i, j, 0])   # normalised


def test_dd_residual_exact_for_integer_data():
    rng = np.random.default_rng(4)
    A = O.pairs(6, 6)
    A[..., 0] = rng.integers(-9, 10, (6, 6))
    X = O.pairs(6, 2)
    X[..., 0] = rng.integers(-9, 10, (6, 2))
    B = O.pairs(6, 2)
    B[..., 0] = A[..., 0] @ X[..., 0]
    R = O.dd_residual(A, X, B)
    assert (R == 0.0).all()
    B[..., 1] = 2.0 ** -70   # only the low parts of b differ: R keeps them
    R = O.dd_residual(A, X, B)
    assert (R[..., 0] == 2.0 ** -70).all() and (R[..., 1] == 0.0).all()
