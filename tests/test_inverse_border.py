"""The bordering behind the device inverse of the verified solve: the midpoint
matrix M (n x n) is embedded as P = [[M, 0], [0, I]] of order npad (the next
multiple of 64) before LU with partial pivoting.  Checked here with LAPACK on
the host: P's pivots in M's columns are M's own (no pivot comes from the
border), P's inverse is [[inv(M), 0], [0, I]], and a singular M still meets an
exactly zero pivot."""
import numpy as np
import pytest
import scipy.linalg


def _border(M, npad):
    n = M.shape[0]
    P = np.eye(npad)
    P[:n, :n] = M
    return P


@pytest.mark.parametrize("n", [2, 17, 64, 65, 130])
def test_bordered_lu_keeps_pivots_and_inverse(n):
    rng = np.random.default_rng(n)
    M = rng.uniform(-1, 1, (n, n))   # general, not diagonally dominant: pivoting matters
    npad = (n + 63) // 64 * 64
    P = _border(M, npad)
    luM, pivM = scipy.linalg.lu_factor(M)
    luP, pivP = scipy.linalg.lu_factor(P)
    assert np.array_equal(pivP[:n], pivM)             # row interchanges only among M's rows
    assert np.array_equal(pivP[n:], np.arange(n, npad))
    assert (luP[n:, :n] == 0).all() and (luP[:n, n:] == 0).all()
    Pinv = scipy.linalg.lu_solve((luP, pivP), np.eye(npad))
    Minv = scipy.linalg.lu_solve((luM, pivM), np.eye(n))
    assert np.abs(Pinv[:n, :n] - Minv).max() <= 1e-9 * np.abs(Minv).max()
    assert np.array_equal(Pinv[n:, n:], np.eye(npad - n))
    assert (Pinv[:n, n:] == 0).all() and (Pinv[n:, :n] == 0).all()


def test_bordered_singular_matrix_meets_a_zero_pivot():
    n = 20
    rng = np.random.default_rng(3)
    M = rng.uniform(-1, 1, (n, n))
    M[:, 7] = 0.0
    P = _border(M, 64)
    with pytest.warns(scipy.linalg.LinAlgWarning):
        lu, piv = scipy.linalg.lu_factor(P)
    assert lu[7, 7] == 0.0   # the zero pivot stays at M's column; the border supplies none
