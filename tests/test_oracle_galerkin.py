"""The oracle's Galerkin matrix (Legendre_Bases_Generator::uhphiphi,
vcp/ldbase.hpp:955-1013) pinned by exact rational arithmetic: kv's interval
product is the tightest outward-rounded hull of the four endpoint products;
the matrix is symmetric and contains the exact sums for point data and for
vertex samples of interval data, with a small width."""
import math
from fractions import Fraction as F

import numpy as np
import pytest

import oracle_ffi as O


def test_interval_product_is_the_tightest_outward_hull():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a = np.sort(rng.standard_normal(2) * 2.0 ** rng.integers(-30, 30, 2))
        b = np.sort(rng.standard_normal(2) * 2.0 ** rng.integers(-30, 30, 2))
        if rng.integers(0, 4) == 0:
            a[0] = 0.0 if a[1] >= 0 else a[0]
        lo, hi = O.imul(a[0], a[1], b[0], b[1])
        prods = [F(x) * F(y) for x in a for y in b]
        pmin, pmax = min(prods), max(prods)
        assert F(lo) <= pmin < F(math.nextafter(lo, math.inf))
        assert F(math.nextafter(hi, -math.inf)) < pmax <= F(hi)


def _ivl(x, rad=0.0):
    P = np.zeros(x.shape + (2,))
    P[..., 0] = x - rad
    P[..., 1] = x + rad
    return P


@pytest.mark.parametrize("m,npts", [(1, 0), (1, 5), (6, 11)])
def test_uhphiphi_contains_exact_sums_and_is_symmetric(m, npts):
    rng = np.random.default_rng(m + npts)
    phi = O.pairs(m, npts)
    phi[...] = _ivl(rng.uniform(-1, 1, (m, npts)))
    w, u = _ivl(rng.uniform(0, 1, npts)), _ivl(rng.uniform(-2, 2, npts))
    A = O.uhphiphi(phi, w, u)
    assert np.array_equal(A, A.transpose(1, 0, 2))
    for i in range(m):
        for j in range(m):
            exact = sum(F(w[k, 0]) * F(u[k, 0]) * F(phi[i, k, 0]) * F(phi[j, k, 0]) for k in range(npts))
            assert F(A[i, j, 0]) <= exact <= F(A[i, j, 1])
            scale = sum(abs(F(w[k, 0]) * F(u[k, 0]) * F(phi[i, k, 0]) * F(phi[j, k, 0])) for k in range(npts))
            assert F(A[i, j, 1]) - F(A[i, j, 0]) <= F(4 * (npts + 2)) * F(2) ** -52 * scale


def test_uhphiphi_contains_vertex_samples_of_interval_data():
    rng = np.random.default_rng(7)
    m, npts = 4, 9
    phi = O.pairs(m, npts)
    phi[...] = _ivl(rng.uniform(-1, 1, (m, npts)), 2.0 ** -20)
    w, u = _ivl(rng.uniform(0, 1, npts), 2.0 ** -30), _ivl(rng.uniform(-2, 2, npts), 2.0 ** -25)
    A = O.uhphiphi(phi, w, u)
    for _ in range(20):
        pv = np.where(rng.integers(0, 2, (m, npts)) == 1, phi[..., 1], phi[..., 0])
        wv = np.where(rng.integers(0, 2, npts) == 1, w[:, 1], w[:, 0])
        uv = np.where(rng.integers(0, 2, npts) == 1, u[:, 1], u[:, 0])
        for i in range(m):
            for j in range(m):
                # phi(i, k) and phi(j, k) are the same interval when i == j: one vertex for both
                exact = sum(F(wv[k]) * F(uv[k]) * F(pv[i, k]) * F(pv[j, k]) for k in range(npts))
                assert F(A[i, j, 0]) <= exact <= F(A[i, j, 1])
