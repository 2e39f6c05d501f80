"""Parity of the inf-sup interval product (VCPHIP_FORM_INFSUP, the product of
pidblas_fma::mulmm, vcp/pidblas_fma.hpp:45-72) with the oracle's restatement
of the reference's sign-case interval FMA: equal values element by element
(np.array_equal, which does not distinguish the sign of a zero end: the
reference's two ifma branches do not agree on it either), and containment
of the exact interval product."""
from fractions import Fraction as F

import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _mixed(rng, r, c):
    """intervals of every sign class, with zero ends"""
    lo = rng.uniform(-1, 1, (r, c)) * 2.0 ** rng.integers(-4, 5, (r, c))
    hi = lo + rng.uniform(0, 1, (r, c)) * 2.0 ** rng.integers(-6, 2, (r, c))
    X = W._pairs(r, c)
    X[..., 0], X[..., 1] = lo, hi
    if r > 3:
        X[0::7, :, 0] = 0.0
        X[0::7, :, 1] = np.abs(X[0::7, :, 1])
        X[1::7, :, 1] = 0.0
        X[1::7, :, 0] = -np.abs(X[1::7, :, 0])
        X[2::7, :, 1] = X[2::7, :, 0]     # point intervals
    assert (X[..., 0] <= X[..., 1]).all()
    return X


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (64, 64, 64), (65, 33, 17), (100, 37, 333), (7, 130, 5),
                                   (256, 128, 1000), (50, 60, 0)])
def test_infsup_bitexact_against_oracle(ctx, m, n, k):
    rng = np.random.default_rng(m * 1000 + n + k)
    IA, IB = _mixed(rng, m, k), _mixed(rng, k, n)
    C = ctx.igemm(IA, IB, vcp_amd.FORM_INFSUP)
    assert np.array_equal(C, O.pidblas_fma_mulmm(IA, IB))


def test_infsup_wide_exponents(ctx):
    rng = np.random.default_rng(3)
    IA, IB = W.wide_exponent_intervals(rng, 200, 300), W.wide_exponent_intervals(rng, 300, 150)
    C = ctx.igemm(IA, IB, vcp_amd.FORM_INFSUP)
    assert np.array_equal(C, O.pidblas_fma_mulmm(IA, IB))


def test_infsup_dev_entry_and_leading_dimensions(ctx):
    rng = np.random.default_rng(4)
    m, n, k = 90, 70, 110
    IA, IB = _mixed(rng, m, k), _mixed(rng, k, n)
    # padded leading dimensions on the device
    lda, ldb, ldc = m + 3, k + 5, m + 1
    dA = torch.zeros((k, lda, 2), dtype=torch.float64, device="cuda")
    dA[:, :m, :] = torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2)))
    dB = torch.zeros((n, ldb, 2), dtype=torch.float64, device="cuda")
    dB[:, :k, :] = torch.from_numpy(np.ascontiguousarray(IB.transpose(1, 0, 2)))
    dC = torch.full((n, ldc, 2), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.igemm_dev(vcp_amd.FORM_INFSUP, m, n, k, dA.data_ptr(), lda, dB.data_ptr(), ldb, dC.data_ptr(), ldc)
    ctx.synchronize()
    C = dC.cpu().numpy()[:, :m, :].transpose(1, 0, 2)
    assert np.array_equal(C, O.pidblas_fma_mulmm(IA, IB))
    assert np.isnan(dC.cpu().numpy()[:, m:, :]).all()   # the padding rows are not written


def test_infsup_contains_exact_interval_product(ctx):
    rng = np.random.default_rng(5)
    m, n, k = 12, 9, 40
    IA, IB = _mixed(rng, m, k), _mixed(rng, k, n)
    C = ctx.igemm(IA, IB, vcp_amd.FORM_INFSUP)
    for i in range(m):
        for j in range(n):
            lo = hi = F(0)
            for t in range(k):
                p = [F(a) * F(b) for a in IA[i, t] for b in IB[t, j]]
                lo += min(p)
                hi += max(p)
            assert F(C[i, j, 0]) <= lo and hi <= F(C[i, j, 1])


def test_infsup_not_wider_than_midrad_on_sign_definite_data(ctx):
    """for intervals away from zero the inf-sup product is the tighter one"""
    rng = np.random.default_rng(6)
    m, n, k = 64, 48, 256
    IA = W.interval_from_midrad(rng.uniform(1, 2, (m, k)), np.ldexp(rng.uniform(1, 2, (m, k)), -10))
    IB = W.interval_from_midrad(rng.uniform(1, 2, (k, n)), np.ldexp(rng.uniform(1, 2, (k, n)), -10))
    Ci = ctx.igemm(IA, IB, vcp_amd.FORM_INFSUP)
    Cp = ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN)
    assert ((Ci[..., 1] - Ci[..., 0]) <= (Cp[..., 1] - Cp[..., 0])).all()


@pytest.mark.parametrize("m,n,k", [(96, 80, 80), (130, 64, 200)])
def test_infsup_two_fma_and_general_slices(ctx, m, n, k):
    """k-slices whose operands all lie strictly on one side of zero take two
    FMAs per chain, the others four: sign-definite data with intervals
    across or touching zero in k = 16..19 (A) and k = 64..66 (B), so both
    loops run in one product, against the oracle."""
    print("infsup mixed slices", m, n, k, flush=True)
    rng = np.random.default_rng(m + n + k)
    IA, IB = W.wide_exponent_intervals(rng, m, k), W.wide_exponent_intervals(rng, k, n)
    IA[:, 16:20, 0] = -np.abs(IA[:, 16:20, 1])          # across zero
    IB[64:66, :, 0] = 0.0                                # touching zero
    IB[64:66, :, 1] = np.abs(IB[64:66, :, 1])
    IB[66, ::2, :] = 0.0                                 # [0, 0]
    C = ctx.igemm(IA, IB, vcp_amd.FORM_INFSUP)
    assert np.array_equal(C, O.pidblas_fma_mulmm(IA, IB))
