"""A seeded sweep of random shapes and exponent ranges through every product
form, each result bit-identical to the oracle's restatement:

- interval x interval, plain and Ozaki-II (pidblas::mulmm);
- interval x point and point x interval (mul_im_m, mul_m_im);
- IA^T IA (mulltmm);
- the RU/RD point product (rmatmul);
- the dd product (pddblas::mulmm via the Ozaki-II dd route).

Shapes run from 1 to 300 in each dimension, including sizes that are not
multiples of any tile; exponent ranges run from narrow to 2^+-60."""
import numpy as np
import pytest

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(20261202)
CASES = [(int(RNG.integers(1, 301)), int(RNG.integers(1, 301)), int(RNG.integers(1, 301)),
          int(RNG.integers(0, 61)), int(RNG.integers(0, 2 ** 31))) for _ in range(12)]


def _iv(rng, r, c, espan):
    return W.wide_exponent_intervals(rng, r, c, emin=-espan, emax=espan)


@pytest.mark.parametrize("m,n,k,espan,seed", CASES)
def test_interval_forms(ctx, m, n, k, espan, seed):
    rng = np.random.default_rng(seed)
    IA, IB = _iv(rng, m, k, espan), _iv(rng, k, n, espan)
    assert np.array_equal(ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN), O.pidblas_mulmm(IA, IB))
    assert np.array_equal(ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI), O.ozaki_igemm(IA, IB, 16))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * 2.0 ** rng.integers(-espan, espan + 1, (k, n)))
    assert np.array_equal(ctx.igemm_ip(IA, B), O.pidblas_mul_im_m(IA, B))
    Bl = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * 2.0 ** rng.integers(-espan, espan + 1, (m, k)))
    assert np.array_equal(ctx.igemm_pi(Bl, IB), O.pidblas_mul_m_im(Bl, IB))
    assert np.array_equal(ctx.igemm_tn(IA), O.pidblas_mulltmm(IA))


@pytest.mark.parametrize("m,n,k,espan,seed", CASES[:6])
def test_point_and_dd_forms(ctx, m, n, k, espan, seed):
    rng = np.random.default_rng(seed + 1)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * 2.0 ** rng.integers(-espan, espan + 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * 2.0 ** rng.integers(-espan, espan + 1, (k, n)))
    CU, CD = ctx.gemm_rurd(A, B)
    assert np.array_equal(CU, O.rgemm(A, B, 1)) and np.array_equal(CD, O.rgemm(A, B, -1))
    Ad, Bd = W.dd_matrix(rng, m, k), W.dd_matrix(rng, k, n)
    assert np.array_equal(ctx.ddgemm(Ad, Bd), O.ozaki_ddgemm(Ad, Bd, O.dd_default_nmod(k)))
