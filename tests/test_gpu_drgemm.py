"""Parity of the plain FP64 directed-rounding kernels (HIP, through the C
ABI) with the oracle's restatement of pidblas / rmatmul: bit-identical,
since both evaluate the same k-ascending rounded-FMA chains."""
import numpy as np
import pytest

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (5, 3, 7), (64, 64, 64), (65, 63, 17), (130, 70, 257), (33, 200, 1), (200, 33, 0)]


def _maps_has(path):
    with open("/proc/self/maps") as f:
        return any(path in line for line in f)


def test_native_library_loaded(ctx):
    assert _maps_has(vcp_amd.LIB_PATH)


def _wide(rng, r, c):
    return np.asfortranarray(rng.uniform(-1, 1, (r, c)) * 2.0 ** rng.integers(-30, 30, (r, c)))


@pytest.mark.parametrize("n", [64, 300])
def test_rounding_kat_one_third(ctx, n):
    # Check_pidblas_rounding.cpp:14-37 / Check_rmatmul_rounding.cpp:53-79
    for i in [0, 1, n // 2, n - 1]:
        A = np.zeros((n, n), order="F")
        B = np.zeros((n, n), order="F")
        A[:, i] = 1.0 / 3.0
        B[i, :] = 1.0 / 3.0
        CU, CD = ctx.gemm_rurd(A, B)
        assert (CU > CD).all()
        assert np.array_equal(CU, O.rgemm(A, B, 1)) and np.array_equal(CD, O.rgemm(A, B, -1))


@pytest.mark.parametrize("n", [64, 300])
def test_rounding_kat_one_plus_2m54(ctx, n):
    # Check_pidblas_rounding.cpp:41-71: exact value 1 + 2^-54
    for i in [2, n // 3, n - 1]:
        A = np.zeros((n, n), order="F")
        B = np.zeros((n, n), order="F")
        A[:, 1] = 1.0
        B[1, :] = 1.0
        A[:, i] = 2.0 ** -27
        B[i, :] = 2.0 ** -27
        CU, CD = ctx.gemm_rurd(A, B)
        assert (CU == 1.0 + 2.0 ** -52).all()
        assert (CD == 1.0).all()


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_rurd_bitexact(ctx, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A, B = _wide(rng, m, k), _wide(rng, k, n)
    CU, CD = ctx.gemm_rurd(A, B)
    assert np.array_equal(CU, O.rgemm(A, B, 1))
    assert np.array_equal(CD, O.rgemm(A, B, -1))
    if O.ref_lib() is not None and k > 0:
        assert np.array_equal(CU, O.ref_rmatmul(A, B, 1))
        assert np.array_equal(CD, O.ref_rmatmul(A, B, -1))


def test_rurd_accumulate(ctx):
    rng = np.random.default_rng(7)
    A, B = _wide(rng, 70, 90), _wide(rng, 90, 50)
    C0 = _wide(rng, 70, 50)
    CU, CD = ctx.gemm_rurd(A, B, Cup=C0.copy(order="F"), Cdn=C0.copy(order="F"), accumulate=True)
    assert np.array_equal(CU, O.rgemm(A, B, 1, C=C0.copy(order="F"), accumulate=True))
    assert np.array_equal(CD, O.rgemm(A, B, -1, C=C0.copy(order="F"), accumulate=True))


def _iv(rng, r, c, kind):
    if kind == "uniform":
        return W.uniform_intervals(rng, r, c, -40)
    if kind == "wide":
        return W.wide_exponent_intervals(rng, r, c)
    if kind == "point":
        x = W.uniform_intervals(rng, r, c, -40)
        x[..., 1] = x[..., 0]
        return x
    if kind == "straddle":   # intervals containing zero
        mid = rng.uniform(-1, 1, (r, c))
        return W.interval_from_midrad(mid, np.abs(rng.uniform(-1, 1, (r, c))))
    raise ValueError(kind)


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("kind", ["uniform", "wide", "point", "straddle"])
def test_igemm_plain_bitexact(ctx, m, n, k, kind):
    rng = np.random.default_rng(hash((m, n, k, kind)) % 2 ** 32)
    IA, IB = _iv(rng, m, k, kind), _iv(rng, k, n, kind)
    C = ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN)
    R = O.pidblas_mulmm(IA, IB)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_igemm_ip_pi_bitexact(ctx, m, n, k):
    rng = np.random.default_rng(m + 7 * n + 13 * k)
    IA = _iv(rng, m, k, "wide")
    B = _wide(rng, k, n)
    assert np.array_equal(ctx.igemm_ip(IA, B), O.pidblas_mul_im_m(IA, B))
    Bl = _wide(rng, m, k)
    IAr = _iv(rng, k, n, "wide")
    assert np.array_equal(ctx.igemm_pi(Bl, IAr), O.pidblas_mul_m_im(Bl, IAr))


@pytest.mark.parametrize("k,n", [(1, 1), (7, 5), (64, 64), (100, 130), (257, 70)])
def test_igemm_tn_vgemm_tn_bitexact(ctx, k, n):
    rng = np.random.default_rng(k * 31 + n)
    IA = _iv(rng, k, n, "wide")
    assert np.array_equal(ctx.igemm_tn(IA), O.pidblas_mulltmm(IA))
    A = _wide(rng, k, n)
    assert np.array_equal(ctx.vgemm_tn(A), O.pidblas_vmulmm(A))


def test_igemm_plain_containment(ctx):
    rng = np.random.default_rng(11)
    m, n, k = 96, 80, 300
    IA, IB = _iv(rng, m, k, "wide"), _iv(rng, k, n, "straddle")
    C = ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ok, _, _ = O.check_containment(IA, IB, C, ii.ravel(), jj.ravel())
    assert ok.all()
