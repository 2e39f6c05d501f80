"""The single plain products read their operands into padded planes once
(midpoint and radius of an interval operand, the point operand itself) and
then load whole k-slices of the planes straight into LDS.  These tests give
every form operands whose leading dimensions exceed their rows, with NaN in
the rows between, and outputs whose padding rows hold a sentinel: the
results must be the oracle's restatements bit for bit (pidblas::mulmm,
mul_im_m, mul_m_im, mulltmm, vmulmm, rmatmul / udmatmul; vcp/pidblas.hpp,
vblas/rmatmul.hpp:100-127, vblas/udmatmul.hpp:738-760), the NaN never read
and the sentinel never written.  k = 0 is included, and both tile heights
(64 and 32 output rows) of the plane-fed kernel."""
import ctypes

import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu

SHAPES = [(70, 45, 33), (129, 3, 200), (5, 130, 1), (64, 64, 16), (40, 20, 0)]
PAD = 7
SENTINEL = 5.0


def _dev_iv(X, ld):
    rows, cols = X.shape[:2]
    buf = np.full((cols, ld, 2), np.nan)
    buf[:, :rows] = X.transpose(1, 0, 2)
    return torch.from_numpy(buf).cuda()


def _dev_pt(X, ld):
    rows, cols = X.shape
    buf = np.full((cols, ld), np.nan)
    buf[:, :rows] = X.T
    return torch.from_numpy(buf).cuda()


def _out_iv(rows, cols, ld):
    return torch.full((cols, ld, 2), SENTINEL, dtype=torch.float64, device="cuda")


def _read_iv(t, rows):
    a = t.cpu().numpy()
    assert (a[:, rows:] == SENTINEL).all()
    return a[:, :rows].transpose(1, 0, 2)


def _wide(rng, r, c):
    return np.asfortranarray(rng.uniform(-1, 1, (r, c)) * 2.0 ** rng.integers(-30, 30, (r, c)))


def _ok(rc, name):
    assert rc == 0, "%s returned %d" % (name, rc)


@pytest.fixture(params=[64, 32])
def plane_rows(ctx, request):
    """Both tile heights of the plane-fed products (the automatic choice
    takes 32 rows for grids this small)."""
    L = vcp_amd.lib()
    assert L.vcphip_ctx_set_param(ctx._h, vcp_amd.PARAM_PLANE_ROWS, request.param) == 0
    yield request.param
    assert L.vcphip_ctx_set_param(ctx._h, vcp_amd.PARAM_PLANE_ROWS, 0) == 0


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_planes_padded_leading_dimensions(ctx, plane_rows, m, n, k):
    print("planes rows=%d m=%d n=%d k=%d" % (plane_rows, m, n, k), flush=True)
    L = vcp_amd.lib()
    rng = np.random.default_rng(m * 7919 + n * 31 + k)
    lda, ldb, ldc = m + PAD, k + PAD, m + PAD
    IA = W.wide_exponent_intervals(rng, m, k, rad_exp=-20)
    IB = W.wide_exponent_intervals(rng, k, n, rad_exp=-20)
    dIA, dIB = _dev_iv(IA, lda), _dev_iv(IB, ldb)
    # interval x interval
    C = _out_iv(m, n, ldc)
    _ok(L.vcphip_igemm_dev(ctx._h, vcp_amd.FORM_PLAIN, m, n, k, dIA.data_ptr(), lda, dIB.data_ptr(), ldb,
                           C.data_ptr(), ldc), "igemm_dev")
    ctx.synchronize()
    assert np.array_equal(_read_iv(C, m), O.pidblas_mulmm(IA, IB))
    # interval x point
    B = _wide(rng, k, n)
    dB = _dev_pt(B, ldb)
    C = _out_iv(m, n, ldc)
    _ok(L.vcphip_igemm_ip_dev(ctx._h, m, n, k, dIA.data_ptr(), lda, dB.data_ptr(), ldb, C.data_ptr(), ldc),
        "igemm_ip_dev")
    ctx.synchronize()
    assert np.array_equal(_read_iv(C, m), O.pidblas_mul_im_m(IA, B))
    # point x interval
    Bl = _wide(rng, m, k)
    dBl = _dev_pt(Bl, lda)
    C = _out_iv(m, n, ldc)
    _ok(L.vcphip_igemm_pi_dev(ctx._h, m, n, k, dBl.data_ptr(), lda, dIB.data_ptr(), ldb, C.data_ptr(), ldc),
        "igemm_pi_dev")
    ctx.synchronize()
    assert np.array_equal(_read_iv(C, m), O.pidblas_mul_m_im(Bl, IB))
    # point RU / RD, accumulating onto existing values (udmatmul)
    A = _wide(rng, m, k)
    dA = _dev_pt(A, lda)
    U0, D0 = _wide(rng, m, n), _wide(rng, m, n)
    dU, dD = _dev_pt(U0, ldc), _dev_pt(D0, ldc)
    dU[:, m:] = SENTINEL
    dD[:, m:] = SENTINEL
    _ok(L.vcphip_gemm_rurd_dev(ctx._h, m, n, k, dA.data_ptr(), lda, dB.data_ptr(), ldb, dU.data_ptr(),
                               dD.data_ptr(), ldc, 1), "gemm_rurd_dev")
    ctx.synchronize()
    u, d = dU.cpu().numpy(), dD.cpu().numpy()
    assert (u[:, m:] == SENTINEL).all() and (d[:, m:] == SENTINEL).all()
    assert np.array_equal(u[:, :m].T, O.rgemm(A, B, 1, C=np.asfortranarray(U0.copy()), accumulate=True))
    assert np.array_equal(d[:, :m].T, O.rgemm(A, B, -1, C=np.asfortranarray(D0.copy()), accumulate=True))


@pytest.mark.parametrize("k,n", [(33, 70), (200, 129), (1, 5), (16, 64), (0, 20)])
def test_planes_symmetric_forms_padded(ctx, k, n):
    print("planes A^T A k=%d n=%d" % (k, n), flush=True)
    L = vcp_amd.lib()
    rng = np.random.default_rng(k * 131 + n)
    lda, ldc = k + PAD, n + PAD
    IA = W.wide_exponent_intervals(rng, k, n, rad_exp=-20)
    C = _out_iv(n, n, ldc)
    _ok(L.vcphip_igemm_tn_dev(ctx._h, k, n, _dev_iv(IA, lda).data_ptr(), lda, C.data_ptr(), ldc), "igemm_tn_dev")
    ctx.synchronize()
    assert np.array_equal(_read_iv(C, n), O.pidblas_mulltmm(IA))
    A = _wide(rng, k, n)
    C = _out_iv(n, n, ldc)
    _ok(L.vcphip_vgemm_tn_dev(ctx._h, k, n, _dev_pt(A, lda).data_ptr(), lda, C.data_ptr(), ldc), "vgemm_tn_dev")
    ctx.synchronize()
    assert np.array_equal(_read_iv(C, n), O.pidblas_vmulmm(A))


def test_plane_rows_parameter(ctx):
    L = vcp_amd.lib()
    v = ctypes.c_longlong(-1)
    assert L.vcphip_ctx_get_param(ctx._h, vcp_amd.PARAM_PLANE_ROWS, ctypes.byref(v)) == 0 and v.value == 0
    for bad in (16, 48, 128, -1):
        assert L.vcphip_ctx_set_param(ctx._h, vcp_amd.PARAM_PLANE_ROWS, bad) == -2
    assert L.vcphip_ctx_get_param(ctx._h, vcp_amd.PARAM_PLANE_ROWS, ctypes.byref(v)) == 0 and v.value == 0
