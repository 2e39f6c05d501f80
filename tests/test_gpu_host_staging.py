"""The host-pointer entry points stage copies above 8 MiB through pinned
buffers, a block of columns at a time.  Checked with leading dimensions larger
than the row count (strided per-column copies in both directions), against the
same product from device pointers, bit for bit, and with the padding of the
output left untouched."""
import ctypes

import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _padded(rng, rows, cols, ld):
    X = O.pairs(ld, cols)
    X[...] = 7.0
    X[:rows] = W.wide_exponent_intervals(rng, rows, cols)
    return X


@pytest.mark.parametrize("form", [vcp_amd.FORM_PLAIN, vcp_amd.FORM_OZAKI])
def test_host_staging_strided(ctx, form):
    rng = np.random.default_rng(81)
    m, k, n = 1500, 1100, 1300          # A 26 MB, B 23 MB, C 31 MB: all above the staging threshold
    lda, ldb, ldc = 1530, 1117, 1509
    A, B = _padded(rng, m, k, lda), _padded(rng, k, n, ldb)
    C = O.pairs(ldc, n)
    C[...] = 5.0
    L = vcp_amd.lib()
    rc = L.vcphip_igemm(ctx._h, form, m, n, k, A.ctypes.data, lda, B.ctypes.data, ldb, C.ctypes.data, ldc)
    assert rc == 0
    # the same product from device pointers
    dA = torch.from_numpy(np.ascontiguousarray(A[:m].transpose(1, 0, 2))).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(B[:k].transpose(1, 0, 2))).cuda()
    dC = torch.empty((n, m, 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.igemm_dev(form, m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, dC.data_ptr(), m)
    ctx.synchronize()
    assert np.array_equal(C[:m], dC.cpu().numpy().transpose(1, 0, 2))
    assert (C[m:] == 5.0).all()   # the rows between m and ldc are not written


def test_host_staging_dense_repeated(ctx):
    """dense operands (one contiguous block per copy), two calls in a row so
    the pinned buffers are reused"""
    rng = np.random.default_rng(82)
    m = k = n = 1200
    A, B = W.wide_exponent_intervals(rng, m, k), W.wide_exponent_intervals(rng, k, n)
    C1 = ctx.igemm(A, B, vcp_amd.FORM_PLAIN)
    A2 = W.wide_exponent_intervals(rng, m, k)
    C2 = ctx.igemm(A2, B, vcp_amd.FORM_PLAIN)
    rows = [0, 777]

    def sub(X):
        P = O.pairs(len(rows), X.shape[1])
        P[...] = X[rows]
        return P

    assert np.array_equal(C1[rows], O.pidblas_mulmm(sub(A), B))
    assert np.array_equal(C2[rows], O.pidblas_mulmm(sub(A2), B))
