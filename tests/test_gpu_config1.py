"""The plain interval product at n = 1024 (mid U[-1, 1], radius 2^-40 |mid|,
seeded): the configuration of the reference's own timing demo of this path
(test_matrix/test_interval_matrix_mul.cpp:46-72, vcp::matrix<interval,
pidblas> A*B).  The plain form is bit-identical to pidblas::mulmm built on
the reference's own vblas rmatmul (oracle/_ref; the oracle's restatement
where that build is absent), every one of 4096 sampled entries of both forms
contains the exact product hull, and the Ozaki radii stay within 2x of the
reference's."""
import numpy as np
import pytest

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_config1_interval_product_n1024(ctx):
    n = 1024
    IA, IB = W.config1(n)
    print("config 1: n=%d plain and Ozaki products" % n, flush=True)
    P = ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN)
    Z = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    ref = O.ref_pidblas_mulmm(IA, IB)
    if ref is None:
        ref = O.pidblas_mulmm(IA, IB)
    assert np.array_equal(P, ref)
    rng = np.random.default_rng(W.SEED0 + 1)
    ii = rng.integers(0, n, 4096)
    jj = rng.integers(0, n, 4096)
    for C in (P, Z):
        ok, _, _ = O.check_containment(IA, IB, C, ii, jj)
        assert ok.all(), "%d of 4096 sampled entries not enclosed" % (~ok).sum()
    ratio = (Z[..., 1] - Z[..., 0]) / np.maximum(ref[..., 1] - ref[..., 0], np.finfo(float).tiny)
    print("radius ratio Ozaki / reference: max %.4f median %.4f" % (ratio.max(), np.median(ratio)), flush=True)
    assert ratio.max() <= 2.0
    # the Ozaki form is the oracle's restatement bit for bit (rows sampled)
    rows = np.array([0, 1, 511, 1023])
    assert np.array_equal(Z[rows], O.ozaki_igemm(np.ascontiguousarray(IA[rows].transpose(1, 0, 2)).transpose(1, 0, 2),
                                                 IB, 16))
