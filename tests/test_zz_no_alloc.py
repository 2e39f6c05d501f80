"""vcphip_igemm_ozaki_noalloc_dev: the Ozaki interval product with its
workspace in the library's fixed device scratch, called through the C ABI.
Two calls in a row, the second smaller (the scratch then holds the first
call's contents), each bit-identical to vcphip_igemm_dev(VCPHIP_FORM_OZAKI)
on the same operands and to the oracle on sampled rows; a shape whose
workspace does not fit is refused with VCPHIP_ERR_SCRATCH before anything
is enqueued (its output buffer stays untouched)."""
import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

import ozaki_cases

pytestmark = pytest.mark.gpu


def _dev(X):
    """column-major AoS pairs on cuda:0 (the (cols, rows, 2) C-contiguous buffer)"""
    return torch.from_numpy(np.ascontiguousarray(X.transpose(1, 0, 2))).cuda()


def _rows(X, rows):
    return np.ascontiguousarray(X[rows].transpose(1, 0, 2)).transpose(1, 0, 2)


def test_noalloc_matches_the_allocating_entry_point():
    rng = np.random.default_rng(2718)
    cases = [(W.wide_exponent_intervals(rng, 700, 900), W.wide_exponent_intervals(rng, 900, 300)),
             ozaki_cases.make("wide_r10_s7")]   # 160 x 160 x 160 after 700 x 300 x 900
    with vcp_amd.Context(0) as ctx:
        for IA, IB in cases:
            m, k = IA.shape[:2]
            n = IB.shape[1]
            print("noalloc %d x %d x %d" % (m, n, k), flush=True)
            dA, dB = _dev(IA), _dev(IB)
            dN = torch.full((n, m, 2), float("nan"), dtype=torch.float64, device="cuda")
            dR = torch.full((n, m, 2), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            ctx.igemm_ozaki_noalloc_dev(m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, dN.data_ptr(), m)
            ctx.igemm_dev(vcp_amd.FORM_OZAKI, m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, dR.data_ptr(), m)
            ctx.synchronize()
            assert torch.equal(dN, dR)
            C = dN.cpu().numpy().transpose(1, 0, 2)
            rows = [0, m // 2, m - 1]
            assert np.array_equal(C[rows], O.ozaki_igemm(_rows(IA, rows), IB, 16))


def test_noalloc_refuses_a_shape_beyond_its_scratch():
    m = n = k = 2048   # planes alone need more than VCPHIP_NOALLOC_SCRATCH_BYTES
    with vcp_amd.Context(0) as ctx:
        dA = torch.zeros((k, m, 2), dtype=torch.float64, device="cuda")
        dB = torch.zeros((n, k, 2), dtype=torch.float64, device="cuda")
        dC = torch.full((n, m, 2), 7.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        with pytest.raises(vcp_amd.VcphipError) as e:
            ctx.igemm_ozaki_noalloc_dev(m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, dC.data_ptr(), m)
        assert e.value.code == -11
        ctx.synchronize()
        assert bool((dC == 7.0).all())
