"""The benchmark workloads at their full sizes, checked through properties that
do not need the whole product on the CPU:

- interval GEMM n=8192 (wide exponents): exact-hull containment of sampled
  entries for the Ozaki-II and the plain form; the Ozaki radii within 2x of
  the plain form's over all 67M entries; two rows of each form bit-identical to
  the oracle's restatement (row scaling and the k-chains depend only on the
  row and on the whole of B, so a row subset is the full product's rows); the
  same for the Ozaki configuration the moduli tuner picks;
- dd GEMM n=4096: error of sampled entries against the exact dot product;
- RU/RD point GEMM n=4096: two rows bit-identical to the oracle, CD <= CU;
- batched R*A - I over the 256 problems of the batched workload: every bound
  below 1, the smallest problems bit-identical to the oracle.
"""
import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dev(X):
    return torch.from_numpy(np.ascontiguousarray(X.transpose(1, 0, 2))).cuda()


def _rows(X, rows):
    P = O.pairs(len(rows), X.shape[1])
    P[...] = X[rows]
    return P


@pytest.fixture(scope="module")
def config3():
    print("generating the n=8192 interval workload", flush=True)
    IA, IB = W.config3(8192)
    return IA, IB


def test_interval_gemm_n8192(ctx, config3):
    IA, IB = config3
    n = IA.shape[0]
    dA, dB = _dev(IA), _dev(IB)
    dC = torch.empty((n, n, 2), dtype=torch.float64, device="cuda")
    dP = torch.empty((n, n, 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    print("ozaki and plain products, n=8192", flush=True)
    ctx.igemm_dev(vcp_amd.FORM_OZAKI, n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dC.data_ptr(), n)
    ctx.igemm_dev(vcp_amd.FORM_PLAIN, n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dP.data_ptr(), n)
    ctx.synchronize()
    # radii over all entries, on the GPU
    wo = dC[..., 1] - dC[..., 0]
    wp = dP[..., 1] - dP[..., 0]
    assert bool(torch.isfinite(wo).all()) and bool((wo >= 0).all()) and bool((wp >= 0).all())
    ratio = wo / torch.clamp(wp, min=torch.finfo(torch.float64).tiny)
    print("radius ratio max %.4f median %.4f" % (float(ratio.max()), float(ratio.median())), flush=True)
    assert float(ratio.max()) <= 2.0
    del wo, wp, ratio
    C = dC.cpu().numpy().transpose(1, 0, 2)
    P = dP.cpu().numpy().transpose(1, 0, 2)
    del dA, dB, dC, dP
    # exact containment of sampled entries
    rng = np.random.default_rng(91)
    ii, jj = rng.integers(0, n, 128), rng.integers(0, n, 128)
    ok_o, _, _ = O.check_containment(IA, IB, C, ii, jj)
    ok_p, _, _ = O.check_containment(IA, IB, P, ii, jj)
    assert ok_o.all() and ok_p.all()
    # two rows bit-identical to the oracle
    rows = [0, 5001]
    print("oracle rows", rows, flush=True)
    IAr = _rows(IA, rows)
    assert np.array_equal(C[rows], O.ozaki_igemm(IAr, IB, 16))
    assert np.array_equal(P[rows], O.pidblas_mulmm(IAr, IB))


def test_interval_gemm_n8192_tuned_configuration(ctx, config3):
    """the configuration the moduli tuner picks for this workload at a radius
    ratio target of 2 (the benchmark's headline): every radius within 2x of
    the plain form's over all 67M entries, 256 sampled entries containing the
    exact hull, and two rows bit-identical to the oracle's restatement with
    the same moduli count and radius scheme"""
    IA, IB = config3
    n = IA.shape[0]
    dA, dB = _dev(IA), _dev(IB)
    dC = torch.empty((n, n, 2), dtype=torch.float64, device="cuda")
    dP = torch.empty((n, n, 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with vcp_amd.Context(0) as c2:
        nmod, radius, sampled = c2.tune_igemm_moduli_dev(n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, 2.0, 0)
        print("tuned: %d moduli, %s radius digits, sampled ratio %.4f"
              % (nmod, vcp_amd.OZ_RADIUS_NAMES[radius], sampled), flush=True)
        plan = c2.ozaki_plan(n, n, n)
        assert plan["moduli"] == nmod and plan["radius"] == radius
        c2.igemm_dev(vcp_amd.FORM_OZAKI, n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dC.data_ptr(), n)
        c2.igemm_dev(vcp_amd.FORM_PLAIN, n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dP.data_ptr(), n)
        c2.synchronize()
    wo = dC[..., 1] - dC[..., 0]
    wp = dP[..., 1] - dP[..., 0]
    ratio = wo / torch.clamp(wp, min=torch.finfo(torch.float64).tiny)
    print("radius ratio max %.4f median %.4f" % (float(ratio.max()), float(ratio.median())), flush=True)
    assert bool(torch.isfinite(wo).all()) and float(ratio.max()) <= 2.0
    del wo, wp, ratio, dP
    C = dC.cpu().numpy().transpose(1, 0, 2)
    del dA, dB, dC
    rng = np.random.default_rng(93)
    ii, jj = rng.integers(0, n, 256), rng.integers(0, n, 256)
    ok, _, _ = O.check_containment(IA, IB, C, ii, jj)
    assert ok.all(), "%d of 256 sampled entries not enclosed" % (~ok).sum()
    rows = [7, 4099]
    print("oracle rows", rows, flush=True)
    assert np.array_equal(C[rows], O.ozaki_igemm(_rows(IA, rows), IB, nmod, radius))


def test_dd_gemm_n4096(ctx):
    """the dd workload at n=4096 on 4096 sampled entries: the error against
    the exact product is within 2^-100 sum |a||b| (rigorous scale), and its
    size in dd ulps (2^-105 |c_ij|) is reported"""
    A, B = W.config4(4096)
    C = ctx.ddgemm(A, B)
    rng = np.random.default_rng(92)
    worst, ulps = 0.0, []
    absA, absB = np.abs(A[..., 0]), np.abs(B[..., 0])
    for i, j in zip(rng.integers(0, 4096, 4096), rng.integers(0, 4096, 4096)):
        h, _, e = O.exact_dd_dot(A[i], B[:, j], C[i, j, 0], C[i, j, 1])
        worst = max(worst, e / float(absA[i] @ absB[:, j]))
        ulps.append(e / (abs(h) * 2.0 ** -105))
    print("dd worst error / sum|a||b|: 2^%.1f; dd ulps max %.3f median %.3f"
          % (np.log2(max(worst, 2.0 ** -200)), max(ulps), float(np.median(ulps))), flush=True)
    assert worst <= 2.0 ** -100


def test_rurd_gemm_n4096(ctx):
    A, B = W.config2(4096)
    CU, CD = ctx.gemm_rurd(A, B)
    assert (CD <= CU).all()
    rows = [3, 4095]
    Ar = np.asfortranarray(A[rows])
    assert np.array_equal(CU[rows], O.rgemm(Ar, B, 1))
    assert np.array_equal(CD[rows], O.rgemm(Ar, B, -1))


def test_batched_residual_workload_256(ctx):
    print("generating the 256 residual problems", flush=True)
    probs = W.config5()
    Rs, IAs = [p[0] for p in probs], [p[1] for p in probs]
    Cs, bounds = ctx.residual_batched(Rs, IAs)
    # every bound below 1 (verified) and above 0: R A - I of a floating-point
    # inverse is never exactly zero, so a zero bound means nothing ran
    assert max(bounds) < 1.0 and min(bounds) > 0.0
    assert all(np.abs(C).max() > 0.0 for C in Cs)
    for i in sorted(range(len(probs)), key=lambda i: probs[i][0].shape[0])[:4]:
        Co, nbo = O.residual(Rs[i], IAs[i])
        assert np.array_equal(Cs[i], Co) and bounds[i] == nbo
