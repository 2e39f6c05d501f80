"""The double-double product (pddhip, the vcp::pddblas product) on int8
MFMA: bit-identical to the oracle's restatement of the Ozaki-II route,
accurate to about one dd ulp against the exact product, and at least as
accurate as the restatement of the reference's Ozaki-I pddblas::mulmm."""
import numpy as np
import pytest

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dd(rng, r, c, kind="uniform"):
    x = W.dd_matrix(rng, r, c)
    if kind == "wide":
        s = np.ldexp(1.0, rng.integers(-30, 30, (r, c)))
        x[..., 0] *= s
        x[..., 1] *= s
    if kind == "zeros":
        x[::4] = 0.0
        x[:, ::5] = 0.0
    if kind == "integers":   # exact products: sums with zero low limbs, lo = 0
        x[..., 0] = rng.integers(-8, 9, (r, c)).astype(np.float64)
        x[..., 1] = 0.0
    return x


@pytest.mark.parametrize("kind,sa,sb", [("integers", 0, 0), ("uniform", -540, -540), ("uniform", -600, -470),
                                        ("uniform", 500, 515), ("wide", -520, -530)])
def test_dd_bitexact_conversion_edges(ctx, kind, sa, sb):
    """exact integer sums, gradual underflow of hi and lo, results next to
    the overflow threshold: the reconstruction's roundings on every path"""
    rng = np.random.default_rng(hash((kind, sa, sb)) % 2 ** 32)
    A, B = _dd(rng, 40, 96, kind), _dd(rng, 96, 36, kind)
    A, B = np.ldexp(A, sa), np.ldexp(B, sb)
    C = ctx.ddgemm(A, B)
    beta_ok = [nm for nm in range(24, 33) if O.ozaki_dd_beta(96, nm) >= 110]
    R = O.ozaki_ddgemm(A, B, beta_ok[0])
    same = (C == R) | (np.isnan(C) & np.isnan(R))
    same &= np.signbit(C) == np.signbit(R)
    bad = np.argwhere(~same.all(axis=2))
    assert len(bad) == 0, "mismatch at %s: gpu %s oracle %s" % (bad[:3].tolist(), C[tuple(bad[0])], R[tuple(bad[0])])


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 64), (70, 33, 100), (256, 256, 256), (300, 257, 513)])
@pytest.mark.parametrize("kind", ["uniform", "wide", "zeros"])
def test_dd_bitexact(ctx, m, n, k, kind):
    rng = np.random.default_rng(hash((m, n, k, kind, "dd")) % 2 ** 32)
    A, B = _dd(rng, m, k, kind), _dd(rng, k, n, kind)
    nmod = vcp_amd.lib() and None
    C = ctx.ddgemm(A, B)
    beta_ok = [nm for nm in range(24, 33) if O.ozaki_dd_beta(k, nm) >= 110]
    R = O.ozaki_ddgemm(A, B, beta_ok[0])
    bad = np.argwhere(~((C == R).all(axis=2)))
    assert len(bad) == 0, "mismatch at %s: gpu %s oracle %s" % (bad[:3].tolist(), C[tuple(bad[0])], R[tuple(bad[0])])
    del nmod


@pytest.mark.parametrize("nmod", [24, 25, 26, 30, 32])
def test_dd_nmod_bitexact(ctx, nmod):
    rng = np.random.default_rng(nmod)
    A, B = _dd(rng, 64, 200), _dd(rng, 200, 48)
    ctx.set_param(vcp_amd.PARAM_DD_NMOD, nmod)
    try:
        C = ctx.ddgemm(A, B)
    finally:
        ctx.set_param(vcp_amd.PARAM_DD_NMOD, 0)
    assert np.array_equal(C, O.ozaki_ddgemm(A, B, nmod))


def test_dd_tn_symmetric_and_bitexact(ctx):
    rng = np.random.default_rng(3)
    A = _dd(rng, 300, 70)
    C = ctx.ddgemm_tn(A)
    assert np.array_equal(C, C.transpose(1, 0, 2))
    beta_ok = [nm for nm in range(24, 33) if O.ozaki_dd_beta(300, nm) >= 110]
    assert np.array_equal(C, O.ozaki_ddgemm(A, None, beta_ok[0]))


def test_dd_accuracy_against_exact_and_pddblas(ctx):
    rng = np.random.default_rng(4)
    m, n, k = 40, 32, 1024
    A, B = _dd(rng, m, k), _dd(rng, k, n)
    C = ctx.ddgemm(A, B)
    P, _, _ = O.pddblas_mulmm(A, B)
    eo, ep = [], []
    for i in range(m):
        for j in range(n):
            h, _, e = O.exact_dd_dot(A[i], B[:, j], C[i, j, 0], C[i, j, 1])
            _, _, e2 = O.exact_dd_dot(A[i], B[:, j], P[i, j, 0], P[i, j, 1])
            scale = float(np.sum(np.abs(A[i, :, 0]) * np.abs(B[:, j, 0])))
            eo.append(e / (scale * 2.0 ** -105))
            ep.append(e2 / (scale * 2.0 ** -105))
    # errors relative to sum |a||b| in units of 2^-105: both well below one
    assert max(eo) <= 1.0 and max(eo) <= max(2.0 * max(ep), 0.05)


def test_dd_nonfinite_operands_match_oracle(ctx):
    """a row / column with an infinite or NaN part gives NaN entries, as the
    oracle; the rest bit-exact"""
    rng = np.random.default_rng(42)
    A, B = W.dd_matrix(rng, 130, 260), W.dd_matrix(rng, 260, 70)
    A[3, 100, 0] = np.inf
    A[90, 1, 1] = np.nan
    B[259, 40, 0] = -np.inf
    C = ctx.ddgemm(A, B)
    nmod = next(t for t in range(24, 33) if O.ozaki_dd_beta(260, t) >= 110)
    assert np.array_equal(C, O.ozaki_ddgemm(A, B, nmod), equal_nan=True)
    assert np.isnan(C[3, 0]).all() and np.isnan(C[0, 40]).all() and np.isfinite(C[0, 0]).all()
    T = W.dd_matrix(rng, 200, 60)
    T[10, 7, 1] = np.inf
    S = ctx.ddgemm_tn(T)
    nmod = next(t for t in range(24, 33) if O.ozaki_dd_beta(200, t) >= 110)
    assert np.array_equal(S, O.ozaki_ddgemm(T, None, nmod), equal_nan=True)
