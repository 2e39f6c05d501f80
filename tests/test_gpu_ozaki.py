"""Parity of the Ozaki-II int8-MFMA interval GEMM (HIP, through the C ABI)
with the oracle's step-by-step restatement: the exact scaled-integer
product rebuilt from the residues must give bit-identical enclosures.
Containment is checked against the exact hull; radii against the plain
FP64 directed-rounding product."""
import numpy as np
import pytest

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

import ozaki_cases

pytestmark = pytest.mark.gpu


def _iv(rng, r, c, kind):
    if kind == "uniform":
        return W.uniform_intervals(rng, r, c, -40)
    if kind == "wide":
        return W.wide_exponent_intervals(rng, r, c)
    if kind == "point":
        x = W.uniform_intervals(rng, r, c, -40)
        x[..., 1] = x[..., 0]
        return x
    if kind == "zerorows":
        x = W.wide_exponent_intervals(rng, r, c)
        x[::3, :, :] = 0.0
        x[:, ::4, :] = 0.0
        return x
    if kind == "straddle":
        mid = rng.uniform(-1, 1, (r, c))
        return W.interval_from_midrad(mid, np.abs(rng.uniform(-1, 1, (r, c))))
    raise ValueError(kind)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 64), (70, 33, 100), (256, 256, 256), (300, 257, 513)])
@pytest.mark.parametrize("kind", ["uniform", "wide", "point", "zerorows", "straddle"])
def test_ozaki_bitexact(ctx, m, n, k, kind):
    rng = np.random.default_rng(hash((m, n, k, kind, "oz")) % 2 ** 32)
    IA, IB = _iv(rng, m, k, kind), _iv(rng, k, n, kind)
    C = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    R = O.ozaki_igemm(IA, IB, 16)
    bad = np.argwhere(~((C == R).all(axis=2)))
    assert len(bad) == 0, "first mismatches %s gpu %s oracle %s" % (bad[:3].tolist(), C[tuple(bad[0])], R[tuple(bad[0])])


@pytest.mark.parametrize("nmod", [8, 9, 10, 11, 12, 14, 17, 18, 20])   # one to four CRT products per wave
def test_ozaki_nmod_bitexact(ctx, nmod):
    rng = np.random.default_rng(nmod)
    IA, IB = _iv(rng, 128, 200, "wide"), _iv(rng, 200, 96, "wide")
    ctx.set_param(vcp_amd.PARAM_OZ_NMOD, nmod)
    try:
        C = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    finally:
        ctx.set_param(vcp_amd.PARAM_OZ_NMOD, 0)
    assert np.array_equal(C, O.ozaki_igemm(IA, IB, nmod))


@pytest.mark.parametrize("variant", [6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18])
def test_ozaki_tile_variants_agree(ctx, variant):
    rng = np.random.default_rng(99)
    IA, IB = _iv(rng, 300, 520, "wide"), _iv(rng, 520, 280, "wide")
    ctx.set_param(vcp_amd.PARAM_OZ_VARIANT, variant)
    try:
        C = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    finally:
        ctx.set_param(vcp_amd.PARAM_OZ_VARIANT, 0)
    assert np.array_equal(C, O.ozaki_igemm(IA, IB, 16))


@pytest.mark.parametrize("kind", ["wide", "straddle"])
def test_ozaki_containment_and_radius(ctx, kind):
    rng = np.random.default_rng(5)
    m, n, k = 128, 96, 1024
    IA, IB = _iv(rng, m, k, kind), _iv(rng, k, n, kind)
    C = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    P = ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN)
    ii = rng.integers(0, m, 512)
    jj = rng.integers(0, n, 512)
    ok, _, _ = O.check_containment(IA, IB, C, ii, jj)
    assert ok.all()
    rad_o = (C[..., 1] - C[..., 0])
    rad_p = (P[..., 1] - P[..., 0])
    ratio = np.max(rad_o / np.maximum(rad_p, np.finfo(float).tiny))
    assert ratio <= 2.0, ratio


def test_ozaki_nonfinite_operands_match_oracle(ctx):
    """rows / columns with infinite or NaN ends: [-inf, +inf] or [NaN, NaN]
    entries, exactly where the oracle puts them; the rest bit-exact"""
    rng = np.random.default_rng(41)
    IA, IB = _iv(rng, 150, 300, "wide"), _iv(rng, 300, 90, "wide")
    IA[7, 20, 1] = np.inf
    IA[100, 299, 0] = -np.inf
    IA[33, 5, :] = np.nan
    IB[0, 17, 1] = np.inf
    IB[150, 60, 0] = np.nan
    C = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    assert np.array_equal(C, O.ozaki_igemm(IA, IB, 16), equal_nan=True)
    assert C[7, 0, 0] == -np.inf and C[7, 0, 1] == np.inf and np.isnan(C[33, 0]).all() and np.isnan(C[0, 60]).all()
    P = ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN)
    assert np.array_equal(P, O.pidblas_mulmm(IA, IB), equal_nan=True)


def test_ozaki_extreme_exponents(ctx):
    """rows near the overflow threshold and columns of subnormal values: the
    scaling, the upward digits and the conversion of tiny and huge products
    stay bit-identical to the oracle and contain the exact hull"""
    rng = np.random.default_rng(43)
    m, n, k = 40, 36, 300
    IA, IB = _iv(rng, m, k, "wide"), _iv(rng, k, n, "wide")
    IA[0] = np.ldexp(IA[0], 960)            # products near 2^1000
    IA[1] = np.ldexp(IA[1], -1000)          # tiny rows
    IB[:, 0] = np.ldexp(IB[:, 0], -1060)    # subnormal column (values down to 2^-1080)
    IB[:, 1] = np.ldexp(IB[:, 1], 40)       # row 0 x column 1 overflows to +-inf ends
    C = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    assert np.array_equal(C, O.ozaki_igemm(IA, IB, 16), equal_nan=True)
    P = ctx.igemm(IA, IB, vcp_amd.FORM_PLAIN)
    assert np.array_equal(P, O.pidblas_mulmm(IA, IB), equal_nan=True)
    ii = np.array([0, 0, 1, 1, 2, 5])
    jj = np.array([0, 2, 0, 3, 0, 7])
    for X in (C, P):
        ok, _, _ = O.check_containment(IA, IB, X, ii, jj)
        assert ok.all()
    # row 0 x column 1 overflows: an infinite end encloses it (the exact-hull
    # checker works on finite bounds only)
    assert C[0, 1, 0] == -np.inf or C[0, 1, 1] == np.inf


@pytest.mark.parametrize("radius", [vcp_amd.OZ_RADIUS_FULL, vcp_amd.OZ_RADIUS_FOLDED, vcp_amd.OZ_RADIUS_HALF,
                                    vcp_amd.OZ_RADIUS_CEIL])
@pytest.mark.parametrize("nmod", [12, 16, 20])
@pytest.mark.parametrize("case", ozaki_cases.NAMES)
def test_ozaki_digit_stress_every_entry(ctx, case, nmod, radius):
    """inputs whose radius-term factors mostly have a zero high digit (the
    low-digit products carry a large share of the radius) and, at 20 moduli,
    midpoint sums beyond 2^127: bit-exact to the oracle with the same moduli
    count and radius scheme, and every entry contains the exact hull"""
    IA, IB = ozaki_cases.make(case)
    print("case", case, "nmod", nmod, "radius", radius, IA.shape, IB.shape, flush=True)
    ctx.set_param(vcp_amd.PARAM_OZ_NMOD, nmod)
    ctx.set_param(vcp_amd.PARAM_OZ_RADIUS, radius)
    try:
        C = ctx.igemm(IA, IB, vcp_amd.FORM_OZAKI)
    finally:
        ctx.set_param(vcp_amd.PARAM_OZ_NMOD, 0)
        ctx.set_param(vcp_amd.PARAM_OZ_RADIUS, 0)
    R = O.ozaki_igemm(IA, IB, nmod, radius)
    bad = np.argwhere(~((C == R).all(axis=2)))
    assert len(bad) == 0, "first mismatches %s gpu %s oracle %s" % (bad[:3].tolist(), C[tuple(bad[0])], R[tuple(bad[0])])
    ii, jj = ozaki_cases.every_entry(C)
    ok, _, _ = O.check_containment(IA, IB, C, ii, jj)
    assert ok.all(), "%d of %d entries not enclosed" % ((~ok).sum(), ok.size)


def test_ozaki_radius_param_checked(ctx):
    for bad in (5, -1, 7):
        with pytest.raises(vcp_amd.VcphipError):
            ctx.set_param(vcp_amd.PARAM_OZ_RADIUS, bad)
    assert ctx.get_param(vcp_amd.PARAM_OZ_RADIUS) == 0
