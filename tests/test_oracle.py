"""Pins of the CPU oracle (no GPU): every oracle function is checked against
something it does not compute itself -- exact rational arithmetic
(fractions.Fraction), the reference's own rmatmul/udmatmul (golden fixtures
and, when built, oracle/_ref), the reference's rounding-check patterns, or
closed forms of the reference (pddblas split parameters)."""
import math
import os
from fractions import Fraction as F

import numpy as np
import pytest

import oracle_ffi as O
from vcp_amd import workloads as W

import ozaki_cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmatmul_ref.npz")


# ---- exact rational helpers (independent of the oracle) --------------------
def round_dir(q, direction):
    """round a Fraction to a double: -1 down, +1 up, 0 nearest-even"""
    x = float(q)                      # nearest (correctly rounded)
    if direction == 0 or F(x) == q:
        return x
    if direction > 0:
        return x if F(x) > q else math.nextafter(x, math.inf)
    return x if F(x) < q else math.nextafter(x, -math.inf)


def chain(a, b, direction, c0=0.0):
    c = c0
    for x, y in zip(a, b):
        c = round_dir(F(x) * F(y) + F(c), direction)
    return c


def test_exact_dot_against_fractions():
    rng = np.random.default_rng(1)
    for trial in range(200):
        k = int(rng.integers(1, 30))
        lo, hi = (-1074, 1000) if trial % 4 == 0 else (-40, 40)
        a = rng.uniform(-1, 1, k) * 2.0 ** rng.integers(lo // 2, hi // 2, k)
        b = rng.uniform(-1, 1, k) * 2.0 ** rng.integers(lo // 2, hi // 2, k)
        ex = sum(F(x) * F(y) for x, y in zip(a, b))
        rd, ru, rn = O.exact_dot(a, b)
        assert rd == round_dir(ex, -1) and ru == round_dir(ex, 1) and rn == round_dir(ex, 0)


def test_exact_hull_against_fractions():
    rng = np.random.default_rng(2)
    for trial in range(100):
        k = int(rng.integers(1, 12))
        a = np.sort(rng.uniform(-1, 1, (k, 2)) * 2.0 ** rng.integers(-20, 20, (k, 1)), axis=1)
        b = np.sort(rng.uniform(-1, 1, (k, 2)) * 2.0 ** rng.integers(-20, 20, (k, 1)), axis=1)
        lo = sum(min(F(x) * F(y) for x in ai for y in bi) for ai, bi in zip(a, b))
        hi = sum(max(F(x) * F(y) for x in ai for y in bi) for ai, bi in zip(a, b))
        okl, oku, hl, hh = O.exact_hull(a, b, round_dir(lo, -1), round_dir(hi, 1))
        assert okl and oku and hl == round_dir(lo, -1) and hh == round_dir(hi, 1)
        # one ulp inside the hull must be rejected (a plausible off-by-one)
        okl, oku, _, _ = O.exact_hull(a, b, math.nextafter(round_dir(lo, 1), math.inf),
                                      math.nextafter(round_dir(hi, -1), -math.inf))
        assert not okl and not oku


@pytest.mark.parametrize("mode", [1, -1, 0])
def test_rgemm_is_the_fma_chain(mode):
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.uniform(-1, 1, (4, 9)) * 2.0 ** rng.integers(-30, 30, (4, 9)))
    B = np.asfortranarray(rng.uniform(-1, 1, (9, 3)) * 2.0 ** rng.integers(-30, 30, (9, 3)))
    C = O.rgemm(A, B, mode)
    for i in range(4):
        for j in range(3):
            assert C[i, j] == chain(A[i], B[:, j], mode)


def test_rgemm_matches_reference_golden():
    g = np.load(GOLDEN)
    for idx in range(int(g["ncases"])):
        A, B = g["A%d" % idx], g["B%d" % idx]
        assert np.array_equal(O.rgemm(A, B, 1), g["CU%d" % idx])
        assert np.array_equal(O.rgemm(A, B, -1), g["CD%d" % idx])
        assert np.array_equal(O.rgemm(A, B, 0), g["CN%d" % idx])
        if "UU%d" % idx in g:
            C0 = g["C0%d" % idx]
            assert np.array_equal(O.rgemm(A, B, 1, C=C0.copy(order="F"), accumulate=True), g["UU%d" % idx])
            assert np.array_equal(O.rgemm(A, B, -1, C=C0.copy(order="F"), accumulate=True), g["UD%d" % idx])
    assert (g["KAT_CU"] == 1 + 2.0 ** -52).all() and (g["KAT_CD"] == 1.0).all()


@pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built (needs the reference sources)")
def test_rgemm_matches_reference_rmatmul_live():
    rng = np.random.default_rng(4)
    for (m, n, k) in [(7, 5, 3), (100, 37, 259), (300, 200, 513)]:
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * 2.0 ** rng.integers(-20, 20, (m, k)))
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * 2.0 ** rng.integers(-20, 20, (k, n)))
        for mode in (1, -1, 0):
            assert np.array_equal(O.rgemm(A, B, mode), O.ref_rmatmul(A, B, mode))


def test_rounding_check_patterns():
    # test_matrix/Check_pidblas_rounding.cpp:14-71, n reduced
    n = 40
    for i in range(n):
        A = np.zeros((n, n), order="F")
        B = np.zeros((n, n), order="F")
        A[:, i] = 1.0 / 3.0
        B[i, :] = 1.0 / 3.0
        assert (O.rgemm(A, B, 1) > O.rgemm(A, B, -1)).all()
    for i in range(2, n):
        A = np.zeros((n, n), order="F")
        B = np.zeros((n, n), order="F")
        A[:, 1] = 1.0
        B[1, :] = 1.0
        A[:, i] = 2.0 ** -27
        B[i, :] = 2.0 ** -27
        assert (O.rgemm(A, B, 1) == 1 + 2.0 ** -52).all() and (O.rgemm(A, B, -1) == 1.0).all()


def test_midrad_contract():
    rng = np.random.default_rng(5)
    cases = [(1.0, 2.0), (-1.0, 1.0), (0.1, 0.1), (5e-324, 5e-324), (5e-324, 1e-323), (-3.0, -2.5),
             (-1.7976931348623157e308, 1.7976931348623157e308)]
    for _ in range(300):
        lo, hi = sorted(rng.uniform(-1, 1, 2) * 2.0 ** rng.integers(-1070, 1000, 2))
        cases.append((lo, hi))
    for lo, hi in cases:
        m, r = O.midrad(lo, hi)
        assert r >= 0
        if lo == hi:
            assert m == lo and r == 0.0
        else:
            assert F(m) - F(r) <= F(lo) and F(m) + F(r) >= F(hi)


def _iv(rng, r, c, kind):
    if kind == "wide":
        return W.wide_exponent_intervals(rng, r, c)
    if kind == "straddle":
        return W.interval_from_midrad(rng.uniform(-1, 1, (r, c)), np.abs(rng.uniform(-1, 1, (r, c))))
    return W.uniform_intervals(rng, r, c, -40)


def _contains_all(IA, IB, C):
    m, n = C.shape[:2]
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ok, _, _ = O.check_containment(IA, IB, C, ii.ravel(), jj.ravel())
    return ok.all()


@pytest.mark.parametrize("kind", ["uniform", "wide", "straddle"])
def test_pidblas_mulmm_encloses_and_follows_op_order(kind):
    rng = np.random.default_rng(6)
    IA, IB = _iv(rng, 20, 33, kind), _iv(rng, 33, 17, kind)
    C = O.pidblas_mulmm(IA, IB)
    assert _contains_all(IA, IB, C)
    # op order: lower = RD(RD-chain(mA mB) - t), t = RU(RU-chain((|mA|+rA) rB) + RU-chain(rA |mB|))
    mA, rA = np.vectorize(O.midrad)(IA[..., 0], IA[..., 1])
    mB, rB = np.vectorize(O.midrad)(IB[..., 0], IB[..., 1])
    sA = np.vectorize(lambda a, r: round_dir(abs(F(a)) + F(r), 1))(mA, rA)
    for i in range(0, 20, 7):
        for j in range(0, 17, 5):
            up = chain(mA[i], mB[:, j], 1)
            dn = chain(mA[i], mB[:, j], -1)
            t1 = chain(sA[i], rB[:, j], 1)
            t2 = chain(rA[i], np.abs(mB[:, j]), 1)
            t = round_dir(F(t1) + F(t2), 1)
            assert C[i, j, 1] == round_dir(F(up) + F(t), 1)
            assert C[i, j, 0] == round_dir(F(dn) - F(t), -1)


def test_pidblas_mixed_and_transposed_forms_enclose():
    rng = np.random.default_rng(7)
    IA = _iv(rng, 15, 22, "wide")
    B = np.asfortranarray(rng.uniform(-1, 1, (22, 9)))
    pts = lambda X: W.interval_from_midrad(X, np.zeros_like(X))  # noqa: E731
    assert _contains_all(IA, pts(B), O.pidblas_mul_im_m(IA, B))
    Bl = np.asfortranarray(rng.uniform(-1, 1, (9, 15)))
    assert _contains_all(pts(Bl), IA, O.pidblas_mul_m_im(Bl, IA))
    IAt = np.ascontiguousarray(IA.transpose(1, 0, 2).transpose(1, 0, 2))
    C = O.pidblas_mulltmm(IA)
    IAT = W._pairs(22, 15)
    IAT[...] = IA.transpose(1, 0, 2)
    assert _contains_all(IAT, IA, C) and np.array_equal(C, C.transpose(1, 0, 2))
    del IAt
    A = np.asfortranarray(rng.uniform(-1, 1, (22, 15)))
    V = O.pidblas_vmulmm(A)
    AT = W._pairs(15, 22)
    AT[...] = pts(A).transpose(1, 0, 2)
    assert _contains_all(AT, pts(A), V)


def test_residual_and_norm_bound():
    rng = np.random.default_rng(8)
    R, IA = W.residual_problem(rng, 40)
    C, nub = O.residual(R, IA)
    n = 40
    IR = W.interval_from_midrad(R, np.zeros_like(R))
    RA = O.pidblas_mul_m_im(R, IA)
    assert _contains_all(IR, IA, RA)
    # C = RA - I on the diagonal, RD / RU
    for i in range(n):
        assert C[i, i, 0] == round_dir(F(RA[i, i, 0]) - 1, -1)
        assert C[i, i, 1] == round_dir(F(RA[i, i, 1]) - 1, 1)
    mag = np.maximum(np.abs(C[..., 0]), np.abs(C[..., 1]))
    exact = max(sum(F(x) for x in row) for row in mag)
    assert F(nub) >= exact and nub < 1.0


def test_ozaki_scaled_integers_within_half():
    rng = np.random.default_rng(9)
    IA = _iv(rng, 9, 40, "wide")
    beta = O.ozaki_beta(40, 16)
    Ap, e = O.ozaki_scaled(IA, beta, rows=True)
    mA, _ = np.vectorize(O.midrad)(IA[..., 0], IA[..., 1])
    for i in range(9):
        mu = np.max(np.abs(mA[i]))
        assert 2.0 ** (e[i] - 1) <= mu < 2.0 ** e[i]
        for c in range(40):
            x = F(mA[i, c]) * F(2) ** (beta - int(e[i]))
            assert abs(x - int(Ap[i, c])) <= F(1, 2) and abs(int(Ap[i, c])) <= 2 ** beta


def test_ozaki_moduli_and_beta():
    mods = O.ozaki_moduli(20)
    assert mods[0] == 256 and all(math.gcd(a, b) == 1 for i, a in enumerate(mods) for b in mods[i + 1:])
    for k in (1, 64, 1000, 8192, 65536):
        for nmod in (12, 16, 20):
            beta = O.ozaki_beta(k, nmod)
            M = math.prod(O.ozaki_moduli(nmod))
            assert k * 2 ** (2 * beta + 1) < M          # no wrap-around in the CRT range
            assert k * 2 ** (2 * beta + 5) >= M or beta == 62  # and not wastefully small


def test_split_quotient_24bit_exact():
    """the splits' residue quotient: floor(y / m) = (8 y ceil(2^29 / m)) >> 32
    (v_mul_hi_u32_u24: both factors below 2^24) for every y < 2^21 and every
    modulus 33 <= m <= 256; a plausible slip (floor instead of ceil, 2^28)
    fails it"""
    y = np.arange(2 ** 21, dtype=np.uint64)
    for m in range(33, 257):
        magic = np.uint64(-(-(2 ** 29) // m))
        assert int(magic) < 2 ** 24
        q = ((np.uint64(8) * y) * magic) >> np.uint64(32)
        assert np.array_equal(q, y // np.uint64(m)), m
    m = 255
    slip = ((np.uint64(8) * y) * np.uint64(2 ** 29 // m)) >> np.uint64(32)
    assert not np.array_equal(slip, y // np.uint64(m))


def test_gemm_epilogue_mod_reduction():
    """the int8 GEMM epilogue's reduction of |acc| <= 2^27 modulo m: t = acc +
    ceil(2^11/m) m 2^16, z = t + (t >> 16)((2^16 mod m) - 2^16), q from the
    24-bit quotient of 8z, r = z - q m equals acc mod m, with t < 2^29 and
    z < 2^21 (the operand ranges of v_mad_i32_i24 and v_mul_hi_u32_u24), for
    every modulus 33 <= m <= 256 on the extremes, the multiples of m around
    them and random sums"""
    rng = np.random.default_rng(27)
    base = np.concatenate([np.array([0, 1, -1, 2 ** 27, -2 ** 27, 2 ** 27 - 1, -2 ** 27 + 1], dtype=np.int64),
                           rng.integers(-2 ** 27, 2 ** 27 + 1, 200000)])
    for m in range(33, 257):
        near = np.array([k * m + d for k in (2 ** 27 // m, -(2 ** 27 // m), 0, 1, -1) for d in (-1, 0, 1)], dtype=np.int64)
        acc = np.concatenate([base, near[np.abs(near) <= 2 ** 27]])
        kbias = ((2048 + m - 1) // m * m) << 16
        c16m = (65536 % m) - 65536
        magic = -(-(2 ** 29) // m)
        t = acc + kbias
        assert t.min() >= 0 and t.max() < 2 ** 29
        th = t >> 16
        assert th.max() < 2 ** 23
        z = t + th * c16m
        assert z.min() >= 0 and z.max() < 2 ** 21
        q = ((8 * z) * magic) >> 32
        r = z - q * m
        assert np.array_equal(r, np.mod(acc, m)), m


def test_crt_symmetric_reconstruction_identity():
    """the reconstruction the kernels use: X = sum_t u_t w_t - q M with
    w_t = (M/m_t)((M/m_t)^-1 mod m_t) and q = floor((sum_t u_t F_t + 2^47) /
    2^48), F_t = floor(2^48 w_t / M), recovers every X the splitting admits,
    including the extremes |X| = (M/2) 2^-1e-6 (interval form) and M/4 (dd
    form); so does the double-precision q = floor(0.5 + sum_t u_t fl(w_t/M)).
    The byte-column sums give the same S and Q, and each equals the centred
    int8 product sum_t (u_t - 128)(v_t - 128) plus 128 sum_t u_t plus the
    column constant 128 sum_t v_t - 128^2 N, with sum_t u_t itself the
    centred product against a column of ones plus 128 N (the matrix-core
    form)"""
    rng = np.random.default_rng(10)
    for nmod in (8, 12, 16, 20, 24, 28, 32):
        mods = O.ozaki_moduli(nmod)
        M = math.prod(mods)
        w = []
        for m in mods:
            Mt = M // m
            w.append(Mt * pow(Mt % m, -1, m))
        f = [float(F(wt, M)) for wt in w]
        assert max(w) < 2 ** (8 * nmod)              # bytes past N are zero: N weight columns
        Fq = [(wt << 48) // M for wt in w]
        assert all(F(wt, M) - F(ft, 2 ** 48) < F(1, 2 ** 48) for wt, ft in zip(w, Fq))
        wb = [[(wt >> (8 * c)) & 255 for wt in w] for c in range(nmod)]
        fb = [[(ft >> (8 * b)) & 255 for ft in Fq] for b in range(6)]
        for vb in wb + fb:                           # centred operands are int8
            assert all(-128 <= v - 128 <= 127 for v in vb)
        edge = int(F(M, 2) * F(2) ** F(-1, 10 ** 6)) if nmod <= 20 else M // 4 - 1
        cases = [0, 1, -1, edge, -edge, edge - 1, -(edge - 1)]
        cases += [int(x) for x in rng.integers(-2 ** 62, 2 ** 62, 100)]
        cases += [(int(rng.integers(0, 2 ** 62)) * int(rng.integers(1, 2 ** 62)) * (2 ** 64 + 1)) % (2 * edge + 1) - edge
                  for _ in range(300)]
        for X in cases:
            assert abs(X) <= edge
            u = [X % m for m in mods]
            qf = 0.5
            for ut, ft in zip(u, f):
                qf = qf + ut * ft
            q = math.floor(qf)
            S = sum(ut * wt for ut, wt in zip(u, w))
            assert S - q * M == X
            # byte columns: col_c = sum_t u_t * byte_c(w_t) < 2^21
            nb = (M.bit_length() + 7) // 8
            cols = [sum(ut * ((wt >> (8 * c)) & 255) for ut, wt in zip(u, w)) for c in range(nb)]
            assert max(cols) < 2 ** 21
            assert sum(col << (8 * c) for c, col in enumerate(cols)) == S
            # fixed-point quotient from the six bytes of F_t
            Q = sum(ut * ft for ut, ft in zip(u, Fq)) + 2 ** 47
            assert Q >> 48 == q and Q < 2 ** 63
            fcols = [sum(ut * ((ft >> (8 * b)) & 255) for ut, ft in zip(u, Fq)) for b in range(6)]
            assert max(fcols) < 2 ** 21
            assert sum(fc << (8 * b) for b, fc in enumerate(fcols)) + 2 ** 47 == Q
            # centred int8 products (column of ones for sum_t u_t)
            ssum = sum((ut - 128) * 1 for ut in u) + 128 * nmod
            assert ssum == sum(u)
            wcols = [sum(ut * v for ut, v in zip(u, vb)) for vb in wb]   # N columns (the top ones may be 0)
            assert wcols[:nb] == cols[:nmod] and not any(wcols[nb:])
            for vb, want in [(wb[c], wcols[c]) for c in range(nmod)] + [(fb[b], fcols[b]) for b in range(6)]:
                cen = sum((ut - 128) * (v - 128) for ut, v in zip(u, vb))
                assert cen + 128 * ssum + 128 * sum(vb) - 128 * 128 * nmod == want


def test_pidblas_fma_is_the_interval_fma_chain():
    """the inf-sup restatement (reference's sign-case ifma under fesetround)
    equals the chain up <- max RU(a b + up), dn <- min RD(a b + dn) over the
    four endpoint products, evaluated in exact rationals"""
    rng = np.random.default_rng(21)
    m, n, k = 6, 5, 9
    def mixed(r, c):
        lo = rng.uniform(-1, 1, (r, c)) * 2.0 ** rng.integers(-3, 4, (r, c))
        hi = lo + np.abs(rng.uniform(0, 1, (r, c))) * 2.0 ** rng.integers(-3, 2, (r, c))
        X = W._pairs(r, c)
        X[..., 0], X[..., 1] = lo, hi
        X[0, :, 0] = 0.0                  # zero lower ends
        X[0, :, 1] = np.abs(X[0, :, 1])
        X[1, :, 1] = 0.0                  # zero upper ends
        X[1, :, 0] = -np.abs(X[1, :, 0])
        X[2, :, 1] = np.minimum(X[2, :, 1], -2.0 ** -4)   # nonpositive rows
        X[2, :, 0] = X[2, :, 1] - 1.0
        assert (X[..., 0] <= X[..., 1]).all()
        return X
    IA, IB = mixed(m, k), mixed(k, n)
    C = O.pidblas_fma_mulmm(IA, IB)
    for i in range(m):
        for j in range(n):
            up, dn = 0.0, 0.0
            for t in range(k):
                prods = [F(a) * F(b) for a in IA[i, t] for b in IB[t, j]]
                up = max(round_dir(p + F(up), 1) for p in prods)
                dn = min(round_dir(p + F(dn), -1) for p in prods)
            assert C[i, j, 1] == up and C[i, j, 0] == dn


@pytest.mark.parametrize("kind", ["uniform", "wide", "straddle"])
def test_ozaki_restatement_encloses_and_is_not_wider(kind):
    rng = np.random.default_rng(11)
    IA, IB = _iv(rng, 24, 300, kind), _iv(rng, 300, 20, kind)
    P = O.pidblas_mulmm(IA, IB)
    for radius in O.OZ_RADIUS_ALL:
        assert _contains_all(IA, IB, O.ozaki_igemm(IA, IB, 16, radius))
    # The radius ratio depends on the moduli count (the scaling error E) and
    # on the data: for k = 300 products over a 41-binade exponent span, E of
    # 16 moduli alone takes about 0.6 of the plain width (max ratio 2.06 on
    # "wide"); with 18 moduli the full scheme stays within 2x.
    Z = O.ozaki_igemm(IA, IB, 18)
    assert _contains_all(IA, IB, Z)
    ratio = (Z[..., 1] - Z[..., 0]) / np.maximum(P[..., 1] - P[..., 0], np.finfo(float).tiny)
    assert ratio.max() <= 2.0


@pytest.mark.parametrize("radius", O.OZ_RADIUS_ALL)
@pytest.mark.parametrize("nmod", [12, 16, 20])
@pytest.mark.parametrize("case", ozaki_cases.NAMES)
def test_ozaki_restatement_encloses_every_entry_on_digit_stress(case, nmod, radius):
    """Every entry contains the exact hull (superaccumulator) on inputs whose
    radius-term factors mostly have a zero high digit, and radii stay within
    2x of the plain form's.  Dropping the low-digit product, or an int128
    midpoint sum that wraps (beta = 62 at 20 moduli), fails here."""
    IA, IB = ozaki_cases.make(case)
    Z = O.ozaki_igemm(IA, IB, nmod, radius)
    ii, jj = ozaki_cases.every_entry(Z)
    ok, _, _ = O.check_containment(IA, IB, Z, ii, jj)
    assert ok.all(), "%d of %d entries not enclosed" % ((~ok).sum(), ok.size)
    P = O.pidblas_mulmm(IA, IB)
    ratio = (Z[..., 1] - Z[..., 0]) / np.maximum(P[..., 1] - P[..., 0], np.finfo(float).tiny)
    assert ratio.max() <= 2.0, ratio.max()


def test_ozaki_midpoint_sum_beyond_int128():
    """20 moduli give beta = 62: |A'|, |B'| up to 2^62, so the exact midpoint
    sum of k = 300 products exceeds 2^127.  Point inputs of that size: the
    enclosure's ends straddle the exact product computed with Python
    integers from the oracle's own scaled operands and lie within a relative
    2^-40 of it."""
    rng = np.random.default_rng(21)
    m, n, k = 6, 5, 300
    IA = W.wide_exponent_intervals(rng, m, k, -1, 0, rad_exp=-60)
    IB = W.wide_exponent_intervals(rng, k, n, -1, 0, rad_exp=-60)
    IA[..., 1] = IA[..., 0] = np.abs(IA[..., 0])   # one sign: no cancellation
    IB[..., 1] = IB[..., 0] = np.abs(IB[..., 0])
    beta = O.ozaki_beta(k, 20)
    assert beta == 62
    Ap, e = O.ozaki_scaled(IA, beta, rows=True)
    Bp, f = O.ozaki_scaled(IB, beta, rows=False)
    Z = O.ozaki_igemm(IA, IB, 20)
    for i in range(m):
        for j in range(n):
            X = sum(int(Ap[i, c]) * int(Bp[c, j]) for c in range(k))
            assert abs(X) >= 2 ** 127
            exact = F(X) * F(2) ** (int(e[i]) + int(f[j]) - 2 * beta)
            assert F(Z[i, j, 0]) <= exact <= F(Z[i, j, 1])
            assert F(Z[i, j, 1]) - F(Z[i, j, 0]) <= abs(exact) * F(2) ** -40


def test_ozaki_point_and_zero_inputs():
    rng = np.random.default_rng(12)
    IA = _iv(rng, 10, 50, "uniform")
    IA[..., 1] = IA[..., 0]
    IB = _iv(rng, 50, 8, "uniform")
    IB[..., 1] = IB[..., 0]
    IA[3] = 0.0
    IB[:, 2] = 0.0
    Z = O.ozaki_igemm(IA, IB, 16)
    assert _contains_all(IA, IB, Z)
    assert (Z[3] == 0).all() and (Z[:, 2] == 0).all()


def test_pddblas_split_parameters_closed_form():
    # vcp/pddblas.hpp:54-78; the k=4096 / 1024 values of the dd workloads
    assert O.dd_params(4096) == (32, 20, 7)
    assert O.dd_params(1024) == (31, 21, 7)
    for k in (1, 2, 3, 100, 8192, 1 << 20):
        s, b, smax = O.dd_params(k)
        rho = math.ceil(math.log2(k)) if k > 1 else 0
        assert s == (51 + rho + 1) // 2 and b == max(1, 52 - s) and smax == 110 // b + 2
        assert 104 + rho - 2 * s <= 53


def test_pddblas_mulmm_accuracy_and_symmetry():
    rng = np.random.default_rng(13)
    A = W.dd_matrix(rng, 12, 64)
    B = W.dd_matrix(rng, 64, 10)
    C, na, nb = O.pddblas_mulmm(A, B)
    assert 2 <= na <= O.dd_params(64)[2] and 2 <= nb <= O.dd_params(64)[2]
    for i in range(12):
        for j in range(10):
            _, _, err = O.exact_dd_dot(A[i], B[:, j], C[i, j, 0], C[i, j, 1])
            scale = float(np.sum(np.abs(A[i, :, 0]) * np.abs(B[:, j, 0])))
            assert err <= 2.0 ** -100 * scale
    At = O.pddblas_mulltmm(A)
    assert np.array_equal(At, At.transpose(1, 0, 2))


def test_ozaki_nonfinite_operands():
    """a row of A / column of B with an infinite end makes its product entries
    [-inf, +inf]; with a NaN end [NaN, NaN]; every other entry is unchanged"""
    rng = np.random.default_rng(40)
    IA, IB = _iv(rng, 6, 9, "wide"), _iv(rng, 9, 5, "wide")
    C0 = O.ozaki_igemm(IA, IB, 16)
    IA[2, 4, 1] = np.inf
    IA[4, 0, 0] = -np.inf
    IB[3, 1, 0] = np.nan
    C = O.ozaki_igemm(IA, IB, 16)
    for i in range(6):
        for j in range(5):
            if j == 1:
                assert np.isnan(C[i, j]).all()
            elif i in (2, 4):
                assert C[i, j, 0] == -np.inf and C[i, j, 1] == np.inf
            else:
                assert np.array_equal(C[i, j], C0[i, j])
    A, B = W.dd_matrix(rng, 5, 8), W.dd_matrix(rng, 8, 4)
    D0 = O.ozaki_ddgemm(A, B, 31)
    A[1, 5, 0] = np.inf
    B[2, 3, 1] = np.nan
    D = O.ozaki_ddgemm(A, B, 31)
    for i in range(5):
        for j in range(4):
            if i == 1 or j == 3:
                assert np.isnan(D[i, j]).all()
            else:
                assert np.array_equal(D[i, j], D0[i, j])
    # A^T A form: a column of A is both an operand row and column
    T = W.dd_matrix(rng, 8, 5)
    T[0, 2, 1] = np.inf
    S = O.ozaki_ddgemm(T, None, 31)
    assert np.isnan(S[2, :]).all() and np.isnan(S[:, 2]).all()
    assert np.isfinite(np.delete(np.delete(S, 2, 0), 2, 1)).all()


def test_reference_built_pidblas_product_matches_the_restatement():
    """pidblas::mulmm built on the reference's own vblas rmatmul (oracle/_ref)
    is bit-identical to the oracle's restatement (both are the k-ascending
    directed chains with the reference's fix-ups)"""
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    from vcp_amd import workloads as W

    rng = np.random.default_rng(77)
    for m, n, k in ((1, 1, 1), (9, 7, 33), (70, 40, 300)):
        IA, IB = W.wide_exponent_intervals(rng, m, k), W.wide_exponent_intervals(rng, k, n)
        assert np.array_equal(O.ref_pidblas_mulmm(IA, IB), O.pidblas_mulmm(IA, IB))


@pytest.mark.parametrize("kind", ["wide", "straddle", "uniform"])
def test_reference_built_mixed_and_transposed_forms_match_the_restatement(kind):
    """pidblas::mul_im_m, mul_m_im, mulltmm and vmulmm built on the
    reference's own vblas rmatmul (oracle/_ref), each following its reference
    function's op order, are bit-identical to the oracle's restatements: an
    op-order or rounding-mode slip in either changes bits here"""
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(78)
    for m, n, k in ((1, 1, 1), (9, 7, 33), (40, 31, 150)):
        IA = _iv(rng, m, k, kind)
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * 2.0 ** rng.integers(-8, 8, (k, n)))
        assert np.array_equal(O.ref_pidblas_mul_im_m(IA, B), O.pidblas_mul_im_m(IA, B))
        Bl = np.asfortranarray(rng.uniform(-1, 1, (n, m)) * 2.0 ** rng.integers(-8, 8, (n, m)))
        assert np.array_equal(O.ref_pidblas_mul_m_im(Bl, IA), O.pidblas_mul_m_im(Bl, IA))
        assert np.array_equal(O.ref_pidblas_mulltmm(IA), O.pidblas_mulltmm(IA))
        A = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
        assert np.array_equal(O.ref_pidblas_vmulmm(A), O.pidblas_vmulmm(A))


def test_reference_built_forms_catch_a_planted_slip():
    """the pin has teeth: a restatement that swaps the rounding of the radius
    product of mul_im_m (RD instead of RU), planted here in numpy on the
    oracle's own intermediate values, differs from the reference-built result"""
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(79)
    IA = W.wide_exponent_intervals(rng, 12, 40, rad_exp=-10)
    B = np.asfortranarray(rng.uniform(-1, 1, (40, 9)))
    ref = O.ref_pidblas_mul_im_m(IA, B)
    mA, rA = np.vectorize(O.midrad)(IA[..., 0], IA[..., 1])
    slip = np.empty_like(ref)
    for i in range(12):
        for j in range(9):
            lo = chain(mA[i], B[:, j], -1)
            hi = chain(mA[i], B[:, j], 1)
            t = chain(rA[i], np.abs(B[:, j]), -1)        # the planted slip: RD chain
            slip[i, j, 1] = round_dir(F(hi) + F(t), 1)
            slip[i, j, 0] = round_dir(F(lo) - F(t), -1)
    assert not np.array_equal(slip, ref)
    # the same construction with the right rounding reproduces it exactly
    good = np.empty_like(ref)
    for i in range(12):
        for j in range(9):
            lo = chain(mA[i], B[:, j], -1)
            hi = chain(mA[i], B[:, j], 1)
            t = chain(rA[i], np.abs(B[:, j]), 1)
            good[i, j, 1] = round_dir(F(hi) + F(t), 1)
            good[i, j, 0] = round_dir(F(lo) - F(t), -1)
    assert np.array_equal(good, ref)
