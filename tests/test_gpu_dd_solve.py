"""pddblas's refinement algorithms with their dd products on the GPU:
linearsolve / inv (vcp/pddblas.hpp:412-486), Cholesky (:488-563) and the
Ogita-Aishima eigensolver (:219-314, :566-606).  vcphip_dd_residual and
vcphip_ddgemm_ta are bit-identical to the oracle; the refined solutions,
inverses, factors and eigenpairs are within dd accuracy of exact rational
arithmetic."""
from fractions import Fraction as F

import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd

pytestmark = pytest.mark.gpu


def dd_rand(rng, r, c, diag=0.0):
    hi = rng.uniform(-1, 1, (r, c))
    if diag:
        hi = hi + diag * np.eye(r, c)
    lo = hi * 2.0 ** -53 * rng.uniform(-1, 1, (r, c))
    s = hi + lo
    X = O.pairs(r, c)
    X[..., 0] = s
    X[..., 1] = lo - (s - hi)
    return X


def val(p):
    return F(p[0]) + F(p[1])


def exact_solve(A, b):
    n = len(A)
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        p = max(range(c, n), key=lambda r: abs(M[r][c]))
        M[c], M[p] = M[p], M[c]
        for r in range(c + 1, n):
            f = M[r][c] / M[c][c]
            if f:
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    x = [F(0)] * n
    for r in range(n - 1, -1, -1):
        x[r] = (M[r][n] - sum(M[r][j] * x[j] for j in range(r + 1, n))) / M[r][r]
    return x


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 3, 17), (64, 1, 64), (130, 70, 257), (256, 256, 256)])
def test_dd_residual_bitexact(ctx, m, n, k):
    print("dd_residual", m, n, k, flush=True)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A, X, B = dd_rand(rng, m, k), dd_rand(rng, k, n), dd_rand(rng, m, n)
    R = ctx.dd_residual(A, X, B)
    assert np.array_equal(R, O.dd_residual(A, X, B))


def test_dd_residual_dev_entry(ctx):
    rng = np.random.default_rng(5)
    m, n, k = 40, 6, 50
    A, X, B = dd_rand(rng, m, k), dd_rand(rng, k, n), dd_rand(rng, m, n)
    t = lambda P: torch.from_numpy(np.ascontiguousarray(P.transpose(1, 0, 2))).cuda()
    dA, dX, dB = t(A), t(X), t(B)
    dR = torch.empty((n, m, 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    L = vcp_amd.lib()
    assert L.vcphip_dd_residual_dev(ctx._h, m, n, k, dA.data_ptr(), m, dX.data_ptr(), k, dB.data_ptr(), m,
                                    dR.data_ptr(), m) == 0
    ctx.synchronize()
    assert np.array_equal(dR.cpu().numpy().transpose(1, 0, 2), O.dd_residual(A, X, B))


@pytest.mark.parametrize("n,nrhs", [(5, 1), (9, 2)])
def test_dd_linearsolve_matches_exact_solution(ctx, n, nrhs):
    rng = np.random.default_rng(10 + n)
    A, b = dd_rand(rng, n, n, diag=3.0), dd_rand(rng, n, nrhs)
    x, it = ctx.dd_linearsolve(A, b)
    assert 2 <= it <= 6
    Ae = [[val(A[i, j]) for j in range(n)] for i in range(n)]
    for j in range(nrhs):
        xe = exact_solve(Ae, [val(b[i, j]) for i in range(n)])
        for i in range(n):
            assert abs(val(x[i, j]) - xe[i]) <= F(2) ** -100 * max(abs(xe[i]), F(1, 2 ** 40))


def test_dd_linearsolve_residual_at_n300(ctx):
    rng = np.random.default_rng(12)
    n = 300
    A, b = dd_rand(rng, n, n, diag=float(n) ** 0.5 * 2), dd_rand(rng, n, 1)
    x, it = ctx.dd_linearsolve(A, b)
    for i in (0, 17, 150, 299):   # exact residual of sampled rows
        r = val(b[i, 0]) - sum(val(A[i, t]) * val(x[t, 0]) for t in range(n))
        scale = sum(abs(val(A[i, t]) * val(x[t, 0])) for t in range(n))
        assert abs(r) <= F(2) ** -98 * scale


def test_dd_inv(ctx):
    rng = np.random.default_rng(13)
    n = 48
    A = dd_rand(rng, n, n, diag=8.0)
    X = ctx.dd_inv(A)
    for i, j in ((0, 0), (3, 7), (47, 47), (20, 5)):
        s = sum(val(A[i, t]) * val(X[t, j]) for t in range(n)) - (1 if i == j else 0)
        assert abs(s) <= F(2) ** -98 * sum(abs(val(A[i, t]) * val(X[t, j])) for t in range(n))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (33, 17, 70), (256, 128, 300)])
def test_ddgemm_ta_bitexact(ctx, m, n, k):
    rng = np.random.default_rng(m + 2 * n + 3 * k)
    A, B = dd_rand(rng, k, m), dd_rand(rng, k, n)
    C = ctx.ddgemm_ta(A, B)
    At = O.pairs(m, k)   # A^T materialised for the oracle
    At[...] = A.transpose(1, 0, 2)
    assert np.array_equal(C, O.ozaki_ddgemm(At, B, O.dd_default_nmod(k)))


def _spd(rng, n):
    G = dd_rand(rng, n, n)
    S = G[..., 0] @ G[..., 0].T + n * np.eye(n)
    L = 2.0 ** -60 * rng.uniform(-1, 1, (n, n))
    L = (L + L.T) * 0.5
    A = O.pairs(n, n)
    A[..., 0] = S
    A[..., 1] = np.where(np.abs(L) <= 2.0 ** -53 * np.abs(S), L, 0.0)
    assert np.array_equal(A[..., 1], A[..., 1].T)
    return A


def test_dd_cholesky(ctx):
    rng = np.random.default_rng(21)
    n = 40
    A = _spd(rng, n)
    U = ctx.dd_cholesky(A)
    assert (U[np.tril_indices(n, -1)] == 0.0).all()
    for i, j in ((0, 0), (5, 9), (39, 39), (12, 30)):
        s = sum(val(U[t, i]) * val(U[t, j]) for t in range(n)) - val(A[i, j])
        scale = sum(abs(val(U[t, i]) * val(U[t, j])) for t in range(n))
        assert abs(s) <= F(2) ** -98 * scale


def test_dd_eigsym(ctx):
    rng = np.random.default_rng(22)
    n = 24
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    S = Q @ np.diag(np.arange(1.0, n + 1)) @ Q.T
    S = (S + S.T) * 0.5
    A = O.pairs(n, n)
    A[..., 0] = S
    lam, X = ctx.dd_eigsym(A)
    anorm = float(np.abs(S).sum(axis=1).max())
    for c in (0, 7, 23):
        for i in (0, 11, 23):
            r = sum(val(A[i, t]) * val(X[t, c]) for t in range(n)) - val(lam[c]) * val(X[i, c])
            assert abs(r) <= F(2) ** -96 * F(anorm)
    for i, j in ((0, 0), (3, 4), (23, 23)):
        g = sum(val(X[t, i]) * val(X[t, j]) for t in range(n)) - (1 if i == j else 0)
        assert abs(g) <= F(2) ** -98 * n
    assert np.allclose(np.sort(lam[:, 0]), np.arange(1.0, n + 1), rtol=0, atol=1e-12)


def test_dd_eigsymge(ctx):
    """A x = lambda B x: exact residuals A x - lambda B x and X^T B X - I"""
    rng = np.random.default_rng(23)
    n = 20
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    S = Q @ np.diag(np.linspace(-3.0, 5.0, n)) @ Q.T
    A = O.pairs(n, n)
    A[..., 0] = (S + S.T) * 0.5
    G = rng.standard_normal((n, n))
    Bm = G @ G.T / n + np.eye(n)
    B = O.pairs(n, n)
    B[..., 0] = (Bm + Bm.T) * 0.5
    lam, X = ctx.dd_eigsymge(A, B)
    anorm = float(np.abs(A[..., 0]).sum(axis=1).max()) + float(np.abs(B[..., 0]).sum(axis=1).max()) * 5
    for c in (0, 9, 19):
        for i in (0, 10, 19):
            r = sum(val(A[i, t]) * val(X[t, c]) for t in range(n)) - \
                val(lam[c]) * sum(val(B[i, t]) * val(X[t, c]) for t in range(n))
            assert abs(r) <= F(2) ** -95 * F(anorm)
    for i, j in ((0, 0), (2, 5), (19, 19)):
        g = sum(val(X[s, i]) * sum(val(B[s, t]) * val(X[t, j]) for t in range(n)) for s in range(n)) - (1 if i == j else 0)
        assert abs(g) <= F(2) ** -96 * n


def _host_lu_refinement(ctx, A, b, itmax=10):
    """pddblas::linearsolve's loop with LAPACK's LU on the host (scipy) and
    the GPU residual: the reference's arrangement, for comparison with the
    all-device vcphip_dd_solve"""
    import scipy.linalg

    n = A.shape[0]
    lu = scipy.linalg.lu_factor(np.asfortranarray(A[..., 0]), check_finite=False)
    x = O.pairs(n, b.shape[1])
    r, prev, it = b, np.inf, 0
    while it < itmax:
        d = scipy.linalg.lu_solve(lu, np.asfortranarray(r[..., 0]), check_finite=False)
        s = x[..., 0] + d
        bb = s - x[..., 0]
        e = (x[..., 0] - (s - bb)) + (d - bb) + x[..., 1]
        h = s + e
        x[..., 1] = e - (h - s)
        x[..., 0] = h
        it += 1
        normd, normx = float(np.abs(d).max()), float(np.abs(x[..., 0]).max())
        if normd == 0.0 or normd <= 0.25e-32 * normx or normd >= 0.5 * prev:
            break
        prev = normd
        r = ctx.dd_residual(A, x, b)
    return x, it


@pytest.mark.parametrize("n,nrhs", [(60, 3), (257, 1)])
def test_dd_solve_agrees_with_host_lu_refinement(ctx, n, nrhs):
    rng = np.random.default_rng(40 + n)
    A, b = dd_rand(rng, n, n, diag=float(n) ** 0.5 * 2), dd_rand(rng, n, nrhs)
    x, it = ctx.dd_linearsolve(A, b)
    xh, ith = _host_lu_refinement(ctx, A, b)
    assert 2 <= it <= 6 and abs(it - ith) <= 1
    # full dd values: the high parts alone may differ by an ulp near a
    # rounding boundary while the values agree to 2^-100
    diff = (x[..., 0] - xh[..., 0]) + (x[..., 1] - xh[..., 1])
    assert np.abs(diff).max() <= 2.0 ** -90 * np.abs(x[..., 0]).max()


def test_dd_solve_at_n1000_exact_rows(ctx):
    print("dd_solve n=1000", flush=True)
    rng = np.random.default_rng(44)
    n = 1000
    A, b = dd_rand(rng, n, n, diag=float(n) ** 0.5 * 2), dd_rand(rng, n, 2)
    x, it = ctx.dd_linearsolve(A, b)
    assert 2 <= it <= 6
    for i, j in ((0, 0), (499, 1), (999, 0)):
        r = val(b[i, j]) - sum(val(A[i, t]) * val(x[t, j]) for t in range(n))
        scale = sum(abs(val(A[i, t]) * val(x[t, j])) for t in range(n))
        assert abs(r) <= F(2) ** -98 * scale


def test_dd_solve_dev_entry_and_leading_dimensions(ctx):
    rng = np.random.default_rng(45)
    n, nrhs, lda, ldb, ldx = 70, 3, 75, 72, 80
    A, b = dd_rand(rng, n, n, diag=20.0), dd_rand(rng, n, nrhs)

    def padded(P, ld):
        t = torch.zeros((P.shape[1], ld, 2), dtype=torch.float64, device="cuda")
        t[:, :P.shape[0], :] = torch.from_numpy(np.ascontiguousarray(P.transpose(1, 0, 2))).cuda()
        return t

    dA, dB = padded(A, lda), padded(b, ldb)
    dX = torch.full((nrhs, ldx, 2), 5.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    it = ctx.dd_solve_dev(n, nrhs, dA.data_ptr(), lda, dB.data_ptr(), ldb, dX.data_ptr(), ldx)
    x, ith = ctx.dd_linearsolve(A, b)
    X = dX.cpu().numpy()
    assert it == ith
    assert np.array_equal(X[:, :n, :].transpose(1, 0, 2), x)
    assert (X[:, n:, :] == 5.0).all()


def test_dd_solve_singular(ctx):
    n = 6
    rng = np.random.default_rng(46)
    A, b = dd_rand(rng, n, n, diag=3.0), dd_rand(rng, n, 1)
    A[:, 4, :] = 0.0
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.dd_linearsolve(A, b)
    assert e.value.code == -10


def test_dd_cholesky_agrees_with_host_arrangement(ctx):
    """the all-device Cholesky against the reference's arrangement restated
    with LAPACK / BLAS on the host and the GPU dd products (ddrefine)"""
    import ddrefine

    rng = np.random.default_rng(23)
    n = 130
    A = _spd(rng, n)
    U = ctx.dd_cholesky(A)
    Uh = ddrefine.cholesky(ctx, A)
    diff = (U[..., 0] - Uh[..., 0]) + (U[..., 1] - Uh[..., 1])
    assert np.abs(diff).max() <= 2.0 ** -90 * np.abs(U[..., 0]).max()
    assert (U[np.tril_indices(n, -1)] == 0.0).all()


def test_dd_cholesky_at_n300_exact_entries(ctx):
    rng = np.random.default_rng(24)
    n = 300
    A = _spd(rng, n)
    U = ctx.dd_cholesky(A)
    for i, j in ((0, 0), (17, 250), (299, 299), (150, 151)):
        s = sum(val(U[t, i]) * val(U[t, j]) for t in range(n)) - val(A[i, j])
        scale = sum(abs(val(U[t, i]) * val(U[t, j])) for t in range(n))
        assert abs(s) <= F(2) ** -98 * scale


def test_dd_cholesky_errors(ctx):
    rng = np.random.default_rng(25)
    n = 12
    A = _spd(rng, n)
    B = O.pairs(n, n)
    B[...] = A
    B[2, 5, 1] = 2.0 ** -80   # breaks symmetry in the low part only
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.dd_cholesky(B)
    assert e.value.code == -2
    C = O.pairs(n, n)
    C[...] = A
    C[..., 0] = -C[..., 0]    # negative definite
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.dd_cholesky(C)
    assert e.value.code == -10


def test_dd_cholesky_dev_leading_dimensions(ctx):
    rng = np.random.default_rng(26)
    n, lda, ldu = 50, 53, 57
    A = _spd(rng, n)
    dA = torch.zeros((n, lda, 2), dtype=torch.float64, device="cuda")
    dA[:, :n, :] = torch.from_numpy(np.ascontiguousarray(A.transpose(1, 0, 2))).cuda()
    dU = torch.full((n, ldu, 2), 3.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.dd_cholesky_dev(n, dA.data_ptr(), lda, dU.data_ptr(), ldu)
    Ud = dU.cpu().numpy()
    assert np.array_equal(Ud[:, :n, :].transpose(1, 0, 2), ctx.dd_cholesky(A))
    assert (Ud[:, n:, :] == 3.0).all()


def _sym_dd(rng, n, spectrum):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    S = Q @ np.diag(spectrum) @ Q.T
    A = O.pairs(n, n)
    A[..., 0] = (S + S.T) * 0.5
    L = 2.0 ** -60 * rng.uniform(-1, 1, (n, n))
    A[..., 1] = np.where(np.abs(L + L.T) * 0.5 <= 2.0 ** -53 * np.abs(A[..., 0]), (L + L.T) * 0.5, 0.0)
    return A


def _same_up_to_sign(X, Y, tol):
    """columns of two dd eigenvector matrices agree up to a sign each"""
    for j in range(X.shape[1]):
        s = 1.0 if float(np.dot(X[:, j, 0], Y[:, j, 0])) >= 0 else -1.0
        d = (X[:, j, 0] - s * Y[:, j, 0]) + (X[:, j, 1] - s * Y[:, j, 1])
        if np.abs(d).max() > tol:
            return False
    return True


def test_dd_eigsym_agrees_with_host_arrangement(ctx):
    """the all-device eigensolver against the reference's arrangement
    restated with host LAPACK and host double-double updates (ddrefine)"""
    import ddrefine

    rng = np.random.default_rng(61)
    n = 96
    A = _sym_dd(rng, n, np.linspace(-4.0, 7.0, n) + 0.01 * rng.standard_normal(n))
    lam, X = ctx.dd_eigsym(A)
    lh, Xh = ddrefine.eigsym(ctx, A)
    dl = (lam[:, 0] - lh[:, 0]) + (lam[:, 1] - lh[:, 1])
    assert np.abs(dl).max() <= 2.0 ** -96 * np.abs(lh[:, 0]).max()
    assert _same_up_to_sign(X, Xh, 2.0 ** -90)


def test_dd_eigsym_at_n300_exact_residuals(ctx):
    rng = np.random.default_rng(62)
    n = 300
    A = _sym_dd(rng, n, np.linspace(1.0, 50.0, n))
    lam, X = ctx.dd_eigsym(A)
    anorm = float(np.abs(A[..., 0]).sum(axis=1).max())
    for c in (0, 151, 299):
        for i in (0, 77, 299):
            r = sum(val(A[i, t]) * val(X[t, c]) for t in range(n)) - val(lam[c]) * val(X[i, c])
            assert abs(r) <= F(2) ** -94 * F(anorm)


def test_dd_eigsymge_agrees_with_host_arrangement(ctx):
    import ddrefine

    rng = np.random.default_rng(63)
    n = 64
    A = _sym_dd(rng, n, np.linspace(-3.0, 5.0, n))
    G = rng.standard_normal((n, n))
    Bm = G @ G.T / n + np.eye(n)
    B = O.pairs(n, n)
    B[..., 0] = (Bm + Bm.T) * 0.5
    lam, X = ctx.dd_eigsymge(A, B)
    lh, Xh = ddrefine.eigsymge(ctx, A, B)
    dl = (lam[:, 0] - lh[:, 0]) + (lam[:, 1] - lh[:, 1])
    assert np.abs(dl).max() <= 2.0 ** -94 * np.abs(lh[:, 0]).max()
    assert _same_up_to_sign(X, Xh, 2.0 ** -88)


def test_dd_eig_errors_and_dev_entry(ctx):
    rng = np.random.default_rng(64)
    n = 16
    A = _sym_dd(rng, n, np.arange(1.0, n + 1))
    N = O.pairs(n, n)
    N[...] = A
    N[3, 9, 0] += 1e-3   # not symmetric
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.dd_eigsym(N)
    assert e.value.code == -2
    Bneg = O.pairs(n, n)
    Bneg[..., 0] = -np.eye(n)   # not positive definite
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.dd_eigsymge(A, Bneg)
    assert e.value.code == -10
    lda, ldx = 19, 21
    dA = torch.zeros((n, lda, 2), dtype=torch.float64, device="cuda")
    dA[:, :n, :] = torch.from_numpy(np.ascontiguousarray(A.transpose(1, 0, 2))).cuda()
    dX = torch.full((n, ldx, 2), 9.0, dtype=torch.float64, device="cuda")
    dl = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    L = vcp_amd.lib()
    assert L.vcphip_dd_eigsym_dev(ctx._h, n, dA.data_ptr(), lda, dl.data_ptr(), dX.data_ptr(), ldx, 4) == 0
    ctx.synchronize()
    lam, X = ctx.dd_eigsym(A)
    Xd = dX.cpu().numpy()
    assert np.array_equal(dl.cpu().numpy(), lam)
    assert np.array_equal(Xd[:, :n, :].transpose(1, 0, 2), X)
    assert (Xd[:, n:, :] == 9.0).all()
