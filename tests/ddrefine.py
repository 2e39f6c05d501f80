"""pddblas's refinement algorithms in the reference's own arrangement:
Cholesky (vcp/pddblas.hpp:488-563) and the Ogita-Aishima symmetric
eigensolvers (oa_step :219-314, eigsym :566-606, eigsymge :611-659) with
their LAPACK / BLAS steps on the host (scipy / numpy, as the reference's
dpotrf / dtrsm / dsyev / dsygv) and every ozaki_acc product a vcphip dd
product (ddgemm, ddgemm_tn for X^T X, ddgemm_ta for X^T Y).  The library's
all-device entries (vcphip_dd_cholesky, vcphip_dd_eigsym / eigsymge) are
checked against these; the elementwise dd arithmetic between the products
is the round-to-nearest double-double arithmetic below, the same formulas
as the device kernels' (kv::dd's own operator bits are not available to pin
against).

Arrays are (rows, cols, 2) AoS column-major pair arrays {hi, lo}."""
import numpy as np


def _pairs(r, c):
    return np.zeros((c, r, 2), dtype=np.float64).transpose(1, 0, 2)


# ---- elementwise double-double arithmetic (numpy, round to nearest) --------
def two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def fast_two_sum(a, b):
    s = a + b
    return s, b - (s - a)


_SPLIT = 134217729.0   # 2^27 + 1 (Veltkamp)


def _split(a):
    t = _SPLIT * a
    hi = t - (t - a)
    return hi, a - hi


def two_prod(a, b):
    """p + e = a b exactly (Dekker; |a|, |b| < 2^996)"""
    p = a * b
    ah, al = _split(a)
    bh, bl = _split(b)
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def dd_add(xh, xl, yh, yl):
    s, e = two_sum(xh, yh)
    t, f = two_sum(xl, yl)
    e = e + t
    s, e = fast_two_sum(s, e)
    e = e + f
    return fast_two_sum(s, e)


def dd_mul(xh, xl, yh, yl):
    p, e = two_prod(xh, yh)
    e = e + (xh * yl + xl * yh)
    return fast_two_sum(p, e)


def dd_mul_d(xh, xl, d):
    p, e = two_prod(xh, d)
    e = e + xl * d
    return fast_two_sum(p, e)


def dd_div(xh, xl, yh, yl):
    """x / y: two correction steps of the leading quotient"""
    q1 = xh / yh
    ph, pl = dd_mul_d(yh, yl, q1)
    rh, rl = dd_add(xh, xl, -ph, -pl)
    q2 = rh / yh
    ph, pl = dd_mul_d(yh, yl, q2)
    rh, rl = dd_add(rh, rl, -ph, -pl)
    q3 = rh / yh
    q1, q2 = fast_two_sum(q1, q2)
    return dd_add(q1, q2, q3, np.zeros_like(q3))


def _set(P, h, l):
    P[..., 0] = h
    P[..., 1] = l
    return P


# ---- Cholesky: pddblas::Cholesky (vcp/pddblas.hpp:488-563) -----------------
def cholesky(ctx, A, itep=2):
    """upper triangular U with U^T U = A to dd accuracy (A symmetric positive
    definite, dd).  dpotrf of the high parts, then itep refinement steps:
    E = A - U^T U (GPU, exactly symmetric), G = Ud^-T E Ud^-1 (host dtrsm),
    W = triu(G, 1) + diag(G) / 2, U += W U (GPU)."""
    import scipy.linalg

    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("Cholesky: matrix must be square")
    if not (np.array_equal(A[..., 0], A[..., 0].T) and np.array_equal(A[..., 1], A[..., 1].T)):
        raise ValueError("Cholesky: matrix must be symmetric")
    Ud = scipy.linalg.cholesky(np.asfortranarray(A[..., 0]), lower=False)
    U = _set(_pairs(n, n), Ud, 0.0)
    for _ in range(itep):
        P = ctx.ddgemm_tn(U)
        E = _set(_pairs(n, n), *dd_add(A[..., 0], A[..., 1], -P[..., 0], -P[..., 1]))
        Udh = np.asfortranarray(U[..., 0])
        G = scipy.linalg.solve_triangular(Udh, E[..., 0], trans="T", lower=False, check_finite=False)
        G = scipy.linalg.solve_triangular(Udh, G.T, trans="T", lower=False, check_finite=False).T
        W = np.triu(G, 1) + np.diag(np.diag(G) * 0.5)
        F = ctx.ddgemm(_set(_pairs(n, n), W, 0.0), U)
        U = _set(_pairs(n, n), *dd_add(U[..., 0], U[..., 1], F[..., 0], F[..., 1]))
    il = np.tril_indices(n, -1)
    U[il[0], il[1], :] = 0.0
    return U


# ---- Ogita-Aishima: pddblas::oa_step / eigsym (vcp/pddblas.hpp:219-314, 566-606)
def oa_step(ctx, A, X, lam, B=None):
    """one refinement step of the eigenpairs (X columns, lam) of symmetric A,
    with the B-inner product when B is given (A x = lambda B x)"""
    n = A.shape[0]
    AX = ctx.ddgemm(A, X)
    S = ctx.ddgemm_ta(X, AX)
    G = ctx.ddgemm_tn(X) if B is None else ctx.ddgemm_ta(X, ctx.ddgemm(B, X))
    dg = np.arange(n)
    lam = np.stack(dd_div(S[dg, dg, 0], S[dg, dg, 1], G[dg, dg, 0], G[dg, dg, 1]), axis=-1)
    # R = I - G
    Rh, Rl = -G[..., 0], -G[..., 1]
    rh, rl = dd_add(Rh[dg, dg], Rl[dg, dg], np.ones(n), np.zeros(n))
    Rh = Rh.copy()
    Rl = Rl.copy()
    Rh[dg, dg] = rh
    Rl[dg, dg] = rl
    normR = float(np.abs(Rh).sum(axis=1).max())
    offS = float((np.abs(S[..., 0]).sum(axis=1) - np.abs(S[dg, dg, 0])).max())
    lmax = float(np.abs(lam[:, 0]).max())
    delta = 2.0 * (offS + normR * lmax)
    # E_ij = (S_ij + lam_j R_ij) / (lam_j - lam_i) where separated, else R_ij / 2
    Lj_h = np.broadcast_to(lam[:, 0][None, :], (n, n))
    Lj_l = np.broadcast_to(lam[:, 1][None, :], (n, n))
    Li_h = np.broadcast_to(lam[:, 0][:, None], (n, n))
    Li_l = np.broadcast_to(lam[:, 1][:, None], (n, n))
    dh, dl = dd_add(Lj_h, Lj_l, -Li_h, -Li_l)
    mh, ml = dd_mul(Lj_h, Lj_l, Rh, Rl)
    nh, nl = dd_add(S[..., 0], S[..., 1], mh, ml)
    with np.errstate(divide="ignore", invalid="ignore"):
        qh, ql = dd_div(nh, nl, dh, dl)
    hh, hl = dd_mul_d(Rh, Rl, 0.5)
    sep = (np.abs(dh) > delta) & ~np.eye(n, dtype=bool)
    E = _set(_pairs(n, n), np.where(sep, qh, hh), np.where(sep, ql, hl))
    XE = ctx.ddgemm(X, E)
    X = _set(_pairs(n, n), *dd_add(X[..., 0], X[..., 1], XE[..., 0], XE[..., 1]))
    return X, lam


def eigsym(ctx, A, itep=4):
    """(lam, V): eigenvalues (n, 2) and eigenvectors (n, n, 2) of symmetric
    dd A: LAPACK (dsyev-class, numpy.linalg.eigh) on the high parts, then
    max(itep, 2) Ogita-Aishima steps with the products on the GPU."""
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("eigsym: matrix must be square")
    if not (np.array_equal(A[..., 0], A[..., 0].T) and np.array_equal(A[..., 1], A[..., 1].T)):
        raise ValueError("eigsym: matrix must be symmetric")
    w, Vd = np.linalg.eigh(A[..., 0])
    X = _set(_pairs(n, n), Vd, 0.0)
    lam = np.stack([w, np.zeros(n)], axis=-1)
    for _ in range(max(itep, 2)):
        X, lam = oa_step(ctx, A, X, lam)
    return lam, X


def eigsymge(ctx, A, B, itep=1):
    """(lam, V) of the generalized problem A x = lambda B x (A symmetric, B
    symmetric positive definite, dd): LAPACK's dsygv-class solver
    (scipy.linalg.eigh) on the high parts, then max(itep, 2) Ogita-Aishima
    steps with the B-inner product (X^T B X = I), products on the GPU."""
    import scipy.linalg

    n = A.shape[0]
    for M, name in ((A, "A"), (B, "B")):
        if M.shape[:2] != (n, n):
            raise ValueError("eigsymge: dimension mismatch")
        if not (np.array_equal(M[..., 0], M[..., 0].T) and np.array_equal(M[..., 1], M[..., 1].T)):
            raise ValueError("eigsymge: matrix %s must be symmetric" % name)
    w, Vd = scipy.linalg.eigh(A[..., 0], B[..., 0])
    X = _set(_pairs(n, n), Vd, 0.0)
    lam = np.stack([w, np.zeros(n)], axis=-1)
    for _ in range(max(itep, 2)):
        X, lam = oa_step(ctx, A, X, lam, B)
    return lam, X
