"""The approximate inverse R = inv(mid(IA)) formed on the device (vcphip_inv_mid:
imid + pdblas::inv, vcp/imats.hpp:298-301, vcp/pdblas.hpp:180-204, with
rocSOLVER getrf/getrs in place of LAPACK's getrf/getri) and the verified solves that form
it themselves (R = NULL).  R is approximate, so it is checked against LAPACK
(numpy) to a tolerance and by R*mid(A) ~ I; the solves built on it are checked
bit for bit against the oracle's restatement given that same R, and by
containment of known solutions."""
import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _system(rng, n, nrhs, rad_exp=-30):
    A = rng.uniform(-1, 1, (n, n)) + 0.5 * n * np.eye(n)
    b = rng.uniform(-1, 1, (n, nrhs))
    IA = W.interval_from_midrad(A, np.ldexp(np.abs(A), rad_exp))
    Ib = W.interval_from_midrad(b, np.ldexp(np.abs(b), rad_exp))
    return IA, Ib


def _mid(IA):
    return 0.5 * IA[..., 0] + 0.5 * IA[..., 1]


@pytest.mark.parametrize("n", [1, 2, 17, 100, 300])
def test_inv_mid_against_lapack(ctx, n):
    print("inv_mid", n, flush=True)
    rng = np.random.default_rng(500 + n)
    IA, _ = _system(rng, n, 1)
    R = ctx.inv_mid(IA)
    M = _mid(IA)
    Rn = np.linalg.inv(M)
    # well-conditioned (diagonally dominant): both inverses within a few
    # hundred ulps of the exact one
    assert np.abs(R - Rn).max() <= 1e-12 * np.abs(Rn).max()
    assert np.abs(R @ M - np.eye(n)).max() <= 1e-12


def test_inv_mid_uses_the_midpoint(ctx):
    """intervals m -+ r with dyadic m and r: the device must invert m itself;
    the tridiagonal (-1, 2, -1) of order 15 has an inverse with entries
    min(i, j)(16 - max(i, j))/16, exact in binary64"""
    n = 15
    T = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    exact = np.array([[min(i, j) * (n + 1 - max(i, j)) / (n + 1) for j in range(1, n + 1)]
                      for i in range(1, n + 1)])
    rad = np.ldexp(np.ones((n, n)), -8)
    IA = W.interval_from_midrad(T, rad)
    assert np.array_equal(_mid(IA), T)
    R = ctx.inv_mid(IA)
    assert np.abs(R - exact).max() <= 64 * np.finfo(float).eps * np.abs(exact).max()


def test_inv_mid_dev_leading_dimensions(ctx):
    n, lda, ldr = 37, 41, 45
    rng = np.random.default_rng(31)
    IA, _ = _system(rng, n, 1)
    dA = torch.zeros((n, lda, 2), dtype=torch.float64, device="cuda")
    dA[:, :n, :] = torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2))).cuda()
    dR = torch.full((n, ldr), 7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.inv_mid_dev(n, dA.data_ptr(), lda, dR.data_ptr(), ldr)
    R = dR.cpu().numpy()
    assert np.array_equal(R[:, :n].T, ctx.inv_mid(IA))
    assert (R[:, n:] == 7.0).all()   # padding rows untouched


def test_inv_mid_singular(ctx):
    n = 5
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    A[:, 2] = 0.0   # an exactly zero pivot
    IA = W.interval_from_midrad(A, np.zeros_like(A))
    Ib = W.interval_from_midrad(np.ones((n, 1)), np.zeros((n, 1)))
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.inv_mid(IA)
    assert e.value.code == -10
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.lss(IA, Ib)
    assert e.value.code == -10


@pytest.mark.parametrize("n,nrhs", [(2, 1), (33, 3), (257, 2)])
def test_lss_with_device_inverse_bitexact(ctx, n, nrhs):
    """lss(R = NULL) is lss with the R that inv_mid returns, and that is the
    oracle's restatement with the same R"""
    print("lss device R", n, nrhs, flush=True)
    rng = np.random.default_rng(2000 + n)
    IA, Ib = _system(rng, n, nrhs)
    X, g = ctx.lss(IA, Ib)
    R = ctx.inv_mid(IA)
    Xo, go, ok = O.lss(IA, Ib, R)
    assert ok and g == go
    assert np.array_equal(X, Xo)
    X2, g2 = ctx.lss(IA, Ib, R)
    assert g2 == g and np.array_equal(X2, X)


def test_lss_with_device_inverse_known_answer(ctx):
    """tridiag(-1, 2, -1) of order 15, b = 1..15: the exact solution (dyadic)
    lies in X"""
    n = 15
    T = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    exact = np.array([[min(i, j) * (n + 1 - max(i, j)) / (n + 1) for j in range(1, n + 1)]
                      for i in range(1, n + 1)])
    b = np.arange(1.0, n + 1).reshape(n, 1)
    IA, Ib = W.interval_from_midrad(T, np.zeros_like(T)), W.interval_from_midrad(b, np.zeros_like(b))
    X, g = ctx.lss(IA, Ib)
    x = exact @ b   # exact: multiples of 1/16 of moderate size
    assert g < 1e-12
    assert (X[..., 0] <= x).all() and (x <= X[..., 1]).all()
    assert (X[..., 1] - X[..., 0]).max() < 1e-11 * np.abs(x).max()


def test_lss_dev_without_r(ctx):
    n, nrhs = 64, 2
    rng = np.random.default_rng(41)
    IA, Ib = _system(rng, n, nrhs)
    dA = torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2))).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(Ib.transpose(1, 0, 2))).cuda()
    dX = torch.zeros((nrhs, n, 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g = ctx.lss_dev(n, nrhs, dA.data_ptr(), n, dB.data_ptr(), n, None, 0, dX.data_ptr(), n)
    X, gh = ctx.lss(IA, Ib)
    assert g == gh
    assert np.array_equal(dX.cpu().numpy().transpose(1, 0, 2), X)


def test_lss_batched_with_device_inverses(ctx):
    """problems whose orders round up to the same multiple of 64 share one
    batched getrf/getrs; each problem's result is still the single solve's
    bit for bit.  The enclosures contain the
    midpoint solution and agree in width with the host-R oracle to 1e-3
    (radii of 2^-30 dominate the widths; the rounding error of mx, which
    depends on R, adds about 1e-5 of them)"""
    sizes = [2, 3, 17, 17, 64, 17, 100, 3]
    nrhss = [1, 2, 1, 1, 3, 2, 1, 1]
    print("lss_batched device R", sizes, flush=True)
    rng = np.random.default_rng(77)
    sys = [_system(rng, n, r) for n, r in zip(sizes, nrhss)]
    IAs, Ibs = [s[0] for s in sys], [s[1] for s in sys]
    for rep in range(2):
        Xs, gs = ctx.lss_batched(IAs, Ibs)
        for p, n in enumerate(sizes):
            assert gs[p] < 1e-3, (rep, p)
            X1, g1 = ctx.lss(IAs[p], Ibs[p])
            assert g1 == gs[p] and np.array_equal(Xs[p], X1), (rep, p)
            xm = np.linalg.solve(_mid(IAs[p]), _mid(Ibs[p]))
            assert (Xs[p][..., 0] <= xm).all() and (xm <= Xs[p][..., 1]).all(), (rep, p)
            Xo, go, ok = O.lss(IAs[p], Ibs[p], np.asfortranarray(np.linalg.inv(_mid(IAs[p]))))
            w, wo = Xs[p][..., 1] - Xs[p][..., 0], Xo[..., 1] - Xo[..., 0]
            assert np.abs(w - wo).max() <= 1e-3 * wo.max(), (rep, p)
    # a mix: caller's R for some problems, the device's for the others
    Rs = [np.asfortranarray(np.linalg.inv(_mid(A))) if p % 2 else None for p, A in enumerate(IAs)]
    Xs, gs = ctx.lss_batched(IAs, Ibs, Rs)
    for p in range(1, len(sizes), 2):
        Xo, go, ok = O.lss(IAs[p], Ibs[p], Rs[p])
        assert ok and go == gs[p] and np.array_equal(Xs[p], Xo), p


def test_lss_batched_device_inverses_over_lanes(ctx):
    """problems of five padded orders (64 ... 320): their getrf/getrs groups
    run side by side on the context stream and three more, and every
    problem's result is still its single solve's, bit for bit, on two calls
    in a row"""
    sizes = [5, 70, 130, 200, 260, 70, 5, 131]
    print("lss_batched device R over lanes", sizes, flush=True)
    rng = np.random.default_rng(78)
    sys = [_system(rng, n, 1) for n in sizes]
    IAs, Ibs = [s[0] for s in sys], [s[1] for s in sys]
    singles = [ctx.lss(IA, Ib) for IA, Ib in zip(IAs, Ibs)]
    for rep in range(2):
        Xs, gs = ctx.lss_batched(IAs, Ibs)
        for p in range(len(sizes)):
            X1, g1 = singles[p]
            assert g1 == gs[p] and np.array_equal(Xs[p], X1), (rep, p)
