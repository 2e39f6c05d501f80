"""The GPU Galerkin matrix (vcphip_uhphiphi, Legendre_Bases_Generator::
uhphiphi at vcp/ldbase.hpp:955-1013) bit-identical to the oracle: random
interval data over several tiles and partial tiles, 2D Legendre bases at
Gauss points (point and interval data), empty and single-point cases, and
the device entry with padded leading dimensions."""
import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _rand(rng, m, npts, rad=2.0 ** -30):
    phi = O.pairs(m, npts)
    v = rng.uniform(-1, 1, (m, npts))
    phi[..., 0] = v - rad * np.abs(v)
    phi[..., 1] = v + rad * np.abs(v)
    wv, uv = rng.uniform(0, 1, npts), rng.uniform(-2, 2, npts)
    w = np.stack([wv, wv * (1 + 2.0 ** -40)], axis=-1)
    u = np.stack([uv - 2.0 ** -35, uv + 2.0 ** -35], axis=-1)
    return phi, w, u


@pytest.mark.parametrize("m,npts", [(1, 1), (3, 0), (64, 16), (65, 17), (130, 70), (200, 33)])
def test_uhphiphi_bitexact_random(ctx, m, npts):
    print("uhphiphi", m, npts, flush=True)
    rng = np.random.default_rng(m * 1000 + npts)
    phi, w, u = _rand(rng, m, npts)
    A = ctx.uhphiphi(phi, w, u)
    assert np.array_equal(A, O.uhphiphi(phi, w, u))


@pytest.mark.parametrize("p,q,rad", [(6, 9, None), (9, 14, -45)])
def test_uhphiphi_legendre_2d(ctx, p, q, rad):
    phi, w, u = W.galerkin(p, q, rad)
    A = ctx.uhphiphi(phi, w, u)
    assert np.array_equal(A, O.uhphiphi(phi, w, u))
    assert (A[..., 0] <= A[..., 1]).all()


def test_uhphiphi_dev_entry(ctx):
    rng = np.random.default_rng(3)
    m, npts = 70, 40
    phi, w, u = _rand(rng, m, npts)
    ldp, lda = m + 3, m + 1
    dP = torch.zeros((npts, ldp, 2), dtype=torch.float64, device="cuda")
    dP[:, :m, :] = torch.from_numpy(np.ascontiguousarray(phi.transpose(1, 0, 2)))
    dW = torch.from_numpy(np.ascontiguousarray(w)).cuda()
    dU = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    dA = torch.full((m, lda, 2), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert vcp_amd.lib().vcphip_uhphiphi_dev(ctx._h, m, npts, dP.data_ptr(), ldp, dW.data_ptr(), dU.data_ptr(),
                                             dA.data_ptr(), lda) == 0
    ctx.synchronize()
    Ad = dA.cpu().numpy()
    assert np.array_equal(Ad[:, :m, :].transpose(1, 0, 2), O.uhphiphi(phi, w, u))
    assert np.isnan(Ad[:, m:, :]).all()


def test_uhphiphi_large_tiles_sampled(ctx):
    """m = 2880: the 64 x 64 tile path; entries of 24 rows spread over every
    tile row against the oracle on those rows, and exact symmetry"""
    rng = np.random.default_rng(4)
    m, npts = 2880, 12
    phi, w, u = _rand(rng, m, npts)
    A = ctx.uhphiphi(phi, w, u)
    idx = np.linspace(0, m - 1, 24).astype(int)
    sub = O.pairs(len(idx), npts)
    sub[...] = phi[idx]
    assert np.array_equal(A[np.ix_(idx, idx)], O.uhphiphi(sub, w, u))
    assert np.array_equal(A, A.transpose(1, 0, 2))


def _with_zero_crossings(rng, phi, u, points):
    """Entries of phi at the given points that straddle or touch zero
    ([-a, b], [0, b], [-b, 0], [-0, 0], [0, 0]) and one solution value
    u = [0, 0]: the stages holding them take the kernel's general product,
    the others its two-product form."""
    m = phi.shape[0]
    for k in points:
        sel = rng.random(m) < 0.3
        n = int(sel.sum())
        a, b = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
        kind = rng.integers(0, 5, n)
        lo = np.choose(kind, [-a, np.zeros(n), -b, np.full(n, -0.0), np.zeros(n)])
        hi = np.choose(kind, [b, b, np.zeros(n), np.zeros(n), np.zeros(n)])
        phi[sel, k, 0] = lo
        phi[sel, k, 1] = hi
    u[points[0]] = 0.0
    return phi, u


@pytest.mark.parametrize("m,npts,tag", [(130, 70, "32x32 tiles"), (2880, 40, "64x64 tiles")])
def test_uhphiphi_two_product_and_general_stages(ctx, m, npts, tag):
    """Stages of 16 points whose operands all lie strictly on one side of
    zero use two products per interval product, the rest the general eight;
    both must give the oracle's bits, in one matrix.  Zero crossings sit at
    points 20-23 and at the last points, so stages 0, 2 and 3 take the
    two-product form and stages 1 and 4 the general one."""
    print("uhphiphi mixed stages", m, npts, tag, flush=True)
    rng = np.random.default_rng(m + npts)
    phi, w, u = _rand(rng, m, npts)
    # negative solution values too: every sign combination of s and phi
    u[::3] = -u[::3, ::-1]
    phi, u = _with_zero_crossings(rng, phi, u, [20, 21, 22, 23, npts - 2, npts - 1])
    A = ctx.uhphiphi(phi, w, u)
    if m <= 256:
        assert np.array_equal(A, O.uhphiphi(phi, w, u))
    else:
        idx = np.linspace(0, m - 1, 24).astype(int)
        sub = O.pairs(len(idx), npts)
        sub[...] = phi[idx]
        assert np.array_equal(A[np.ix_(idx, idx)], O.uhphiphi(sub, w, u))
        assert np.array_equal(A, A.transpose(1, 0, 2))


def test_uhphiphi_two_product_signed_zero_results(ctx):
    """Products that underflow to zero in the two-product form: tiny operands
    of every sign combination, whose rounded products are -0, +0 or the
    smallest subnormal, bit-identical to the oracle (signed zeros included)."""
    rng = np.random.default_rng(11)
    m, npts = 40, 33
    v = rng.uniform(0.5, 1, (m, npts)) * 2.0 ** -540 * rng.choice([-1.0, 1.0], (m, npts))
    phi = O.pairs(m, npts)
    phi[..., 0] = np.where(v > 0, v, v * (1 + 2.0 ** -20))
    phi[..., 1] = np.where(v > 0, v * (1 + 2.0 ** -20), v)
    wv = rng.uniform(0.5, 1, npts)
    w = np.stack([wv, wv], axis=-1)
    uv = rng.uniform(0.5, 1, npts) * rng.choice([-1.0, 1.0], npts)
    u = np.stack([uv, uv], axis=-1)
    A = ctx.uhphiphi(phi, w, u)
    R = O.uhphiphi(phi, w, u)
    assert np.array_equal(A, R)
    assert np.array_equal(np.signbit(A), np.signbit(R))
