"""Parity of the verified linear solve (vcphip_lss, imats::linearsolve at
vcp/imats.hpp:292-362 after its LAPACK step) with the oracle's restatement:
bit-identical enclosures and norm bounds, the exact-inverse known answer, the
reference's verification failure, device pointers with padded leading
dimensions, tiny right-hand sides (the scaled path of the directed division)
and problems of the batched residual workload."""
import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _system(rng, n, nrhs, rad_exp=-30, scale=1.0):
    A = rng.uniform(-1, 1, (n, n)) + 0.5 * n * np.eye(n)
    b = rng.uniform(-1, 1, (n, nrhs)) * scale
    IA = W.interval_from_midrad(A, np.ldexp(np.abs(A), rad_exp))
    Ib = W.interval_from_midrad(b, np.ldexp(np.abs(b), rad_exp))
    return IA, Ib, np.asfortranarray(np.linalg.inv(A))


def _eq(X, Y):
    return np.array_equal(X, Y)


@pytest.mark.parametrize("n,nrhs", [(2, 1), (3, 2), (33, 3), (64, 1), (100, 1), (257, 2), (520, 1)])
def test_lss_bitexact_against_oracle(ctx, n, nrhs):
    print("lss", n, nrhs, flush=True)
    rng = np.random.default_rng(1000 * n + nrhs)
    IA, Ib, R = _system(rng, n, nrhs)
    Xo, go, ok = O.lss(IA, Ib, R)
    assert ok
    X, g = ctx.lss(IA, Ib, R)
    assert g == go
    assert _eq(X, Xo)
    assert (X[..., 0] <= X[..., 1]).all()


def test_lss_exact_inverse_known_answer(ctx):
    n = 15
    T = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    R = np.asfortranarray([[min(i, j) * (n + 1 - max(i, j)) / (n + 1) for j in range(1, n + 1)]
                           for i in range(1, n + 1)])
    b = np.arange(1.0, n + 1).reshape(n, 1)
    IA, Ib = W.interval_from_midrad(T, np.zeros_like(T)), W.interval_from_midrad(b, np.zeros_like(b))
    X, g = ctx.lss(IA, Ib, R)
    assert g == 0.0
    assert np.array_equal(X[..., 0], R @ b) and np.array_equal(X[..., 1], R @ b)


def test_lss_verification_failure(ctx):
    n = 40
    rng = np.random.default_rng(7)
    IA, Ib, R = _system(rng, n, 1)
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.lss(IA, Ib, np.asfortranarray(2.5 * R))
    assert e.value.code == -6
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.lss(IA, Ib, np.zeros((n, n), order="F"))
    assert e.value.code == -6
    # the context stays usable
    X, g = ctx.lss(IA, Ib, R)
    assert _eq(X, O.lss(IA, Ib, R)[0])


def test_lss_rejects_scalar_systems(ctx):
    one = W.interval_from_midrad(np.ones((1, 1)), np.zeros((1, 1)))
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.lss(one, one, np.ones((1, 1), order="F"))
    assert e.value.code == -1


def test_lss_tiny_right_hand_side(ctx):
    """||y|| below 2^-960: the directed quotient takes its scaled path"""
    rng = np.random.default_rng(8)
    IA, Ib, R = _system(rng, 50, 2, scale=2.0 ** -1000)
    Xo, go, ok = O.lss(IA, Ib, R)
    assert ok
    X, g = ctx.lss(IA, Ib, R)
    assert g == go and _eq(X, Xo)


def test_lss_dev_entry_and_leading_dimensions(ctx):
    rng = np.random.default_rng(9)
    n, nrhs = 90, 2
    IA, Ib, R = _system(rng, n, nrhs)
    Xo, go, _ = O.lss(IA, Ib, R)
    lda, ldb, ldr, ldx = n + 3, n + 1, n + 5, n + 2
    dA = torch.zeros((n, lda, 2), dtype=torch.float64, device="cuda")
    dA[:, :n, :] = torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2)))
    dB = torch.zeros((nrhs, ldb, 2), dtype=torch.float64, device="cuda")
    dB[:, :n, :] = torch.from_numpy(np.ascontiguousarray(Ib.transpose(1, 0, 2)))
    dR = torch.zeros((n, ldr), dtype=torch.float64, device="cuda")
    dR[:, :n] = torch.from_numpy(np.ascontiguousarray(R.T))
    dX = torch.full((nrhs, ldx, 2), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g = ctx.lss_dev(n, nrhs, dA.data_ptr(), lda, dB.data_ptr(), ldb, dR.data_ptr(), ldr, dX.data_ptr(), ldx)
    Xd = dX.cpu().numpy()
    assert g == go
    assert _eq(Xd[:, :n, :].transpose(1, 0, 2), Xo)
    assert np.isnan(Xd[:, n:, :]).all()


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_lss_on_residual_workload_problems(ctx, idx):
    """problems of the batched residual workload (A = n^2 tridiag + 0.01 G G^T / n)"""
    sizes = [64, 181, 400]
    rng = np.random.default_rng(W.SEED0 + 60 + idx)
    R, IA = W.residual_problem(rng, sizes[idx])
    b = rng.uniform(-1, 1, (sizes[idx], 1))
    Ib = W.interval_from_midrad(b, np.ldexp(np.abs(b), -50))
    Xo, go, ok = O.lss(IA, Ib, R)
    assert ok
    X, g = ctx.lss(IA, Ib, R)
    assert g == go and _eq(X, Xo)


# ---- batched form -----------------------------------------------------------
def _mixed_set(seed, sizes, nrhss):
    rng = np.random.default_rng(seed)
    sys = [_system(rng, n, r) for n, r in zip(sizes, nrhss)]
    return [s[0] for s in sys], [s[1] for s in sys], [s[2] for s in sys]


def test_lss_batched_bitexact_against_oracle_and_single(ctx):
    sizes, nrhss = [2, 3, 17, 64, 100, 257, 31], [1, 2, 1, 3, 1, 2, 1]
    print("lss_batched", sizes, flush=True)
    IAs, Ibs, Rs = _mixed_set(11, sizes, nrhss)
    for rep in range(2):   # the second call replays the residual's graph
        Xs, gs = ctx.lss_batched(IAs, Ibs, Rs)
        for p in range(len(sizes)):
            Xo, go, ok = O.lss(IAs[p], Ibs[p], Rs[p])
            assert ok and gs[p] == go, (rep, p)
            assert _eq(Xs[p], Xo), (rep, p)
            Xsingle, gsingle = ctx.lss(IAs[p], Ibs[p], Rs[p])
            assert _eq(Xs[p], Xsingle) and gsingle == gs[p]


def test_lss_batched_residual_workload(ctx):
    """48 problems of the batched residual workload's size distribution,
    each against the single-problem form (bit-identical), a few against
    the oracle"""
    sizes = W.config5_sizes(48)
    rng = np.random.default_rng(W.SEED0 + 70)
    probs = [W.residual_problem(rng, n) for n in sizes]
    Rs, IAs = [p[0] for p in probs], [p[1] for p in probs]
    Ibs = []
    for n in sizes:
        b = rng.uniform(-1, 1, (n, 1))
        Ibs.append(W.interval_from_midrad(b, np.ldexp(np.abs(b), -50)))
    Xs, gs = ctx.lss_batched(IAs, Ibs, Rs)
    for p in range(len(sizes)):
        Xsingle, gsingle = ctx.lss(IAs[p], Ibs[p], Rs[p])
        assert gs[p] == gsingle and _eq(Xs[p], Xsingle), p
    for p in np.argsort(sizes)[:3]:
        Xo, go, ok = O.lss(IAs[p], Ibs[p], Rs[p])
        assert ok and go == gs[p] and _eq(Xs[p], Xo)


def test_lss_batched_one_failing_problem(ctx):
    sizes, nrhss = [20, 33, 40], [1, 1, 2]
    IAs, Ibs, Rs = _mixed_set(12, sizes, nrhss)
    Rs[1] = np.asfortranarray(2.5 * Rs[1])
    dA = [torch.from_numpy(np.ascontiguousarray(A.transpose(1, 0, 2))).cuda() for A in IAs]
    dB = [torch.from_numpy(np.ascontiguousarray(b.transpose(1, 0, 2))).cuda() for b in Ibs]
    dR = [torch.from_numpy(np.ascontiguousarray(R.T)).cuda() for R in Rs]
    dX = [torch.zeros((r, n, 2), dtype=torch.float64, device="cuda") for n, r in zip(sizes, nrhss)]
    torch.cuda.synchronize()
    with pytest.raises(vcp_amd.VcphipError) as e:
        ctx.lss_batched_dev(sizes, nrhss, [t.data_ptr() for t in dA], sizes, [t.data_ptr() for t in dB], sizes,
                            [t.data_ptr() for t in dR], sizes, [t.data_ptr() for t in dX], sizes)
    assert e.value.code == -6
    for p in (0, 2):
        Xo, go, ok = O.lss(IAs[p], Ibs[p], Rs[p])
        assert _eq(dX[p].cpu().numpy().transpose(1, 0, 2), Xo)
    assert np.isnan(dX[1].cpu().numpy()).all()
