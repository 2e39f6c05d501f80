"""Batched verified residual R_p * IA_p - I with the ||.||_inf bound (the
R*A - I segment of imats::linearsolve) in one persistent, graph-captured
launch: bit-identical to the oracle's restatement, graph replays identical
to the first (eager-built) result."""
import numpy as np
import pytest

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu

SIZES = [1, 5, 64, 65, 100, 130, 200, 17, 64, 3]


def _problems(sizes, seed):
    rng = np.random.default_rng(seed)
    return [W.residual_problem(rng, n) for n in sizes]


def test_resid_bitexact_against_oracle(ctx):
    probs = _problems(SIZES, 1)
    Cs, nub = ctx.residual_batched([p[0] for p in probs], [p[1] for p in probs])
    for (R, IA), C, nb in zip(probs, Cs, nub):
        Co, nbo = O.residual(R, IA)
        assert np.array_equal(C, Co)
        assert nb == nbo and nb < 1.0


@pytest.mark.parametrize("mode", [vcp_amd.BATCH_TILES_64, vcp_amd.BATCH_TILES_EDGE, vcp_amd.BATCH_TILES_32])
def test_resid_every_tile_shape_bitexact(ctx, mode):
    print("residual tile shape", vcp_amd.BATCH_TILES_NAMES[mode], flush=True)
    probs = _problems(SIZES + [31, 32, 33, 96, 97, 160], 4)
    ctx.set_param(vcp_amd.PARAM_BATCH_TILES, mode)
    try:
        Cs, nub = ctx.residual_batched([p[0] for p in probs], [p[1] for p in probs])
    finally:
        ctx.set_param(vcp_amd.PARAM_BATCH_TILES, 0)
    for (R, IA), C, nb in zip(probs, Cs, nub):
        Co, nbo = O.residual(R, IA)
        assert np.array_equal(C, Co)
        assert nb == nbo and nb < 1.0


def test_resid_encloses_exact_product(ctx):
    probs = _problems([40, 77], 2)
    Cs, nub = ctx.residual_batched([p[0] for p in probs], [p[1] for p in probs])
    for (R, IA), C, nb in zip(probs, Cs, nub):
        n = R.shape[0]
        IR = W.interval_from_midrad(R, np.zeros_like(R))
        # C + I with the diagonal widened outward by one ulp must contain R*IA
        S = vcp_amd.as_pairs(C)
        d = np.arange(n)
        S[d, d, 0] = np.nextafter(C[d, d, 0] + 1.0, -np.inf)
        S[d, d, 1] = np.nextafter(C[d, d, 1] + 1.0, np.inf)
        ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        ok, _, _ = O.check_containment(IR, IA, S, ii.ravel(), jj.ravel())
        assert ok.all()
        mag = np.maximum(np.abs(C[..., 0]), np.abs(C[..., 1]))
        assert nb >= mag.sum(axis=1).max() * (1 - 1e-12)


def test_resid_graph_replays_match(ctx):
    torch = pytest.importorskip("torch")
    sizes = [64, 96, 33, 128]
    probs = _problems(sizes, 3)
    dev = torch.device("cuda:0")
    dR = [torch.from_numpy(np.ascontiguousarray(R.T)).to(dev) for R, _ in probs]
    dA = [torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2))).to(dev) for _, IA in probs]
    dC = [torch.empty((n, n, 2), dtype=torch.float64, device=dev) for n in sizes]
    nub = torch.empty(len(sizes), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    args = (sizes, [t.data_ptr() for t in dR], sizes, [t.data_ptr() for t in dA], sizes,
            [t.data_ptr() for t in dC], sizes, nub.data_ptr())
    vcp_amd.residual_batched_dev(ctx, *args)
    ctx.synchronize()
    first = [c.cpu().clone() for c in dC]
    first_n = nub.cpu().clone()
    for rep in range(5):
        for c in dC:
            c.zero_()            # eager work between replays
        nub.zero_()
        torch.cuda.synchronize()
        vcp_amd.residual_batched_dev(ctx, *args)
        ctx.synchronize()
        for c, f in zip(dC, first):
            assert torch.equal(c.cpu(), f)
        assert torch.equal(nub.cpu(), first_n)
    for (R, IA), f, nb in zip(probs, first, first_n.tolist()):
        Co, nbo = O.residual(R, IA)
        assert np.array_equal(f.numpy().transpose(1, 0, 2), Co) and nb == nbo


def test_resid_empty_batch(ctx):
    Cs, nub = ctx.residual_batched([], [])
    assert Cs == [] and nub == []
