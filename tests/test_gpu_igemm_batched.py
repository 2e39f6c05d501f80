"""Batched interval products C_p = IA_p * IB_p of uneven sizes in one
persistent launch (vcphip_igemm_batched: the plain form of pidblas::mulmm,
vcp/pidblas.hpp:239-294, per problem).  Every C_p must be bit-identical to the
single plain product and to the oracle's restatement; the graph-captured
device form must give the same results on every replay and after its problem
set changes."""
import numpy as np
import pytest
import torch

import oracle_ffi as O
import vcp_amd
from vcp_amd import workloads as W

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (3, 70, 5), (64, 64, 64), (65, 130, 257), (200, 33, 7), (300, 300, 300), (0, 5, 5),
          (7, 0, 3), (9, 4, 0), (129, 65, 1)]


def _set(rng, shapes):
    IAs = [W.wide_exponent_intervals(rng, m, k) for m, n, k in shapes]
    IBs = [W.wide_exponent_intervals(rng, k, n) for m, n, k in shapes]
    return IAs, IBs


def _dev(X):
    return torch.from_numpy(np.ascontiguousarray(X.transpose(1, 0, 2))).cuda()


def test_batched_equals_single_and_oracle(ctx):
    print("igemm_batched", SHAPES, flush=True)
    rng = np.random.default_rng(71)
    IAs, IBs = _set(rng, SHAPES)
    Cs = ctx.igemm_batched(IAs, IBs)
    for p, (m, n, k) in enumerate(SHAPES):
        C1 = ctx.igemm(IAs[p], IBs[p], vcp_amd.FORM_PLAIN)
        assert np.array_equal(Cs[p], C1), p
        if m * n * k and m * n * k <= 300 * 300 * 300:
            assert np.array_equal(Cs[p], O.pidblas_mulmm(IAs[p], IBs[p])), p


def test_batched_dev_graph_replays_and_recapture(ctx):
    rng = np.random.default_rng(72)
    shapes = [(130, 90, 70), (64, 257, 33), (17, 17, 400)]
    IAs, IBs = _set(rng, shapes)
    dA, dB = [_dev(A) for A in IAs], [_dev(B) for B in IBs]
    dC = [torch.zeros((n, m, 2), dtype=torch.float64, device="cuda") for m, n, k in shapes]
    args = ([s[0] for s in shapes], [s[1] for s in shapes], [s[2] for s in shapes],
            [t.data_ptr() for t in dA], [s[0] for s in shapes], [t.data_ptr() for t in dB], [s[2] for s in shapes],
            [t.data_ptr() for t in dC], [s[0] for s in shapes])
    ref = [O.pidblas_mulmm(A, B) for A, B in zip(IAs, IBs)]
    torch.cuda.synchronize()
    for rep in range(3):   # capture, then two replays, with the outputs cleared between them
        for t in dC:
            t.zero_()
        torch.cuda.synchronize()
        ctx.igemm_batched_dev(*args)
        ctx.synchronize()
        for p in range(len(shapes)):
            assert np.array_equal(dC[p].cpu().numpy().transpose(1, 0, 2), ref[p]), (rep, p)
    # a different problem set (one size changed): a new tile list and graph
    shapes2 = [(130, 90, 70), (64, 200, 33), (17, 17, 400)]
    IB1 = W.wide_exponent_intervals(rng, 33, 200)
    dB1 = _dev(IB1)
    dC1 = torch.zeros((200, 64, 2), dtype=torch.float64, device="cuda")
    args2 = list(args)
    args2[1] = [s[1] for s in shapes2]
    args2[5] = [dB[0].data_ptr(), dB1.data_ptr(), dB[2].data_ptr()]
    args2[7] = [dC[0].data_ptr(), dC1.data_ptr(), dC[2].data_ptr()]
    torch.cuda.synchronize()
    ctx.igemm_batched_dev(*args2)
    ctx.synchronize()
    assert np.array_equal(dC1.cpu().numpy().transpose(1, 0, 2), O.pidblas_mulmm(IAs[1], IB1))
    assert np.array_equal(dC[0].cpu().numpy().transpose(1, 0, 2), ref[0])


def test_batched_galerkin_block_sizes(ctx):
    """48 square products of the batched-residual size distribution, each
    against the single plain product"""
    rng = np.random.default_rng(73)
    sizes = W.config5_sizes(48)
    IAs = [W.uniform_intervals(rng, n, n, -40) for n in sizes]
    IBs = [W.uniform_intervals(rng, n, n, -40) for n in sizes]
    Cs = ctx.igemm_batched(IAs, IBs)
    for p in range(len(sizes)):
        assert np.array_equal(Cs[p], ctx.igemm(IAs[p], IBs[p], vcp_amd.FORM_PLAIN)), p


@pytest.mark.parametrize("mode", [vcp_amd.BATCH_TILES_64, vcp_amd.BATCH_TILES_EDGE, vcp_amd.BATCH_TILES_32])
def test_batched_every_tile_shape_bitexact(ctx, mode):
    """the 64 x 64 and 32 x 64 tiles run the same chains for every output:
    each shape gives the single product's bits, on row counts below, at and
    above multiples of 32 and 64"""
    shapes = SHAPES + [(31, 40, 50), (32, 33, 17), (33, 64, 9), (96, 70, 40), (97, 5, 130), (160, 129, 65)]
    print("igemm_batched tile shape", vcp_amd.BATCH_TILES_NAMES[mode], flush=True)
    rng = np.random.default_rng(75)
    IAs, IBs = _set(rng, shapes)
    ctx.set_param(vcp_amd.PARAM_BATCH_TILES, mode)
    try:
        assert ctx.get_param(vcp_amd.PARAM_BATCH_TILES) == mode
        Cs = ctx.igemm_batched(IAs, IBs)
    finally:
        ctx.set_param(vcp_amd.PARAM_BATCH_TILES, 0)
    for p, (m, n, k) in enumerate(shapes):
        assert np.array_equal(Cs[p], ctx.igemm(IAs[p], IBs[p], vcp_amd.FORM_PLAIN)), (mode, p)
        if m * n * k and m * n * k <= 300 * 300 * 300:
            assert np.array_equal(Cs[p], O.pidblas_mulmm(IAs[p], IBs[p])), (mode, p)


def test_batch_tiles_param_rejects_unknown_shapes(ctx):
    for bad in (-1, 4):
        with pytest.raises(vcp_amd.VcphipError):
            ctx.set_param(vcp_amd.PARAM_BATCH_TILES, bad)
    assert ctx.get_param(vcp_amd.PARAM_BATCH_TILES) == 0


def test_tune_batched_shapes_then_replays_bitexact(ctx, tmp_path):
    """the shape tuner times the set under each shape per size class, keeps
    the fastest, and the graph captured with its choice gives the single
    products' bits; the choice survives a save / load"""
    rng = np.random.default_rng(76)
    sizes = [65, 70, 90, 97, 130, 150, 200, 260, 300, 33]
    IAs = [W.uniform_intervals(rng, n, n, -40) for n in sizes]
    IBs = [W.uniform_intervals(rng, n, n, -40) for n in sizes]
    dA, dB = [_dev(A) for A in IAs], [_dev(B) for B in IBs]
    dC = [torch.zeros((n, n, 2), dtype=torch.float64, device="cuda") for n in sizes]
    args = (sizes, sizes, sizes, [t.data_ptr() for t in dA], sizes, [t.data_ptr() for t in dB], sizes,
            [t.data_ptr() for t in dC], sizes)
    torch.cuda.synchronize()
    ms, padded = ctx.tune_igemm_batched_dev(*args)
    print("tuned batched shapes: %.3f ms, padded fraction %.4f" % (ms, padded), flush=True)
    assert ms > 0.0 and 0.0 <= padded < 0.5
    ref = [ctx.igemm(A, B, vcp_amd.FORM_PLAIN) for A, B in zip(IAs, IBs)]
    for rep in range(3):   # capture with the tuned shapes, then replays
        for t in dC:
            t.zero_()
        torch.cuda.synchronize()
        ctx.igemm_batched_dev(*args)
        ctx.synchronize()
        for p in range(len(sizes)):
            assert np.array_equal(dC[p].cpu().numpy().transpose(1, 0, 2), ref[p]), (rep, p)
    path = tmp_path / "tune.txt"
    ctx.tune_save(str(path))
    lines = [ln for ln in path.read_text().splitlines() if ln.startswith("tiles igemm")]
    assert lines, path.read_text()
    with vcp_amd.Context(0) as c2:
        c2.tune_load(str(path))
        Cs = c2.igemm_batched(IAs, IBs)
    for p in range(len(sizes)):
        assert np.array_equal(Cs[p], ref[p]), p


def test_batched_rejects_bad_dimensions(ctx):
    rng = np.random.default_rng(74)
    A, B = W.wide_exponent_intervals(rng, 5, 6), W.wide_exponent_intervals(rng, 6, 7)
    L = vcp_amd.lib()
    import ctypes
    ia = lambda xs: (ctypes.c_int * 1)(*xs)
    vp = lambda xs: (ctypes.c_void_p * 1)(*xs)
    C = O.pairs(5, 7)
    rc = L.vcphip_igemm_batched(ctx._h, 1, ia([5]), ia([7]), ia([6]), vp([A.ctypes.data]), ia([4]),
                                vp([B.ctypes.data]), ia([6]), vp([C.ctypes.data]), ia([5]))
    assert rc == -1   # lda < m
