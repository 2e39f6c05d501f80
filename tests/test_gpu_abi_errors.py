"""Argument checks of the C entry points with a live context: bad sizes,
leading dimensions and null pointers return VCPHIP_ERR_DIMENSION (-1) or
VCPHIP_ERR_ARGUMENT (-2) before any device work, and empty problems return 0
without touching their (null) pointers."""
import ctypes

import pytest

import vcp_amd

pytestmark = pytest.mark.gpu

DIM, ARG = -1, -2


def _ia(*xs):
    return (ctypes.c_int * len(xs))(*xs)


def _vp(*xs):
    return (ctypes.c_void_p * len(xs))(*xs)


def test_solver_argument_checks(ctx):
    L, h = vcp_amd.lib(), ctx._h
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    g = ctypes.c_double()
    # verified solve: n < 2, short leading dimensions, null operands; R = NULL is allowed
    assert L.vcphip_lss(h, 1, 1, p, 1, p, 1, None, 1, p, 1, ctypes.byref(g)) == DIM
    assert L.vcphip_lss(h, 4, 1, p, 3, p, 4, None, 4, p, 4, ctypes.byref(g)) == DIM
    assert L.vcphip_lss(h, 4, 1, p, 4, p, 4, p, 3, p, 4, ctypes.byref(g)) == DIM
    assert L.vcphip_lss(h, 4, 1, None, 4, p, 4, None, 4, p, 4, ctypes.byref(g)) == ARG
    assert L.vcphip_lss_dev(h, 4, 0, p, 4, p, 4, None, 4, p, 4, None) == DIM
    # approximate inverse
    assert L.vcphip_inv_mid(h, -1, p, 1, p, 1) == DIM
    assert L.vcphip_inv_mid(h, 4, p, 3, p, 4) == DIM
    assert L.vcphip_inv_mid(h, 0, None, 1, None, 1) == 0
    assert L.vcphip_inv_mid_dev(h, 4, None, 4, p, 4) == ARG
    # dd solve, Cholesky, eigensolvers
    it = ctypes.c_int(7)
    assert L.vcphip_dd_solve(h, -1, 1, p, 1, p, 1, p, 1, 10, ctypes.byref(it)) == DIM
    assert L.vcphip_dd_solve(h, 4, 1, p, 4, p, 3, p, 4, 10, ctypes.byref(it)) == DIM
    assert L.vcphip_dd_solve(h, 4, 1, p, 4, p, 4, p, 4, -1, ctypes.byref(it)) == ARG
    assert L.vcphip_dd_solve(h, 0, 1, None, 1, None, 1, None, 1, 10, ctypes.byref(it)) == 0 and it.value == 0
    assert L.vcphip_dd_cholesky(h, 4, p, 3, p, 4, 2) == DIM
    assert L.vcphip_dd_cholesky_dev(h, 4, None, 4, p, 4, 2) == ARG
    assert L.vcphip_dd_eigsym(h, 4, p, 4, p, p, 3, 4) == DIM
    assert L.vcphip_dd_eigsymge(h, 4, p, 4, None, 4, p, p, 4, 1) == ARG
    assert L.vcphip_dd_eigsymge_dev(h, 4, p, 4, p, 3, p, p, 4, 1) == DIM


def test_batched_argument_checks(ctx):
    L, h = vcp_amd.lib(), ctx._h
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    one = _ia(4)
    # count < 0, null arrays, a short leading dimension, a null operand of a non-empty problem
    assert L.vcphip_igemm_batched(h, -1, None, None, None, None, None, None, None, None, None) == ARG
    assert L.vcphip_igemm_batched(h, 1, one, one, one, None, one, _vp(p), one, _vp(p), one) == ARG
    assert L.vcphip_igemm_batched_dev(h, 1, one, one, one, _vp(p), _ia(3), _vp(p), one, _vp(p), one) == DIM
    assert L.vcphip_igemm_batched_dev(h, 1, one, one, one, _vp(None), one, _vp(p), one, _vp(p), one) == ARG
    assert L.vcphip_igemm_batched(h, 0, None, None, None, None, None, None, None, None, None) == 0
    # an empty problem (m = 0) needs no pointers
    z = _ia(0)
    assert L.vcphip_igemm_batched_dev(h, 1, z, one, one, _vp(None), _ia(1), _vp(None), one, _vp(None), _ia(1)) == 0
    ctx.synchronize()
    # batched solve: sizes below 2 and short leading dimensions
    g = (ctypes.c_double * 1)()
    assert L.vcphip_lss_batched(h, 1, _ia(1), _ia(1), _vp(p), _ia(1), _vp(p), _ia(1), None, None, _vp(p), _ia(1),
                                g) == DIM
    assert L.vcphip_lss_batched_dev(h, 1, one, _ia(1), _vp(p), _ia(3), _vp(p), one, None, None, _vp(p), one, g) == DIM
