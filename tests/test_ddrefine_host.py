"""The host double-double arithmetic of tests/ddrefine.py (the elementwise
steps between the GPU products of the refinement algorithms), against exact
rational arithmetic."""
from fractions import Fraction as F

import numpy as np

import ddrefine as D


def _dd(rng, n, spread=30):
    h = rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-spread, spread, n)
    l = h * 2.0 ** -53 * rng.uniform(-1, 1, n)
    return D.fast_two_sum(h, l)


def _v(h, l):
    return F(float(h)) + F(float(l))


def test_two_sum_and_two_prod_are_exact():
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, 500) * 2.0 ** rng.integers(-40, 40, 500)
    b = rng.uniform(-1, 1, 500) * 2.0 ** rng.integers(-40, 40, 500)
    s, e = D.two_sum(a, b)
    p, f = D.two_prod(a, b)
    for i in range(500):
        assert F(s[i]) + F(e[i]) == F(a[i]) + F(b[i])
        assert F(p[i]) + F(f[i]) == F(a[i]) * F(b[i])


def test_dd_add_mul_div_accuracy():
    rng = np.random.default_rng(2)
    xh, xl = _dd(rng, 400)
    yh, yl = _dd(rng, 400)
    ops = {
        "add": (D.dd_add(xh, xl, yh, yl), lambda x, y: x + y, lambda x, y: abs(x) + abs(y)),
        "mul": (D.dd_mul(xh, xl, yh, yl), lambda x, y: x * y, lambda x, y: abs(x * y)),
        "div": (D.dd_div(xh, xl, yh, yl), lambda x, y: x / y, lambda x, y: abs(x / y)),
    }
    for name, ((rh, rl), f, scale) in ops.items():
        for i in range(400):
            x, y = _v(xh[i], xl[i]), _v(yh[i], yl[i])
            err = abs(_v(rh[i], rl[i]) - f(x, y))
            assert err <= F(2) ** -102 * scale(x, y), (name, i)
            assert abs(rl[i]) <= 2.0 ** -52 * abs(rh[i]) or rh[i] == 0.0, (name, i)
