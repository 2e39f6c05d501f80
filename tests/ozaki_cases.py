"""Inputs that stress the upward radius-digit bound of the Ozaki interval form.

The digit bound writes each radius-term factor as two 7-bit upward digits
against its row (A) or column (B) maximum.  Entries far below that maximum
have a zero high digit and a non-zero low digit, so the product of the two
low digits is the largest part of their contribution: a bound that drops or
underestimates that part loses containment on exactly these inputs.

  constructed  one 1 x 1000 point row (a_0 = 1, the rest 0.99 * 2/127) against a
               1000 x 1 column with zero midpoints, r_0 = 0.999 * 2^-10 and the
               rest 0.99 * 2^-10 / 127: the low-digit products carry about 11%
               of the exact radius
  wide_r10     mid = +-m 2^e, e in [-6, 0], radius 2^-10 |mid|, n = 256
  wide_r10_s7  the same distribution, n = 160, seed 7
  near_r2      e in [-3, 0], radius 2^-2 |mid|: large radii and midpoint
               products near 2^124 for 20 moduli (beta = 62)
"""
import numpy as np

from vcp_amd import workloads as W

NAMES = ["constructed", "wide_r10", "wide_r10_s7", "near_r2"]


def make(name):
    if name == "constructed":
        k = 1000
        am = np.full((1, k), 0.99 * 2.0 / 127.0)
        am[0, 0] = 1.0
        bm = np.zeros((k, 1))
        br = np.full((k, 1), 0.99 * 2.0 ** -10 / 127.0)
        br[0, 0] = 0.999 * 2.0 ** -10
        return W.interval_from_midrad(am, np.zeros_like(am)), W.interval_from_midrad(bm, br)
    if name == "wide_r10":
        rng = np.random.default_rng(0)
        return (W.wide_exponent_intervals(rng, 256, 256, -6, 0, rad_exp=-10),
                W.wide_exponent_intervals(rng, 256, 256, -6, 0, rad_exp=-10))
    if name == "wide_r10_s7":
        rng = np.random.default_rng(7)
        return (W.wide_exponent_intervals(rng, 160, 160, -6, 0, rad_exp=-10),
                W.wide_exponent_intervals(rng, 160, 160, -6, 0, rad_exp=-10))
    if name == "near_r2":
        rng = np.random.default_rng(3)
        return (W.wide_exponent_intervals(rng, 200, 300, -3, 0, rad_exp=-2),
                W.wide_exponent_intervals(rng, 300, 180, -3, 0, rad_exp=-2))
    raise ValueError(name)


def every_entry(C):
    m, n = C.shape[:2]
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    return ii.ravel(), jj.ravel()
