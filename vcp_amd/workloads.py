"""Synthetic inputs of the benchmarked workloads (seeded, reproducible).

Each generator returns host arrays in VCP's layout (column-major; intervals
as AoS {inf, sup}, dd as AoS {hi, lo}).  Seeds are 0x5EED0000 + workload id.

  1  interval n=1024: mid U[-1,1], radius 2^-40 |mid|
  2  point n=4096: A, B U[-1,1] (RU/RD pair), plus the interval product
     with radius 2^-40 |mid|
  3  interval n=8192: mid = +-m 2^e, m U[1,2), e U{-20..20}, radius 2^-45 |mid|
  4  dd n=4096: hi U[-1,1], lo = hi 2^-53 u, u U[-1,1] (renormalised)
  5  256 residual problems R*A - I, n log-uniform in [64, 1024],
     A = n^2 tridiag(-1,2,-1) + 0.01 G G^T / n, R ~ inv(mid A), rad A = 2^-50 |A|
"""
import numpy as np

SEED0 = 0x5EED0000


def _pairs(rows, cols):
    return np.zeros((cols, rows, 2), dtype=np.float64).transpose(1, 0, 2)


def interval_from_midrad(mid, rad):
    """[mid - rad, mid + rad] with endpoints rounded outward by one ulp when
    the nearest-rounded sum is inexact (so the stated radius is contained)."""
    lo = mid - rad
    hi = mid + rad
    lo = np.where(lo + rad != mid, np.nextafter(lo, -np.inf), lo)
    hi = np.where(hi - rad != mid, np.nextafter(hi, np.inf), hi)
    out = _pairs(*mid.shape)
    out[..., 0] = lo
    out[..., 1] = hi
    return out


def uniform_intervals(rng, rows, cols, rad_exp):
    mid = rng.uniform(-1.0, 1.0, (rows, cols))
    return interval_from_midrad(mid, np.ldexp(np.abs(mid), rad_exp))


def wide_exponent_intervals(rng, rows, cols, emin=-20, emax=20, rad_exp=-45):
    mant = rng.uniform(1.0, 2.0, (rows, cols))
    sign = np.where(rng.integers(0, 2, (rows, cols)) == 1, -1.0, 1.0)
    e = rng.integers(emin, emax + 1, (rows, cols))
    mid = sign * np.ldexp(mant, e)
    return interval_from_midrad(mid, np.ldexp(np.abs(mid), rad_exp))


def config1(n=1024):
    rng = np.random.default_rng(SEED0 + 1)
    return uniform_intervals(rng, n, n, -40), uniform_intervals(rng, n, n, -40)


def config2(n=4096):
    rng = np.random.default_rng(SEED0 + 2)
    A = np.asfortranarray(rng.uniform(-1.0, 1.0, (n, n)))
    B = np.asfortranarray(rng.uniform(-1.0, 1.0, (n, n)))
    return A, B


def config3(n=8192):
    rng = np.random.default_rng(SEED0 + 3)
    return wide_exponent_intervals(rng, n, n), wide_exponent_intervals(rng, n, n)


def dd_matrix(rng, rows, cols):
    hi = rng.uniform(-1.0, 1.0, (rows, cols))
    lo = hi * 2.0 ** -53 * rng.uniform(-1.0, 1.0, (rows, cols))
    s = hi + lo                      # FastTwoSum renormalisation
    e = lo - (s - hi)
    out = _pairs(rows, cols)
    out[..., 0] = s
    out[..., 1] = e
    return out


def config4(n=4096):
    rng = np.random.default_rng(SEED0 + 4)
    return dd_matrix(rng, n, n), dd_matrix(rng, n, n)


def residual_problem(rng, n):
    G = rng.standard_normal((n, n))
    A = n * n * (2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)) + 0.01 * (G @ G.T) / n
    R = np.asfortranarray(np.linalg.inv(A))
    IA = interval_from_midrad(A, np.ldexp(np.abs(A), -50))
    return R, IA


def config5_sizes(count=256, seed=SEED0 + 5):
    rng = np.random.default_rng(seed)
    return [int(round(np.exp(rng.uniform(np.log(64), np.log(1024))))) for _ in range(count)]


def config5(count=256):
    sizes = config5_sizes(count)
    rng = np.random.default_rng(SEED0 + 50)
    return [residual_problem(rng, n) for n in sizes]


def galerkin(p, q, rad_exp=None):
    """Operands of Legendre_Bases_Generator::uhphiphi for a 2D tensor basis:
    phi(i, k) = f_{i1}(x_{k1}) f_{i2}(x_{k2}) with the Dirichlet-type Legendre
    functions f_n = (P_{n+1} - P_{n-1}) / sqrt(2 (2n + 1)), n = 1..p, at
    q x q Gauss-Legendre points of [-1, 1]^2; w_k the product weights and
    u_k = sin(pi x) cos(pi y / 2) + 2 at the points.  Returns phi (p^2, q^2, 2)
    and w, u (q^2, 2) as intervals (point intervals, or radii 2^rad_exp |x|)."""
    from numpy.polynomial import legendre as L

    x, wt = L.leggauss(q)
    f = np.zeros((p, q))
    for n in range(1, p + 1):
        cp = np.zeros(n + 2)
        cp[n + 1] = 1.0
        cm = np.zeros(n + 2)
        cm[n - 1] = 1.0
        f[n - 1] = (L.legval(x, cp) - L.legval(x, cm)) / np.sqrt(2.0 * (2 * n + 1))
    phi_v = np.einsum("ak,bl->abkl", f, f).reshape(p * p, q * q)
    X, Y = np.meshgrid(x, x, indexing="ij")
    w_v = np.outer(wt, wt).reshape(-1)
    u_v = (np.sin(np.pi * X) * np.cos(0.5 * np.pi * Y) + 2.0).reshape(-1)

    def iv(v):
        rad = np.zeros_like(v) if rad_exp is None else np.ldexp(np.abs(v), rad_exp)
        return interval_from_midrad(v.reshape(v.shape[0], -1), rad.reshape(v.shape[0], -1))

    phi = iv(phi_v)
    w = iv(w_v.reshape(-1, 1)).reshape(-1, 2)
    u = iv(u_v.reshape(-1, 1)).reshape(-1, 2)
    return phi, np.ascontiguousarray(w), np.ascontiguousarray(u)
