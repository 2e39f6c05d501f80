"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product path
(vcp_amd, libvcphip) never imports it.

Arrays follow the oracle's conventions: column-major, intervals and
double-doubles as AoS pairs, numpy float64 with Fortran order and a trailing
axis of length 2 for pairs (shape (rows, cols, 2), strides so that element
(i, j) pair sits at 2*(i + ld*j)).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_p_dbl = ctypes.POINTER(ctypes.c_double)
_p_int = ctypes.POINTER(ctypes.c_int)
_p_i64 = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_rgemm": [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _c_int, _c_int],
            "orc_rsyrk": [_c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _c_int],
            "orc_midrad": [_c_dbl, _c_dbl, _p_dbl, _p_dbl],
            "orc_pidblas_mulmm": [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int],
            "orc_pidblas_fma_mulmm": [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int],
            "orc_pidblas_mul_im_m": [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int],
            "orc_pidblas_mul_m_im": [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int],
            "orc_pidblas_mulltmm": [_c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int],
            "orc_pidblas_vmulmm": [_c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int],
            "orc_residual": [_c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl],
            "orc_imul": [_c_dbl, _c_dbl, _c_dbl, _c_dbl, _p_dbl, _p_dbl],
            "orc_uhphiphi": [_c_int, _c_int, _p_dbl, _c_int, _p_dbl, _p_dbl, _p_dbl, _c_int],
            "orc_lss": [_c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl],
            "orc_exact_hull_entry": [_c_int, _p_dbl, _c_int, _p_dbl, _c_int, _c_dbl, _c_dbl,
                                     _p_int, _p_int, _p_dbl, _p_dbl],
            "orc_exact_dot": [_c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _p_dbl, _p_dbl],
            "orc_exact_dd_dot": [_c_int, _p_dbl, _c_int, _p_dbl, _c_int, _c_dbl, _c_dbl,
                                 _p_dbl, _p_dbl, _p_dbl],
            "orc_check_containment": [_c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _c_int,
                                      _p_int, _p_int, _p_int, _p_dbl, _p_dbl],
            "orc_ozaki_moduli": [_c_int, _p_int],
            "orc_ozaki_beta": [_c_int, _c_int],
            "orc_ozaki_igemm": [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _c_int,
                                _c_int],
            "orc_ozaki_scaled_rows": [_c_int, _c_int, _p_dbl, _c_int, _c_int, _p_i64, _p_int],
            "orc_ozaki_scaled_cols": [_c_int, _c_int, _p_dbl, _c_int, _c_int, _p_i64, _p_int],
            "orc_dd_split_shift": [_c_int],
            "orc_dd_split_bits": [_c_int],
            "orc_dd_max_slices": [_c_int],
            "orc_pddblas_mulmm": [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int,
                                  _p_int, _p_int],
            "orc_pddblas_mulltmm": [_c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int],
            "orc_set_threads": [_c_int],
            "orc_get_threads": [],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None if name in ("orc_rgemm", "orc_rsyrk", "orc_midrad", "orc_exact_hull_entry",
                                         "orc_exact_dot", "orc_exact_dd_dot", "orc_check_containment",
                                         "orc_ozaki_scaled_rows", "orc_ozaki_scaled_cols", "orc_imul",
                                         "orc_set_threads") else _c_int
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous or a.size == 0)
    return a.ctypes.data_as(_p_dbl)


def _ld_pair(a):
    """leading dimension (in elements) of an (rows, cols, 2) pair array laid
    out as AoS column-major: pair (i, j) at 2*(i + ld*j)."""
    assert a.ndim == 3 and a.shape[2] == 2
    if a.size == 0:
        return max(a.shape[0], 1)
    assert a.strides[2] == 8 and a.strides[0] == 16
    return max(a.strides[1] // 16 if a.shape[1] > 1 else a.shape[0], 1)


def pairs(rows, cols):
    """zeroed AoS column-major pair matrix (rows, cols, 2)."""
    buf = np.zeros((cols, rows, 2), dtype=np.float64)
    return buf.transpose(1, 0, 2)


def as_pairs(x):
    """copy an (rows, cols, 2) array into the AoS column-major layout."""
    out = pairs(x.shape[0], x.shape[1])
    out[...] = x
    return out


def fmat(rows, cols):
    return np.zeros((rows, cols), dtype=np.float64, order="F")


def rgemm(A, B, mode, C=None, accumulate=False):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m, k = A.shape
    n = B.shape[1]
    if C is None:
        C = fmat(m, n)
    lib().orc_rgemm(m, n, k, _dp(A), max(m, 1), _dp(B), max(k, 1), _dp(C), max(m, 1), mode, int(accumulate))
    return C


def midrad(inf, sup):
    m = ctypes.c_double()
    r = ctypes.c_double()
    lib().orc_midrad(inf, sup, ctypes.byref(m), ctypes.byref(r))
    return m.value, r.value


def pidblas_mulmm(IA, IB):
    m, k = IA.shape[:2]
    n = IB.shape[1]
    C = pairs(m, n)
    lib().orc_pidblas_mulmm(m, n, k, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), _dp(IB.transpose(1, 0, 2)),
                            _ld_pair(IB), _dp(C.transpose(1, 0, 2)), _ld_pair(C))
    return C


def pidblas_fma_mulmm(IA, IB):
    """pidblas_fma::mulmm: inf-sup interval FMA chains (vcp/pidblas_fma.hpp:45-72)"""
    m, k = IA.shape[:2]
    n = IB.shape[1]
    C = pairs(m, n)
    rc = lib().orc_pidblas_fma_mulmm(m, n, k, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), _dp(IB.transpose(1, 0, 2)),
                                     _ld_pair(IB), _dp(C.transpose(1, 0, 2)), _ld_pair(C))
    assert rc == 0
    return C


def pidblas_mul_im_m(IA, B):
    B = np.asfortranarray(B, dtype=np.float64)
    m, k = IA.shape[:2]
    n = B.shape[1]
    C = pairs(m, n)
    lib().orc_pidblas_mul_im_m(m, n, k, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), _dp(B), max(k, 1),
                               _dp(C.transpose(1, 0, 2)), _ld_pair(C))
    return C


def pidblas_mul_m_im(B, IA):
    B = np.asfortranarray(B, dtype=np.float64)
    m, k = B.shape
    n = IA.shape[1]
    C = pairs(m, n)
    lib().orc_pidblas_mul_m_im(m, n, k, _dp(B), max(m, 1), _dp(IA.transpose(1, 0, 2)), _ld_pair(IA),
                               _dp(C.transpose(1, 0, 2)), _ld_pair(C))
    return C


def pidblas_mulltmm(IA):
    k, n = IA.shape[:2]
    C = pairs(n, n)
    lib().orc_pidblas_mulltmm(k, n, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), _dp(C.transpose(1, 0, 2)),
                              _ld_pair(C))
    return C


def pidblas_vmulmm(A):
    A = np.asfortranarray(A, dtype=np.float64)
    k, n = A.shape
    C = pairs(n, n)
    lib().orc_pidblas_vmulmm(k, n, _dp(A), max(k, 1), _dp(C.transpose(1, 0, 2)), _ld_pair(C))
    return C


def residual(R, IA):
    R = np.asfortranarray(R, dtype=np.float64)
    n = R.shape[0]
    C = pairs(n, n)
    nub = ctypes.c_double()
    lib().orc_residual(n, _dp(R), max(n, 1), _dp(IA.transpose(1, 0, 2)), _ld_pair(IA),
                       _dp(C.transpose(1, 0, 2)), _ld_pair(C), ctypes.byref(nub))
    return C, nub.value


def imul(al, au, bl, bu):
    lo, hi = ctypes.c_double(), ctypes.c_double()
    lib().orc_imul(al, au, bl, bu, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def uhphiphi(phi, w, u):
    """Galerkin matrix of Legendre_Bases_Generator::uhphiphi: phi (m, npts, 2)
    AoS column-major intervals, w and u (npts, 2)"""
    m, npts = phi.shape[:2]
    A = pairs(m, m)
    wv = np.ascontiguousarray(w, dtype=np.float64).reshape(npts, 2)
    uv = np.ascontiguousarray(u, dtype=np.float64).reshape(npts, 2)
    rc = lib().orc_uhphiphi(m, npts, _dp(phi.transpose(1, 0, 2)), _ld_pair(phi), _dp(wv), _dp(uv),
                            _dp(A.transpose(1, 0, 2)), _ld_pair(A))
    if rc:
        raise ValueError("orc_uhphiphi: bad dimensions")
    return A


def lss(IA, Ib, R):
    """imats::linearsolve after its LAPACK step (R ~ inv(mid IA)): returns
    (X, g, ok) with X the enclosure, g >= ||R IA - I||_inf and ok False
    when g >= 1 (the reference's verification failure; X is then zero)."""
    R = np.asfortranarray(R, dtype=np.float64)
    n, nrhs = Ib.shape[:2]
    X = pairs(n, nrhs)
    g = ctypes.c_double()
    rc = lib().orc_lss(n, nrhs, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), _dp(Ib.transpose(1, 0, 2)), _ld_pair(Ib),
                       _dp(R), max(n, 1), _dp(X.transpose(1, 0, 2)), _ld_pair(X), ctypes.byref(g))
    if rc < 0:
        raise ValueError("orc_lss: bad dimensions")
    return X, g.value, rc == 0


def exact_dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rd, ru, rn = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_exact_dot(len(a), _dp(a), 1, _dp(b), 1, ctypes.byref(rd), ctypes.byref(ru), ctypes.byref(rn))
    return rd.value, ru.value, rn.value


def exact_hull(IArow, IBcol, c_lower=0.0, c_upper=0.0):
    """IArow: (k, 2) intervals, IBcol: (k, 2) intervals."""
    a = np.ascontiguousarray(IArow, dtype=np.float64)
    b = np.ascontiguousarray(IBcol, dtype=np.float64)
    okl, oku = ctypes.c_int(), ctypes.c_int()
    hl, hh = ctypes.c_double(), ctypes.c_double()
    lib().orc_exact_hull_entry(a.shape[0], _dp(a), 1, _dp(b), 1, c_lower, c_upper,
                               ctypes.byref(okl), ctypes.byref(oku), ctypes.byref(hl), ctypes.byref(hh))
    return bool(okl.value), bool(oku.value), hl.value, hh.value


def check_containment(IA, IB, C, ii, jj):
    """exact containment of sampled entries (i, j) of IA*IB by C."""
    k = IA.shape[1]
    ii = np.ascontiguousarray(ii, dtype=np.int32)
    jj = np.ascontiguousarray(jj, dtype=np.int32)
    ns = len(ii)
    ok = np.zeros(ns, dtype=np.int32)
    hl = np.zeros(ns)
    hh = np.zeros(ns)
    lib().orc_check_containment(k, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), _dp(IB.transpose(1, 0, 2)),
                                _ld_pair(IB), _dp(C.transpose(1, 0, 2)), _ld_pair(C), ns,
                                ii.ctypes.data_as(_p_int), jj.ctypes.data_as(_p_int), ok.ctypes.data_as(_p_int),
                                _dp(hl), _dp(hh))
    return ok.astype(bool), hl, hh


def exact_dd_dot(arow, bcol, ch=0.0, cl=0.0):
    """arow, bcol: (k, 2) dd values; returns nearest dd of the exact dot
    product and |exact - (ch + cl)| rounded up."""
    a = np.ascontiguousarray(arow, dtype=np.float64)
    b = np.ascontiguousarray(bcol, dtype=np.float64)
    h, l, e = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_exact_dd_dot(a.shape[0], _dp(a), 1, _dp(b), 1, ch, cl, ctypes.byref(h), ctypes.byref(l),
                           ctypes.byref(e))
    return h.value, l.value, e.value


def ozaki_moduli(nmod):
    out = (ctypes.c_int * nmod)()
    cnt = lib().orc_ozaki_moduli(nmod, out)
    return list(out)[:cnt]


def ozaki_beta(k, nmod):
    return lib().orc_ozaki_beta(k, nmod)


OZ_RADIUS_FULL = 1
OZ_RADIUS_FOLDED = 2
OZ_RADIUS_HALF = 3
OZ_RADIUS_CEIL = 4
OZ_RADIUS_ALL = (OZ_RADIUS_FULL, OZ_RADIUS_FOLDED, OZ_RADIUS_HALF, OZ_RADIUS_CEIL)
OZ_FOLDED_MINK = 4096   # the library's untuned default: folded from this k


def ozaki_igemm(IA, IB, nmod, radius=None):
    """radius: one of OZ_RADIUS_ALL; None follows the library's untuned
    default (folded for k >= 4096, else full)"""
    m, k = IA.shape[:2]
    n = IB.shape[1]
    if radius is None:
        radius = OZ_RADIUS_FOLDED if k >= OZ_FOLDED_MINK else OZ_RADIUS_FULL
    C = pairs(m, n)
    rc = lib().orc_ozaki_igemm(m, n, k, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), _dp(IB.transpose(1, 0, 2)),
                               _ld_pair(IB), _dp(C.transpose(1, 0, 2)), _ld_pair(C), nmod, radius)
    if rc != 0:
        raise ValueError("orc_ozaki_igemm failed: %d" % rc)
    return C


def ozaki_scaled(IA, beta, rows=True):
    """scaled integer operand of the Ozaki pipeline: rows of IA (m x k) or
    columns (k x n); returns (int64 matrix, exponents)."""
    r, c = IA.shape[:2]
    out = np.zeros((r, c), dtype=np.int64, order="F")
    ex = np.zeros(r if rows else c, dtype=np.int32)
    f = lib().orc_ozaki_scaled_rows if rows else lib().orc_ozaki_scaled_cols
    f(r, c, _dp(IA.transpose(1, 0, 2)), _ld_pair(IA), beta, out.ctypes.data_as(_p_i64), ex.ctypes.data_as(_p_int))
    return out, ex


def dd_params(k):
    L = lib()
    return L.orc_dd_split_shift(k), L.orc_dd_split_bits(k), L.orc_dd_max_slices(k)


def pddblas_mulmm(A, B):
    m, k = A.shape[:2]
    n = B.shape[1]
    C = pairs(m, n)
    na, nb = ctypes.c_int(), ctypes.c_int()
    lib().orc_pddblas_mulmm(m, n, k, _dp(A.transpose(1, 0, 2)), _ld_pair(A), _dp(B.transpose(1, 0, 2)),
                            _ld_pair(B), _dp(C.transpose(1, 0, 2)), _ld_pair(C), ctypes.byref(na), ctypes.byref(nb))
    return C, na.value, nb.value


def pddblas_mulltmm(A):
    k, n = A.shape[:2]
    C = pairs(n, n)
    lib().orc_pddblas_mulltmm(k, n, _dp(A.transpose(1, 0, 2)), _ld_pair(A), _dp(C.transpose(1, 0, 2)), _ld_pair(C))
    return C


def set_threads(n):
    lib().orc_set_threads(int(n))


# ---- oracle/_ref: the reference's own vblas GEMMs (built by `make ref`) ----
_REF_DIR = os.path.join(_HERE, "_ref")
_ref = None


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in (line + " ")
    except OSError:
        pass
    return False


def ref_lib():
    """The reference's rmatmul/udmatmul compiled from its own sources, or
    None when it has not been built (it needs the reference tree)."""
    global _ref
    if _ref is None:
        name = "libvblas_ref_avx512.so" if _cpu_has_avx512() else "libvblas_ref_avx2.so"
        path = os.path.join(_REF_DIR, name)
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_rmatmul.argtypes = [_c_int, _c_int, _c_int, _p_dbl, _p_dbl, _p_dbl, _c_int]
        L.ref_rmatmul.restype = None
        L.ref_udmatmul.argtypes = [_c_int, _c_int, _c_int, _p_dbl, _p_dbl, _p_dbl, _p_dbl]
        L.ref_udmatmul.restype = _c_int
        L.ref_backend.restype = _c_int
        L.ref_pidblas_mulmm.argtypes = [_c_int, _c_int, _c_int, _p_dbl, _p_dbl, _p_dbl]
        L.ref_pidblas_mulmm.restype = None
        for name in ("ref_pidblas_mul_im_m", "ref_pidblas_mul_m_im"):
            getattr(L, name).argtypes = [_c_int, _c_int, _c_int, _p_dbl, _p_dbl, _p_dbl]
            getattr(L, name).restype = None
        for name in ("ref_pidblas_mulltmm", "ref_pidblas_vmulmm"):
            getattr(L, name).argtypes = [_c_int, _c_int, _p_dbl, _p_dbl]
            getattr(L, name).restype = None
        _ref = L
    return _ref


def ref_rmatmul(A, B, mode):
    """reference rmatmul: C = A*B rounded by mode (A m x k, B k x n)."""
    L = ref_lib()
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m, k = A.shape
    n = B.shape[1]
    C = fmat(m, n)
    L.ref_rmatmul(m, k, n, _dp(A), _dp(B), _dp(C), mode)
    return C


def ref_pidblas_mulmm(IA, IB):
    """pidblas::mulmm with the reference's own vblas rmatmul for its four
    products (oracle/_ref); None when the reference build is absent."""
    L = ref_lib()
    if L is None:
        return None
    m, k = IA.shape[:2]
    n = IB.shape[1]
    A = np.ascontiguousarray(IA.transpose(1, 0, 2))
    B = np.ascontiguousarray(IB.transpose(1, 0, 2))
    C = pairs(m, n)
    Cv = np.zeros((n, m, 2))
    L.ref_pidblas_mulmm(m, n, k, _dp(A), _dp(B), _dp(Cv))
    C[...] = Cv.transpose(1, 0, 2)
    return C


def _ref_pairs_out(rows, cols):
    return np.zeros((cols, rows, 2))


def _ref_pairs_in(X):
    return np.ascontiguousarray(X.transpose(1, 0, 2))


def _ref_pairs_back(Cv):
    C = pairs(Cv.shape[1], Cv.shape[0])
    C[...] = Cv.transpose(1, 0, 2)
    return C


def ref_pidblas_mul_im_m(IA, B):
    """pidblas::mul_im_m (C = IA B, B point) on the reference's rmatmul; None
    when the reference build is absent."""
    L = ref_lib()
    if L is None:
        return None
    m, k = IA.shape[:2]
    n = B.shape[1]
    Bf = np.asfortranarray(B, dtype=np.float64)
    Cv = _ref_pairs_out(m, n)
    L.ref_pidblas_mul_im_m(m, n, k, _dp(_ref_pairs_in(IA)), _dp(Bf), _dp(Cv))
    return _ref_pairs_back(Cv)


def ref_pidblas_mul_m_im(B, IA):
    """pidblas::mul_m_im (C = B IA, B point) on the reference's rmatmul"""
    L = ref_lib()
    if L is None:
        return None
    m, k = B.shape
    n = IA.shape[1]
    Bf = np.asfortranarray(B, dtype=np.float64)
    Cv = _ref_pairs_out(m, n)
    L.ref_pidblas_mul_m_im(m, n, k, _dp(Bf), _dp(_ref_pairs_in(IA)), _dp(Cv))
    return _ref_pairs_back(Cv)


def ref_pidblas_mulltmm(IA):
    """pidblas::mulltmm (IA^T IA) on the reference's rmatmul"""
    L = ref_lib()
    if L is None:
        return None
    k, n = IA.shape[:2]
    Cv = _ref_pairs_out(n, n)
    L.ref_pidblas_mulltmm(k, n, _dp(_ref_pairs_in(IA)), _dp(Cv))
    return _ref_pairs_back(Cv)


def ref_pidblas_vmulmm(A):
    """pidblas::vmulmm (verified A^T A of a point matrix) on the reference's rmatmul"""
    L = ref_lib()
    if L is None:
        return None
    k, n = A.shape
    Af = np.asfortranarray(A, dtype=np.float64)
    Cv = _ref_pairs_out(n, n)
    L.ref_pidblas_vmulmm(k, n, _dp(Af), _dp(Cv))
    return _ref_pairs_back(Cv)


def ref_udmatmul(A, B, CU, CD):
    """reference udmatmul (accumulating RU/RD); returns False if unavailable."""
    L = ref_lib()
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    B = np.asfortranarray(B, dtype=np.float64).copy(order="F")
    m, k = A.shape
    n = B.shape[1]
    return L.ref_udmatmul(m, k, n, _dp(A), _dp(B), _dp(CU), _dp(CD)) == 0


def ozaki_dd_beta(k, nmod):
    L = lib()
    L.orc_ozaki_dd_beta.argtypes = [_c_int, _c_int]
    return L.orc_ozaki_dd_beta(k, nmod)


def dd_default_nmod(k):
    """the dd product's default moduli count: the fewest in 24..32 with beta >= 110"""
    return next(t for t in range(24, 33) if ozaki_dd_beta(k, t) >= 110)


def dd_residual(A, X, B, nmod=None):
    """R = B - A X for dd (m, k, 2), (k, n, 2), (m, n, 2) pair arrays"""
    L = lib()
    L.orc_dd_residual.argtypes = [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl,
                                  _c_int, _c_int]
    L.orc_dd_residual.restype = _c_int
    m, k = A.shape[:2]
    n = X.shape[1]
    if nmod is None:
        nmod = dd_default_nmod(k)
    R = pairs(m, n)
    rc = L.orc_dd_residual(m, n, k, _dp(A.transpose(1, 0, 2)), _ld_pair(A), _dp(X.transpose(1, 0, 2)), _ld_pair(X),
                           _dp(B.transpose(1, 0, 2)), _ld_pair(B), _dp(R.transpose(1, 0, 2)), _ld_pair(R), nmod)
    if rc < 0:
        raise ValueError("orc_dd_residual failed: %d" % rc)
    return R


def ozaki_ddgemm(A, B, nmod):
    """dd C = A*B (or A^T A when B is None) through the exact scaled-integer
    product of the Ozaki-II route."""
    L = lib()
    L.orc_ozaki_ddgemm.argtypes = [_c_int, _c_int, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _p_dbl, _c_int, _c_int]
    L.orc_ozaki_ddgemm.restype = _c_int
    if B is None:
        k, n = A.shape[:2]
        C = pairs(n, n)
        rc = L.orc_ozaki_ddgemm(n, n, k, _dp(A.transpose(1, 0, 2)), _ld_pair(A), None, 0,
                                _dp(C.transpose(1, 0, 2)), _ld_pair(C), nmod)
    else:
        m, k = A.shape[:2]
        n = B.shape[1]
        C = pairs(m, n)
        rc = L.orc_ozaki_ddgemm(m, n, k, _dp(A.transpose(1, 0, 2)), _ld_pair(A), _dp(B.transpose(1, 0, 2)),
                                _ld_pair(B), _dp(C.transpose(1, 0, 2)), _ld_pair(C), nmod)
    if rc != 0:
        raise ValueError("orc_ozaki_ddgemm failed: %d" % rc)
    return C
