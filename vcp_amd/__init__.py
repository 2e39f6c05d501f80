"""vcp_amd -- MI355X-native verified dense matrix product for VCP.

Python host binding of libvcphip (the C ABI in include/vcphip.h).  The
library is the product: every call below goes to hand-written gfx950 HIP
kernels; there is no CPU fallback, and a missing or unloadable library
raises.

Host arrays follow VCP's storage (vcp/mats.hpp:411-417): column-major; an
interval matrix is a float64 array of shape (rows, cols, 2) whose memory is
the AoS column-major sequence {inf, sup} of kv::interval<double>; a dd
matrix is the same with {hi, lo} (kv::dd).  Use ``pairs``/``as_pairs`` to
make one.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "Context", "VcphipError", "pairs", "as_pairs", "FORM_PLAIN", "FORM_OZAKI", "FORM_INFSUP"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvcphip.so")

FORM_PLAIN = 0
FORM_OZAKI = 1
FORM_INFSUP = 2
PARAM_OZ_NMOD = 1
PARAM_OZ_VARIANT = 2
PARAM_STAGE_TIMING = 3
PARAM_DD_NMOD = 4
PARAM_OZ_RADIUS = 5
PARAM_BATCH_TILES = 6
PARAM_PLANE_ROWS = 7
BATCH_TILES_64 = 1
BATCH_TILES_EDGE = 2
BATCH_TILES_32 = 3
BATCH_TILES_NAMES = {1: "64x64", 2: "64x64+32x64 edge", 3: "32x64"}
NOALLOC_SCRATCH_BYTES = 100 << 20
OZ_RADIUS_FULL = 1
OZ_RADIUS_FOLDED = 2
OZ_RADIUS_HALF = 3
OZ_RADIUS_CEIL = 4
OZ_RADIUS_NAMES = {1: "full", 2: "folded", 3: "half", 4: "ceil"}

ERRORS = {
    -1: "dimension", -2: "argument", -3: "device", -4: "oom", -5: "rounding",
    -6: "verification", -7: "unsupported", -8: "range", -9: "io", -10: "singular", -11: "scratch",
}

_c_int = ctypes.c_int
_c_ll = ctypes.c_longlong
_vp = ctypes.c_void_p
_p_int = ctypes.POINTER(ctypes.c_int)
_p_dbl = ctypes.POINTER(ctypes.c_double)

# exported symbols and their argument types (every symbol of include/vcphip.h)
SIGNATURES = {
    "vcphip_ctx_create": [_c_int, ctypes.POINTER(_vp)],
    "vcphip_ctx_destroy": [_vp],
    "vcphip_ctx_set_stream": [_vp, _vp],
    "vcphip_ctx_synchronize": [_vp],
    "vcphip_ctx_set_param": [_vp, _c_int, _c_ll],
    "vcphip_ctx_get_param": [_vp, _c_int, ctypes.POINTER(_c_ll)],
    "vcphip_ctx_stage_times": [_vp, _p_dbl, _c_int, _p_int, ctypes.c_char_p, _c_int],
    "vcphip_strerror": [_c_int],
    "vcphip_version": [],
    "vcphip_ozaki_config": [_c_int, _c_int, _p_int, _p_int, _c_int],
    "vcphip_igemm": [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_dev": [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_ip": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_ip_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_pi": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_pi_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_tn": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_tn_dev": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_vgemm_tn": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_vgemm_tn_dev": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_gemm_rurd": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int],
    "vcphip_gemm_rurd_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int],
    "vcphip_ddgemm": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_ddgemm_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_ddgemm_tn": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_igemm_batched": [_vp, _c_int, _p_int, _p_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int],
    "vcphip_igemm_batched_dev": [_vp, _c_int, _p_int, _p_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int],
    "vcphip_inv_mid": [_vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_dd_solve": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _p_int],
    "vcphip_dd_cholesky": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int],
    "vcphip_dd_eigsym": [_vp, _c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int],
    "vcphip_dd_eigsym_dev": [_vp, _c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int],
    "vcphip_dd_eigsymge": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int],
    "vcphip_dd_eigsymge_dev": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int],
    "vcphip_dd_cholesky_dev": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int],
    "vcphip_dd_solve_dev": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _p_int],
    "vcphip_inv_mid_dev": [_vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_lss": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _p_dbl],
    "vcphip_lss_dev": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _p_dbl],
    "vcphip_lss_batched": [_vp, _c_int, _p_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int, _p_dbl],
    "vcphip_lss_batched_dev": [_vp, _c_int, _p_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int,
                               _p_dbl],
    "vcphip_ddgemm_ta": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_ddgemm_ta_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_dd_residual": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_dd_residual_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_uhphiphi": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _c_int],
    "vcphip_uhphiphi_dev": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _c_int],
    "vcphip_mat_info": [ctypes.c_char_p, _p_int, _p_int, _p_int, _p_int],
    "vcphip_mat_read": [ctypes.c_char_p, _c_int, _c_int, _c_int, _vp, _c_int],
    "vcphip_mat_write": [ctypes.c_char_p, _c_int, _c_int, _c_int, _vp, _c_int, _c_int],
    "vcphip_crc64": [_vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)],
    "vcphip_residual_batched": [_vp, _c_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int, _p_dbl],
    "vcphip_residual_batched_dev": [_vp, _c_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int, _vp],
    "vcphip_tune_igemm_tile": [_vp, _c_int, _c_int, _c_int, _p_int, _p_dbl],
    "vcphip_tune_ddgemm_tile": [_vp, _c_int, _c_int, _c_int, _p_int, _p_dbl],
    "vcphip_tune_igemm_moduli_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, ctypes.c_double, _c_int,
                                     _p_int, _p_int, _p_dbl],
    "vcphip_tune_residual_dev": [_vp, _c_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int, _p_dbl,
                                 _p_dbl],
    "vcphip_tune_igemm_batched_dev": [_vp, _c_int, _p_int, _p_int, _p_int, _vp, _p_int, _vp, _p_int, _vp, _p_int,
                                      _p_dbl, _p_dbl],
    "vcphip_ozaki_plan": [_vp, _c_int, _c_int, _c_int, _p_int, _p_int, _p_int, _p_int],
    "vcphip_igemm_ozaki_noalloc_dev": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int],
    "vcphip_tune_save": [_vp, ctypes.c_char_p],
    "vcphip_tune_load": [_vp, ctypes.c_char_p],
}

_lib = None


class VcphipError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        msg = lib().vcphip_strerror(code).decode() if _lib is not None else str(code)
        super().__init__("%s failed: %s (%d)" % (where, msg, code))


def lib():
    """Load libvcphip.so (built by vcp_amd.build); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libvcphip.so not built (run `python -m vcp_amd.build`): %s" % LIB_PATH)
        # PyTorch-ROCm ships its own HIP runtime with the same soname
        # (libamdhip64.so.7) but loads it by another name; loading torch first
        # makes this library bind to that runtime, so one process never holds
        # two HIP runtimes (device pointers from torch are then native here).
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_char_p if name == "vcphip_strerror" else ctypes.c_int
        _lib = L
    return _lib


def _check(rc, where):
    if rc != 0:
        raise VcphipError(rc, where)


def pairs(rows, cols):
    """zeroed (rows, cols, 2) float64 array in AoS column-major layout."""
    return np.zeros((cols, rows, 2), dtype=np.float64).transpose(1, 0, 2)


def as_pairs(x):
    out = pairs(x.shape[0], x.shape[1])
    out[...] = x
    return out


def _pair_ptr(a):
    """pointer and leading dimension of an AoS column-major pair array"""
    assert a.dtype == np.float64 and a.ndim == 3 and a.shape[2] == 2
    if a.size == 0:
        return a.ctypes.data_as(_vp), max(a.shape[0], 1)
    assert a.strides[2] == 8 and a.strides[0] == 16, "not an AoS column-major pair array"
    ld = a.strides[1] // 16 if a.shape[1] > 1 else max(a.shape[0], 1)
    return a.ctypes.data_as(_vp), max(ld, 1)


def _mat_ptr(a):
    assert a.dtype == np.float64 and a.ndim == 2 and (a.flags.f_contiguous or a.size == 0)
    return a.ctypes.data_as(_vp), max(a.shape[0], 1)


# ---- matrix files (vcp/vcp_portable_fio.hpp, vcp/vcp_fio.hpp) --------------
MAT_INT32, MAT_DOUBLE, MAT_INTERVAL, MAT_DD, MAT_INTERVAL_DD = 1, 2, 3, 4, 5
_MAT_WORDS = {MAT_INT32: 1, MAT_DOUBLE: 1, MAT_INTERVAL: 2, MAT_DD: 2, MAT_INTERVAL_DD: 4}


def mat_info(path):
    """(type, rows, cols, portable) of a .vcpmat or .matrix_* file"""
    t, r, c, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().vcphip_mat_info(os.fsencode(path), ctypes.byref(t), ctypes.byref(r), ctypes.byref(c), ctypes.byref(p)),
           "vcphip_mat_info")
    return t.value, r.value, c.value, bool(p.value)


def _mat_array(mtype, rows, cols):
    if mtype == MAT_INT32:
        return np.zeros((rows, cols), dtype=np.int32, order="F")
    w = _MAT_WORDS[mtype]
    if w == 1:
        return np.zeros((rows, cols), dtype=np.float64, order="F")
    return np.zeros((cols, rows, w), dtype=np.float64).transpose(1, 0, 2)


def mat_read(path):
    """(array, type): int32 / float64 (rows, cols) Fortran-order arrays, or
    AoS column-major (rows, cols, 2) pairs (interval, dd) and (rows, cols, 4)
    for interval<dd>"""
    mtype, rows, cols, _ = mat_info(path)
    X = _mat_array(mtype, rows, cols)
    _check(lib().vcphip_mat_read(os.fsencode(path), mtype, rows, cols, X.ctypes.data_as(_vp), max(rows, 1)),
           "vcphip_mat_read")
    return X, mtype


def mat_write(path, X, mtype, portable=True):
    """write X (shaped as mat_read returns it) as a .vcpmat (portable) or
    native .matrix_* file"""
    rows, cols = X.shape[:2]
    Y = _mat_array(mtype, rows, cols)
    Y[...] = X
    _check(lib().vcphip_mat_write(os.fsencode(path), mtype, rows, cols, Y.ctypes.data_as(_vp), max(rows, 1),
                                  int(bool(portable))), "vcphip_mat_write")


def ozaki_config(k, nmod=0):
    beta = ctypes.c_int()
    mods = (ctypes.c_int * 64)()
    n = lib().vcphip_ozaki_config(k, nmod, ctypes.byref(beta), mods, 64)
    if n < 0:
        raise VcphipError(n, "vcphip_ozaki_config")
    return beta.value, list(mods)[:n]


class Context:
    """A libvcphip context on one device (one stream, cached workspace)."""

    def __init__(self, device=0):
        self._h = _vp()
        _check(lib().vcphip_ctx_create(device, ctypes.byref(self._h)), "vcphip_ctx_create")

    def close(self):
        if self._h:
            lib().vcphip_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_param(self, key, value):
        _check(lib().vcphip_ctx_set_param(self._h, key, int(value)), "vcphip_ctx_set_param")

    def get_param(self, key):
        v = _c_ll()
        _check(lib().vcphip_ctx_get_param(self._h, key, ctypes.byref(v)), "vcphip_ctx_get_param")
        return v.value

    def set_stream(self, stream_ptr):
        _check(lib().vcphip_ctx_set_stream(self._h, _vp(stream_ptr) if stream_ptr else None), "vcphip_ctx_set_stream")

    def synchronize(self):
        _check(lib().vcphip_ctx_synchronize(self._h), "vcphip_ctx_synchronize")

    def stage_times(self):
        ms = (ctypes.c_double * 32)()
        cnt = ctypes.c_int()
        names = ctypes.create_string_buffer(512)
        _check(lib().vcphip_ctx_stage_times(self._h, ms, 32, ctypes.byref(cnt), names, 512), "stage_times")
        nm = names.value.decode().split(";") if cnt.value else []
        return dict(zip(nm, list(ms)[:cnt.value]))

    # ---- host-pointer products --------------------------------------------
    def igemm(self, IA, IB, form=FORM_PLAIN):
        """C = IA * IB (pidblas::mulmm)."""
        m, k = IA.shape[:2]
        k2, n = IB.shape[:2]
        if k != k2:
            raise ValueError("dimension mismatch")
        C = pairs(m, n)
        pa, lda = _pair_ptr(IA)
        pb, ldb = _pair_ptr(IB)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_igemm(self._h, form, m, n, k, pa, lda, pb, ldb, pc, ldc), "vcphip_igemm")
        return C

    def igemm_ip(self, IA, B):
        """C = IA * B, B point (pidblas::mul_im_m)."""
        m, k = IA.shape[:2]
        n = B.shape[1]
        C = pairs(m, n)
        pa, lda = _pair_ptr(IA)
        pb, ldb = _mat_ptr(B)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_igemm_ip(self._h, m, n, k, pa, lda, pb, ldb, pc, ldc), "vcphip_igemm_ip")
        return C

    def igemm_pi(self, B, IA):
        """C = B * IA, B point (pidblas::mul_m_im)."""
        m, k = B.shape
        n = IA.shape[1]
        C = pairs(m, n)
        pb, ldb = _mat_ptr(B)
        pa, lda = _pair_ptr(IA)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_igemm_pi(self._h, m, n, k, pb, ldb, pa, lda, pc, ldc), "vcphip_igemm_pi")
        return C

    def igemm_tn(self, IA):
        """C = IA^T IA (pidblas::mulltmm)."""
        k, n = IA.shape[:2]
        C = pairs(n, n)
        pa, lda = _pair_ptr(IA)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_igemm_tn(self._h, k, n, pa, lda, pc, ldc), "vcphip_igemm_tn")
        return C

    def vgemm_tn(self, A):
        """verified A^T A of a point matrix (pidblas::vmulmm)."""
        k, n = A.shape
        C = pairs(n, n)
        pa, lda = _mat_ptr(A)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_vgemm_tn(self._h, k, n, pa, lda, pc, ldc), "vcphip_vgemm_tn")
        return C

    def gemm_rurd(self, A, B, Cup=None, Cdn=None, accumulate=False):
        """(RU(A*B), RD(A*B)) as k-ascending FMA chains (rmatmul +1/-1;
        udmatmul when accumulate)."""
        m, k = A.shape
        n = B.shape[1]
        if Cup is None:
            Cup = np.zeros((m, n), order="F")
        if Cdn is None:
            Cdn = np.zeros((m, n), order="F")
        pa, lda = _mat_ptr(A)
        pb, ldb = _mat_ptr(B)
        pu, ldc = _mat_ptr(Cup)
        pd, _ = _mat_ptr(Cdn)
        _check(lib().vcphip_gemm_rurd(self._h, m, n, k, pa, lda, pb, ldb, pu, pd, ldc, int(accumulate)),
               "vcphip_gemm_rurd")
        return Cup, Cdn

    def ddgemm(self, A, B):
        """C = A * B for double-double matrices (pddblas::mulmm)."""
        m, k = A.shape[:2]
        n = B.shape[1]
        C = pairs(m, n)
        pa, lda = _pair_ptr(A)
        pb, ldb = _pair_ptr(B)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_ddgemm(self._h, m, n, k, pa, lda, pb, ldb, pc, ldc), "vcphip_ddgemm")
        return C

    def uhphiphi(self, phi, w, u):
        """Galerkin matrix A(i, j) = sum_k ((w_k u_k) phi(i, k)) phi(j, k) in
        interval arithmetic (Legendre_Bases_Generator::uhphiphi, vcp/ldbase.hpp:
        955-1013): phi (m, npts, 2), w and u (npts, 2) intervals."""
        m, npts = phi.shape[:2]
        A = pairs(m, m)
        pp, ldp = _pair_ptr(phi)
        wv = np.ascontiguousarray(w, dtype=np.float64).reshape(npts, 2)
        uv = np.ascontiguousarray(u, dtype=np.float64).reshape(npts, 2)
        pa, lda = _pair_ptr(A)
        _check(lib().vcphip_uhphiphi(self._h, m, npts, pp, ldp, wv.ctypes.data_as(_vp), uv.ctypes.data_as(_vp), pa,
                                     lda), "vcphip_uhphiphi")
        return A

    def ddgemm_ta(self, A, B):
        """C = A^T B for double-double matrices (A k x m, B k x n)."""
        k, m = A.shape[:2]
        n = B.shape[1]
        C = pairs(m, n)
        pa, lda = _pair_ptr(A)
        pb, ldb = _pair_ptr(B)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_ddgemm_ta(self._h, m, n, k, pa, lda, pb, ldb, pc, ldc), "vcphip_ddgemm_ta")
        return C

    def dd_residual(self, A, X, B):
        """R = B - A X for dd matrices (pddblas::linearsolve's residual)."""
        m, k = A.shape[:2]
        n = X.shape[1]
        R = pairs(m, n)
        pa, lda = _pair_ptr(A)
        px, ldx = _pair_ptr(X)
        pb, ldb = _pair_ptr(B)
        pr, ldr = _pair_ptr(R)
        _check(lib().vcphip_dd_residual(self._h, m, n, k, pa, lda, px, ldx, pb, ldb, pr, ldr), "vcphip_dd_residual")
        return R

    def dd_linearsolve(self, A, b, itmax=10):
        """x = A \\ b for dd matrices by pddblas::linearsolve's refinement
        (vcp/pddblas.hpp:412-474), every step on the device (vcphip_dd_solve):
        LU of the high parts (rocSOLVER getrf), x += the LU correction in dd
        (getrs), stop when the correction is below 0.25e-32 |x| or
        stagnates, else r = b - A x by the dd Ozaki product.  Returns
        (x, iterations); raises VcphipError (VCPHIP_ERR_SINGULAR) on a zero
        pivot."""
        n = A.shape[0]
        if A.shape[1] != n or b.shape[0] != n:
            raise ValueError("dimension mismatch")
        nrhs = b.shape[1]
        x = pairs(n, nrhs)
        if n == 0 or nrhs == 0:
            return x, 0
        pa, lda = _pair_ptr(A)
        pb, ldb = _pair_ptr(b)
        px, ldx = _pair_ptr(x)
        it = ctypes.c_int()
        _check(lib().vcphip_dd_solve(self._h, n, nrhs, pa, lda, pb, ldb, px, ldx, itmax, ctypes.byref(it)),
               "vcphip_dd_solve")
        return x, it.value

    def dd_solve_dev(self, n, nrhs, dA, lda, dB, ldb, dX, ldx, itmax=10):
        """device-pointer form of dd_linearsolve(); returns the iterations"""
        it = ctypes.c_int()
        _check(lib().vcphip_dd_solve_dev(self._h, n, nrhs, _vp(dA), lda, _vp(dB), ldb, _vp(dX), ldx, itmax,
                                         ctypes.byref(it)), "vcphip_dd_solve_dev")
        return it.value

    def dd_inv(self, A):
        """inv(A) for a dd matrix: A X = I by dd_linearsolve (pddblas::inv,
        vcp/pddblas.hpp:477-486)."""
        n = A.shape[0]
        eye = pairs(n, n)
        eye[..., 0] = np.eye(n)
        return self.dd_linearsolve(A, eye)[0]

    def dd_cholesky(self, A, itep=2):
        """upper U with U^T U = A to dd accuracy (pddblas::Cholesky,
        vcp/pddblas.hpp:488-563), every step on the device
        (vcphip_dd_cholesky).  Raises VcphipError: VCPHIP_ERR_ARGUMENT when
        A is not symmetric, VCPHIP_ERR_SINGULAR when it is not positive
        definite."""
        n = A.shape[0]
        if A.shape[1] != n:
            raise ValueError("Cholesky: matrix must be square")
        U = pairs(n, n)
        if n == 0:
            return U
        pa, lda = _pair_ptr(A)
        pu, ldu = _pair_ptr(U)
        _check(lib().vcphip_dd_cholesky(self._h, n, pa, lda, pu, ldu, itep), "vcphip_dd_cholesky")
        return U

    def dd_cholesky_dev(self, n, dA, lda, dU, ldu, itep=2):
        """device-pointer form of dd_cholesky()"""
        _check(lib().vcphip_dd_cholesky_dev(self._h, n, _vp(dA), lda, _vp(dU), ldu, itep), "vcphip_dd_cholesky_dev")

    def dd_eigsym(self, A, itep=4):
        """(eigenvalues (n, 2), eigenvectors (n, n, 2)) of a symmetric dd
        matrix (pddblas::eigsym, Ogita-Aishima refinement, vcp/pddblas.hpp:
        219-314, 566-606), every step on the device (vcphip_dd_eigsym)."""
        return self._dd_eig(A, None, itep)

    def dd_eigsymge(self, A, B, itep=1):
        """(eigenvalues, eigenvectors) of A x = lambda B x for symmetric dd A,
        B, B positive definite (pddblas::eigsymge, vcp/pddblas.hpp:611-659),
        every step on the device (vcphip_dd_eigsymge)."""
        return self._dd_eig(A, B, itep)

    def _dd_eig(self, A, B, itep):
        n = A.shape[0]
        if A.shape[1] != n or (B is not None and B.shape[:2] != (n, n)):
            raise ValueError("eigsym: dimension mismatch")
        lam, X = pairs(n, 1), pairs(n, n)
        if n == 0:
            return lam[:, 0, :], X
        pa, lda = _pair_ptr(A)
        px, ldx = _pair_ptr(X)
        pl = lam.ctypes.data_as(_vp)
        if B is None:
            _check(lib().vcphip_dd_eigsym(self._h, n, pa, lda, pl, px, ldx, itep), "vcphip_dd_eigsym")
        else:
            pb, ldb = _pair_ptr(B)
            _check(lib().vcphip_dd_eigsymge(self._h, n, pa, lda, pb, ldb, pl, px, ldx, itep), "vcphip_dd_eigsymge")
        return np.ascontiguousarray(lam[:, 0, :]), X

    def ddgemm_tn(self, A):
        """C = A^T A for a double-double matrix (pddblas::mulltmm)."""
        k, n = A.shape[:2]
        C = pairs(n, n)
        pa, lda = _pair_ptr(A)
        pc, ldc = _pair_ptr(C)
        _check(lib().vcphip_ddgemm_tn(self._h, k, n, pa, lda, pc, ldc), "vcphip_ddgemm_tn")
        return C

    def residual_batched(self, Rs, IAs):
        """[(R_i IA_i - I, ||.||_inf upper bound)] for a list of problems."""
        cnt = len(Rs)
        ns = (ctypes.c_int * cnt)(*[R.shape[0] for R in Rs])
        Cs = [pairs(R.shape[0], R.shape[0]) for R in Rs]
        rp = (_vp * cnt)(*[_mat_ptr(R)[0] for R in Rs])
        ap = (_vp * cnt)(*[_pair_ptr(A)[0] for A in IAs])
        cp = (_vp * cnt)(*[_pair_ptr(C)[0] for C in Cs])
        ldr = (ctypes.c_int * cnt)(*[_mat_ptr(R)[1] for R in Rs])
        lda = (ctypes.c_int * cnt)(*[_pair_ptr(A)[1] for A in IAs])
        ldc = (ctypes.c_int * cnt)(*[_pair_ptr(C)[1] for C in Cs])
        nub = (ctypes.c_double * cnt)()
        _check(lib().vcphip_residual_batched(self._h, cnt, ns, rp, ldr, ap, lda, cp, ldc, nub),
               "vcphip_residual_batched")
        return Cs, list(nub)

    def igemm_batched(self, IAs, IBs):
        """[IA_p * IB_p] for interval matrices of uneven sizes in one
        persistent launch (the plain form of pidblas::mulmm per problem,
        bit-identical to igemm(..., FORM_PLAIN))."""
        cnt = len(IAs)
        if len(IBs) != cnt:
            raise ValueError("igemm_batched: list lengths differ")
        for A, B in zip(IAs, IBs):
            if A.shape[1] != B.shape[0]:
                raise ValueError("igemm_batched: dimension mismatch")
        Cs = [pairs(A.shape[0], B.shape[1]) for A, B in zip(IAs, IBs)]
        if cnt == 0:
            return Cs
        ia = lambda xs: (ctypes.c_int * cnt)(*xs)
        vp = lambda xs: (_vp * cnt)(*xs)
        _check(lib().vcphip_igemm_batched(self._h, cnt, ia([A.shape[0] for A in IAs]), ia([B.shape[1] for B in IBs]),
                                          ia([A.shape[1] for A in IAs]),
                                          vp([_pair_ptr(A)[0] for A in IAs]), ia([_pair_ptr(A)[1] for A in IAs]),
                                          vp([_pair_ptr(B)[0] for B in IBs]), ia([_pair_ptr(B)[1] for B in IBs]),
                                          vp([_pair_ptr(C)[0] for C in Cs]), ia([_pair_ptr(C)[1] for C in Cs])),
               "vcphip_igemm_batched")
        return Cs

    def igemm_batched_dev(self, m, n, k, dA, lda, dB, ldb, dC, ldc):
        """device-pointer form: lists of sizes, pointers and leading
        dimensions; asynchronous on the context stream"""
        cnt = len(m)
        ia = lambda xs: (ctypes.c_int * cnt)(*xs)
        vp = lambda xs: (_vp * cnt)(*[_vp(x) for x in xs])
        _check(lib().vcphip_igemm_batched_dev(self._h, cnt, ia(m), ia(n), ia(k), vp(dA), ia(lda), vp(dB), ia(ldb),
                                              vp(dC), ia(ldc)), "vcphip_igemm_batched_dev")

    def tune_igemm_batched_dev(self, m, n, k, dA, lda, dB, ldb, dC, ldc):
        """tile shape per size class of the batched interval products of this
        set (arguments as igemm_batched_dev); returns (ms per launch of the
        set, padded fraction of the tiles' work)"""
        cnt = len(m)
        ia = lambda xs: (ctypes.c_int * cnt)(*xs)
        vp = lambda xs: (_vp * cnt)(*[_vp(x) for x in xs])
        ms, pad = ctypes.c_double(), ctypes.c_double()
        _check(lib().vcphip_tune_igemm_batched_dev(self._h, cnt, ia(m), ia(n), ia(k), vp(dA), ia(lda), vp(dB),
                                                   ia(ldb), vp(dC), ia(ldc), ctypes.byref(ms), ctypes.byref(pad)),
               "vcphip_tune_igemm_batched_dev")
        return ms.value, pad.value

    def inv_mid(self, IA):
        """R = inv(mid(IA)) formed on the device (imid + pdblas::inv,
        vcp/imats.hpp:298-301, vcp/pdblas.hpp:180-204): kv's midpoints, then
        rocSOLVER dgetrf and dgetrs against the identity.  Returns R (n, n), column-major; raises
        VcphipError (VCPHIP_ERR_SINGULAR) on an exactly zero pivot."""
        n = IA.shape[0]
        if IA.shape[1] != n:
            raise ValueError("inv_mid: matrix must be square")
        R = np.zeros((n, n), dtype=np.float64, order="F")
        if n == 0:
            return R
        pa, lda = _pair_ptr(IA)
        pr, ldr = _mat_ptr(R)
        _check(lib().vcphip_inv_mid(self._h, n, pa, lda, pr, ldr), "vcphip_inv_mid")
        return R

    def inv_mid_dev(self, n, dA, lda, dR, ldr):
        """device-pointer form of inv_mid()"""
        _check(lib().vcphip_inv_mid_dev(self._h, n, _vp(dA), lda, _vp(dR), ldr), "vcphip_inv_mid_dev")

    def lss(self, IA, Ib, R=None):
        """Verified solution of A x = b (imats::linearsolve, vcp/imats.hpp:
        292-362): returns (X, g), X (n, nrhs, 2) containing A^-1 b for every
        A in IA and b in Ib, g >= ||R IA - I||_inf.  R = None has the library
        form R = inv(mid(IA)) on the device (as inv_mid), the reference's
        step at :298-301; else R is the caller's approximate inverse.
        Raises VcphipError (VCPHIP_ERR_VERIFICATION) when g >= 1, as the
        reference throws vcp::verification_error, and VCPHIP_ERR_SINGULAR
        when the device R meets a zero pivot (its lapack_error)."""
        n = IA.shape[0]
        if IA.shape[1] != n or Ib.shape[0] != n:
            raise ValueError("dimension mismatch")
        nrhs = Ib.shape[1]
        X = pairs(n, nrhs)
        pa, lda = _pair_ptr(IA)
        pb, ldb = _pair_ptr(Ib)
        if R is None:
            pr, ldr = None, n
        else:
            R = np.asfortranarray(R, dtype=np.float64)
            pr, ldr = _mat_ptr(R)
        px, ldx = _pair_ptr(X)
        g = ctypes.c_double()
        _check(lib().vcphip_lss(self._h, n, nrhs, pa, lda, pb, ldb, pr, ldr, px, ldx, ctypes.byref(g)), "vcphip_lss")
        return X, g.value

    def lss_batched(self, IAs, Ibs, Rs=None):
        """vcphip_lss for many systems of uneven size in one call: returns
        ([X_p], [g_p]); raises VcphipError (VCPHIP_ERR_VERIFICATION) when
        any g_p >= 1.  Rs = None (or a None entry) has the library form
        those R_p = inv(mid(IA_p)) on the device."""
        cnt = len(IAs)
        if Rs is None:
            Rs = [None] * cnt
        Rs = [None if R is None else np.asfortranarray(R, dtype=np.float64) for R in Rs]
        Xs = [pairs(b.shape[0], b.shape[1]) for b in Ibs]
        ia = lambda xs: (ctypes.c_int * cnt)(*xs)
        vp = lambda xs: (_vp * cnt)(*xs)
        g = (ctypes.c_double * cnt)()
        rp = [(None, A.shape[0]) if R is None else _mat_ptr(R) for R, A in zip(Rs, IAs)]
        _check(lib().vcphip_lss_batched(self._h, cnt, ia([A.shape[0] for A in IAs]), ia([b.shape[1] for b in Ibs]),
                                        vp([_pair_ptr(A)[0] for A in IAs]), ia([_pair_ptr(A)[1] for A in IAs]),
                                        vp([_pair_ptr(b)[0] for b in Ibs]), ia([_pair_ptr(b)[1] for b in Ibs]),
                                        vp([r[0] for r in rp]), ia([r[1] for r in rp]),
                                        vp([_pair_ptr(X)[0] for X in Xs]), ia([_pair_ptr(X)[1] for X in Xs]), g),
               "vcphip_lss_batched")
        return Xs, list(g)

    def lss_batched_dev(self, n, nrhs, dA, lda, dB, ldb, dR, ldr, dX, ldx):
        """device-pointer form: lists of sizes, pointers and leading
        dimensions; returns [g_p] (raises when any g_p >= 1).  dR (with
        ldr) may be None, or hold None entries: those R_p are formed on the
        device."""
        cnt = len(n)
        ia = lambda xs: (ctypes.c_int * cnt)(*xs)
        vp = lambda xs: (_vp * cnt)(*[_vp(x) for x in xs])
        g = (ctypes.c_double * cnt)()
        _check(lib().vcphip_lss_batched_dev(self._h, cnt, ia(n), ia(nrhs), vp(dA), ia(lda), vp(dB), ia(ldb),
                                            None if dR is None else vp(dR), None if ldr is None else ia(ldr), vp(dX),
                                            ia(ldx), g), "vcphip_lss_batched_dev")
        return list(g)

    # ---- device-pointer products (torch tensors or raw pointers) -----------
    def lss_dev(self, n, nrhs, dA, lda, dB, ldb, dR, ldr, dX, ldx):
        """device-pointer form of lss(); returns g (raises when g >= 1)"""
        g = ctypes.c_double()
        _check(lib().vcphip_lss_dev(self._h, n, nrhs, _vp(dA), lda, _vp(dB), ldb, _vp(dR), ldr, _vp(dX), ldx,
                                    ctypes.byref(g)), "vcphip_lss_dev")
        return g.value

    def igemm_dev(self, form, m, n, k, dA, lda, dB, ldb, dC, ldc):
        _check(lib().vcphip_igemm_dev(self._h, form, m, n, k, _vp(dA), lda, _vp(dB), ldb, _vp(dC), ldc),
               "vcphip_igemm_dev")

    def igemm_pi_dev(self, m, n, k, dB, ldb, dA, lda, dC, ldc):
        _check(lib().vcphip_igemm_pi_dev(self._h, m, n, k, _vp(dB), ldb, _vp(dA), lda, _vp(dC), ldc),
               "vcphip_igemm_pi_dev")

    def gemm_rurd_dev(self, m, n, k, dA, lda, dB, ldb, dU, dD, ldc, accumulate=False):
        _check(lib().vcphip_gemm_rurd_dev(self._h, m, n, k, _vp(dA), lda, _vp(dB), ldb, _vp(dU), _vp(dD), ldc,
                                          int(accumulate)), "vcphip_gemm_rurd_dev")

    # ---- autotuner ----
    def tune_igemm_tile(self, m, n, k):
        """fastest int8 GEMM tile variant of the interval Ozaki form at (m, n, k): (variant, ms)"""
        v, ms = ctypes.c_int(), ctypes.c_double()
        _check(lib().vcphip_tune_igemm_tile(self._h, m, n, k, ctypes.byref(v), ctypes.byref(ms)),
               "vcphip_tune_igemm_tile")
        return v.value, ms.value

    def tune_ddgemm_tile(self, m, n, k):
        """fastest int8 GEMM tile variant of the dd form at (m, n, k): (variant, ms)"""
        v, ms = ctypes.c_int(), ctypes.c_double()
        _check(lib().vcphip_tune_ddgemm_tile(self._h, m, n, k, ctypes.byref(v), ctypes.byref(ms)),
               "vcphip_tune_ddgemm_tile")
        return v.value, ms.value

    def tune_igemm_moduli_dev(self, m, n, k, dA, lda, dB, ldb, max_ratio=2.0, sample_rows=0):
        """cheapest (moduli, radius scheme) keeping sampled radii within max_ratio
        of the plain form: (nmod, radius, ratio), radius one of OZ_RADIUS_*"""
        nm, rad, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().vcphip_tune_igemm_moduli_dev(self._h, m, n, k, _vp(dA), lda, _vp(dB), ldb, float(max_ratio),
                                                  int(sample_rows), ctypes.byref(nm), ctypes.byref(rad),
                                                  ctypes.byref(r)),
               "vcphip_tune_igemm_moduli_dev")
        return nm.value, rad.value, r.value

    def igemm_ozaki_noalloc_dev(self, m, n, k, dA, lda, dB, ldb, dC, ldc):
        """the Ozaki-form product of igemm_dev with its workspace in the
        library's fixed device scratch (no allocation; m, n, k <= 1024 fit)"""
        _check(lib().vcphip_igemm_ozaki_noalloc_dev(self._h, m, n, k, _vp(dA), lda, _vp(dB), ldb, _vp(dC), ldc),
               "vcphip_igemm_ozaki_noalloc_dev")

    def ozaki_plan(self, m, n, k):
        """moduli, radius scheme, tile variant and int8 plane products an
        interval Ozaki product of this shape uses in this context"""
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().vcphip_ozaki_plan(self._h, m, n, k, *[ctypes.byref(x) for x in v]), "vcphip_ozaki_plan")
        return {"moduli": v[0].value, "radius": v[1].value, "variant": v[2].value, "products": v[3].value}

    def tune_save(self, path):
        _check(lib().vcphip_tune_save(self._h, os.fsencode(path)), "vcphip_tune_save")

    def tune_load(self, path):
        _check(lib().vcphip_tune_load(self._h, os.fsencode(path)), "vcphip_tune_load")

    def ddgemm_dev(self, m, n, k, dA, lda, dB, ldb, dC, ldc):
        _check(lib().vcphip_ddgemm_dev(self._h, m, n, k, _vp(dA), lda, _vp(dB), ldb, _vp(dC), ldc),
               "vcphip_ddgemm_dev")


def tune_residual_dev(ctx, ns, R_ptrs, ldr, A_ptrs, lda, C_ptrs, ldc, normub_ptr):
    """autotune the batched residual for this problem set: the tile shape per
    size class, then the persistent grid; returns (workgroups per CU, ms of
    the set's tile launches, padded fraction of the tiles' work)"""
    cnt = len(ns)
    arr = lambda T, xs: (T * cnt)(*xs)  # noqa: E731
    per, ms, pad = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _check(lib().vcphip_tune_residual_dev(ctx.handle, cnt, arr(ctypes.c_int, ns), arr(_vp, R_ptrs),
                                          arr(ctypes.c_int, ldr), arr(_vp, A_ptrs), arr(ctypes.c_int, lda),
                                          arr(_vp, C_ptrs), arr(ctypes.c_int, ldc), _vp(normub_ptr),
                                          ctypes.byref(per), ctypes.byref(ms), ctypes.byref(pad)),
           "vcphip_tune_residual_dev")
    return per.value, ms.value, pad.value


def residual_batched_dev(ctx, ns, R_ptrs, ldr, A_ptrs, lda, C_ptrs, ldc, normub_ptr):
    """device-pointer batched residual (pointers as ints), on ctx's stream"""
    cnt = len(ns)
    arr = lambda T, xs: (T * cnt)(*xs)  # noqa: E731
    _check(lib().vcphip_residual_batched_dev(ctx.handle, cnt, arr(ctypes.c_int, ns), arr(_vp, R_ptrs),
                                             arr(ctypes.c_int, ldr), arr(_vp, A_ptrs), arr(ctypes.c_int, lda),
                                             arr(_vp, C_ptrs), arr(ctypes.c_int, ldc), _vp(normub_ptr)),
           "vcphip_residual_batched_dev")
