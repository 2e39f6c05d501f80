/*
 * oracle.c -- CPU restatement of VCP's verified dense matrix product.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the product path never links it.
 *
 * Build: oracle/Makefile (gcc -O2 -mfma -frounding-math -ffp-contract=off
 * -fopenmp).  Directed rounding uses the x86 MXCSR through fesetround(),
 * exactly as kv::hwround does for the reference (vcp/pidblas.hpp:116-134);
 * every rounding-sensitive loop sets the mode at the top of its (OpenMP)
 * region and reads its operands from memory after that call.
 *
 * Exact values come from a fixed-point superaccumulator over the whole
 * double exponent range: every product of two doubles is an exact 106-bit
 * integer times a power of two and is added without rounding.
 */
#include "oracle.h"

#include <fenv.h>
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#pragma STDC FENV_ACCESS ON

typedef unsigned __int128 u128;
typedef __int128 i128;

static int g_threads = 0;
void orc_set_threads(int nthreads) { g_threads = nthreads > 0 ? nthreads : 0; }
int orc_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

static int fe_of(int mode) { return mode > 0 ? FE_UPWARD : (mode < 0 ? FE_DOWNWARD : FE_TONEAREST); }

static double* dalloc(size_t n) {
    double* p = (double*)malloc((n ? n : 1) * sizeof(double));
    return p;
}

/* ------------------------------------------------------------------------ */
/* Directed-rounding point GEMM: the rmatmul contract                       */
/* vblas/README.md:167-170: each C element is the k-ascending sequential    */
/* chain of rounded FMAs; vblas/rmatmul_nosimd.hpp:100-108 (init from 0).   */
/* ------------------------------------------------------------------------ */
#define RG_MB 256
#define RG_NB 4
void orc_rgemm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
               double* C, int ldc, int mode, int accumulate) {
    if (m <= 0 || n <= 0) return;
    const int rm = fe_of(mode);
    const int nt = orc_get_threads();
    const int nblk = (n + RG_NB - 1) / RG_NB;
    const int mblk = (m + RG_MB - 1) / RG_MB;
#pragma omp parallel num_threads(nt)
    {
        const int old = fegetround();
        fesetround(rm);
        double acc[RG_NB][RG_MB];
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int jb = 0; jb < nblk; ++jb) {
            for (int ib = 0; ib < mblk; ++ib) {
                const int j0 = jb * RG_NB, i0 = ib * RG_MB;
                const int jn = (n - j0) < RG_NB ? (n - j0) : RG_NB;
                const int in = (m - i0) < RG_MB ? (m - i0) : RG_MB;
                for (int jj = 0; jj < jn; ++jj)
                    for (int i = 0; i < in; ++i)
                        acc[jj][i] = accumulate ? C[(size_t)(i0 + i) + (size_t)ldc * (j0 + jj)] : 0.0;
                for (int kk = 0; kk < k; ++kk) {
                    const double* a = A + i0 + (size_t)lda * kk;
                    for (int jj = 0; jj < jn; ++jj) {
                        const double b = B[kk + (size_t)ldb * (j0 + jj)];
                        double* c = acc[jj];
                        for (int i = 0; i < in; ++i) c[i] = fma(a[i], b, c[i]);
                    }
                }
                for (int jj = 0; jj < jn; ++jj)
                    for (int i = 0; i < in; ++i)
                        C[(size_t)(i0 + i) + (size_t)ldc * (j0 + jj)] = acc[jj][i];
            }
        }
        fesetround(old);
    }
}

static void transpose(int r, int c, const double* A, int lda, double* At /* c x r */) {
    for (int j = 0; j < c; ++j)
        for (int i = 0; i < r; ++i) At[j + (size_t)c * i] = A[i + (size_t)lda * j];
}

/* pdblas::mulltmm (vcp/pdblas.hpp:119-157): dsyrk 'U','T' then mirror.  The
 * chain for (i,j) and (j,i) is the same sequence of products, so a full GEMM
 * of A^T by A is already symmetric. */
void orc_rsyrk(int k, int n, const double* A, int lda, double* C, int ldc, int mode) {
    double* At = dalloc((size_t)n * k);
    transpose(k, n, A, lda, At);
    orc_rgemm(n, n, k, At, n, A, lda, C, ldc, mode, 0);
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) C[i + (size_t)ldc * j] = C[j + (size_t)ldc * i];
    free(At);
}

/* ------------------------------------------------------------------------ */
/* midrad                                                                   */
/* ------------------------------------------------------------------------ */
/* Restates the contract of kv's midrad as used by pidblas_assist::midrad
 * (vcp/pidblas.hpp:43-60): m is the upward-rounded average, r the upward-
 * rounded larger distance to an endpoint, point intervals give (x, 0).
 * Runs under FE_UPWARD set by the caller; operands come from memory. */
static void midrad_block(size_t cnt, const double* src /* AoS inf,sup */, double* M, double* R) {
    for (size_t t = 0; t < cnt; ++t) {
        const double lo = src[2 * t], hi = src[2 * t + 1];
        if (lo == hi) { M[t] = lo; R[t] = 0.0; continue; }
        const double h1 = 0.5 * lo;
        const double h2 = 0.5 * hi;
        const double mm = h1 + h2;
        const double r1 = hi - mm;
        const double r2 = mm - lo;
        M[t] = mm;
        R[t] = r1 > r2 ? r1 : r2;
    }
}

void orc_midrad(double inf, double sup, double* m, double* r) {
    double buf[2];
    volatile double* vb = buf;
    const int old = fegetround();
    fesetround(FE_UPWARD);
    vb[0] = inf; vb[1] = sup;
    midrad_block(1, buf, m, r);
    fesetround(old);
}

/* interval matrix (rows x cols, lda) -> mid / rad matrices (rows x cols, packed) */
static void midrad_mat(int rows, int cols, const double* IA, int lda, double* M, double* R) {
    const int nt = orc_get_threads();
#pragma omp parallel num_threads(nt)
    {
        const int old = fegetround();
        fesetround(FE_UPWARD);
#pragma omp for schedule(static)
        for (int j = 0; j < cols; ++j)
            midrad_block((size_t)rows, IA + 2 * (size_t)lda * j, M + (size_t)rows * j, R + (size_t)rows * j);
        fesetround(old);
    }
}

/* elementwise helpers running under an explicit mode (pidblas_assist
 * addmm_sequential / subsmmA_sequential, vcp/pidblas.hpp:61-84) */
static void ew_abs(size_t n, double* X) { for (size_t i = 0; i < n; ++i) X[i] = fabs(X[i]); }
static void ew_add(size_t n, double* X, const double* Y, int mode) {
    const int old = fegetround();
    fesetround(fe_of(mode));
    for (size_t i = 0; i < n; ++i) X[i] += Y[i];
    fesetround(old);
}
static void ew_sub(size_t n, double* X, const double* Y, int mode) {
    const int old = fegetround();
    fesetround(fe_of(mode));
    for (size_t i = 0; i < n; ++i) X[i] -= Y[i];
    fesetround(old);
}
static void pack_interval(int m, int n, const double* L, const double* U, double* C, int ldc) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            C[2 * ((size_t)i + (size_t)ldc * j)] = L[i + (size_t)m * j];
            C[2 * ((size_t)i + (size_t)ldc * j) + 1] = U[i + (size_t)m * j];
        }
}

/* pidblas::mulmm, vcp/pidblas.hpp:239-294 (same op order) */
int orc_pidblas_mulmm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                      double* C, int ldc) {
    if (m < 0 || n < 0 || k < 0) return -1;
    const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
    double *mA = dalloc(mk), *rA = dalloc(mk), *mB = dalloc(kn), *rB = dalloc(kn);
    double *lC = dalloc(mn), *uC = dalloc(mn), *tmp = dalloc(mn), *t2 = dalloc(mn);
    midrad_mat(m, k, A, lda, mA, rA);                              /* :260 */
    midrad_mat(k, n, B, ldb, mB, rB);                              /* :261 */
    orc_rgemm(m, n, k, mA, m, mB, k, lC, m, -1, 0);                /* :264-266 */
    orc_rgemm(m, n, k, mA, m, mB, k, uC, m, 1, 0);                 /* :269-271 */
    ew_abs(mk, mA);                                                /* :273 */
    ew_abs(kn, mB);                                                /* :275 */
    ew_add(mk, mA, rA, 1);                                         /* :277 */
    orc_rgemm(m, n, k, mA, m, rB, k, tmp, m, 1, 0);                /* :279 */
    orc_rgemm(m, n, k, rA, m, mB, k, t2, m, 1, 0);                 /* :281 */
    ew_add(mn, tmp, t2, 1);                                        /* :283 */
    ew_add(mn, uC, tmp, 1);                                        /* :284 */
    ew_sub(mn, lC, tmp, -1);                                       /* :285-286 */
    pack_interval(m, n, lC, uC, C, ldc);                           /* :289-293 */
    free(mA); free(rA); free(mB); free(rB); free(lC); free(uC); free(tmp); free(t2);
    return 0;
}

/* pidblas::mul_im_m, vcp/pidblas.hpp:90-140: C = IA * B (B point) */
int orc_pidblas_mul_im_m(int m, int n, int k, const double* IA, int lda, const double* B, int ldb,
                         double* C, int ldc) {
    const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
    double *mA = dalloc(mk), *rA = dalloc(mk), *aB = dalloc(kn);
    double *lC = dalloc(mn), *uC = dalloc(mn), *tmp = dalloc(mn);
    midrad_mat(m, k, IA, lda, mA, rA);                             /* :113 */
    orc_rgemm(m, n, k, mA, m, B, ldb, lC, m, -1, 0);               /* :116-118 */
    orc_rgemm(m, n, k, mA, m, B, ldb, uC, m, 1, 0);                /* :121-123 */
    for (int j = 0; j < n; ++j)                                    /* :125-127 */
        for (int i = 0; i < k; ++i) aB[i + (size_t)k * j] = fabs(B[i + (size_t)ldb * j]);
    orc_rgemm(m, n, k, rA, m, aB, k, tmp, m, 1, 0);                /* :129 */
    ew_add(mn, uC, tmp, 1);                                        /* :130 */
    ew_sub(mn, lC, tmp, -1);                                       /* :131-132 */
    pack_interval(m, n, lC, uC, C, ldc);
    free(mA); free(rA); free(aB); free(lC); free(uC); free(tmp);
    return 0;
}

/* pidblas::mul_m_im, vcp/pidblas.hpp:142-196: C = B * IA (B point m x k) */
int orc_pidblas_mul_m_im(int m, int n, int k, const double* B, int ldb, const double* IA, int lda,
                         double* C, int ldc) {
    const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
    double *mA = dalloc(kn), *rA = dalloc(kn), *aB = dalloc(mk);
    double *lC = dalloc(mn), *uC = dalloc(mn), *tmp = dalloc(mn);
    midrad_mat(k, n, IA, lda, mA, rA);                             /* :165 */
    orc_rgemm(m, n, k, B, ldb, mA, k, lC, m, -1, 0);               /* :168-174 */
    orc_rgemm(m, n, k, B, ldb, mA, k, uC, m, 1, 0);                /* :177-179 */
    for (int j = 0; j < k; ++j)                                    /* :181-183 */
        for (int i = 0; i < m; ++i) aB[i + (size_t)m * j] = fabs(B[i + (size_t)ldb * j]);
    orc_rgemm(m, n, k, aB, m, rA, k, tmp, m, 1, 0);                /* :185 */
    ew_add(mn, uC, tmp, 1);                                        /* :186 */
    ew_sub(mn, lC, tmp, -1);                                       /* :187-188 */
    pack_interval(m, n, lC, uC, C, ldc);
    free(mA); free(rA); free(aB); free(lC); free(uC); free(tmp);
    return 0;
}

/* pidblas::mulltmm, vcp/pidblas.hpp:296-358: C = IA^T IA (IA is k x n) */
int orc_pidblas_mulltmm(int k, int n, const double* IA, int lda, double* C, int ldc) {
    const size_t kn = (size_t)k * n, nn = (size_t)n * n;
    double *mA = dalloc(kn), *rA = dalloc(kn), *mAt = dalloc(kn), *S = dalloc(nn);
    double *lC = dalloc(nn), *uC = dalloc(nn), *tmp = dalloc(nn);
    midrad_mat(k, n, IA, lda, mA, rA);                             /* :306 */
    orc_rsyrk(k, n, mA, k, lC, n, -1);                             /* :308-310 */
    orc_rsyrk(k, n, mA, k, uC, n, 1);                              /* :313-315 */
    ew_abs(kn, mA);                                                /* :317 */
    transpose(k, n, mA, k, mAt);                                   /* :319 */
    orc_rgemm(n, n, k, mAt, n, rA, k, tmp, n, 1, 0);               /* :321 */
    orc_rsyrk(k, n, rA, k, S, n, 1);                               /* :323 */
    {
        const int old = fegetround();
        fesetround(FE_UPWARD);                                     /* still RU (:313) */
        for (int i = 0; i < n; ++i)                                /* :324-328 */
            for (int j = i; j < n; ++j)
                tmp[i + (size_t)n * j] += tmp[j + (size_t)n * i] + S[i + (size_t)n * j];
        fesetround(old);
    }
    for (int i = 0; i < n; ++i)                                    /* :329-333 */
        for (int j = i + 1; j < n; ++j) tmp[j + (size_t)n * i] = tmp[i + (size_t)n * j];
    ew_add(nn, uC, tmp, 1);                                        /* :336 */
    ew_sub(nn, lC, tmp, -1);                                       /* :337-338 */
    pack_interval(n, n, lC, uC, C, ldc);                           /* :341-346 */
    for (int i = 0; i < n; ++i)                                    /* :348-352 */
        for (int j = i + 1; j < n; ++j) {
            C[2 * ((size_t)j + (size_t)ldc * i)] = C[2 * ((size_t)i + (size_t)ldc * j)];
            C[2 * ((size_t)j + (size_t)ldc * i) + 1] = C[2 * ((size_t)i + (size_t)ldc * j) + 1];
        }
    free(mA); free(rA); free(mAt); free(S); free(lC); free(uC); free(tmp);
    return 0;
}

/* pidblas::vmulmm, vcp/pidblas.hpp:198-236: IC = C^T C for point C (k x n) */
int orc_pidblas_vmulmm(int k, int n, const double* A, int lda, double* C, int ldc) {
    const size_t nn = (size_t)n * n;
    double *U = dalloc(nn), *L = dalloc(nn);
    orc_rsyrk(k, n, A, lda, U, n, 1);                              /* :202-203 */
    orc_rsyrk(k, n, A, lda, L, n, -1);                             /* :219-220 */
    pack_interval(n, n, L, U, C, ldc);
    free(U); free(L);
    return 0;
}

/* vcp::ifma, the branch without AVX-512 (vcp/ifma.hpp:58-137): the interval
   FMA [cd, cu] + [ad, au] * [bd, bu], each end one directed FMA of the
   endpoint product the signs select.  An operand is classified as
   nonnegative (lower >= 0), straddling (lower <= 0 <= upper) or nonpositive,
   tested in that order (ifma.hpp:67, :86, :111). */
enum { ORC_POS = 0, ORC_MIX = 1, ORC_NEG = 2 };
static int sign_class(double lo, double hi) { return lo >= 0 ? ORC_POS : (hi >= 0 ? ORC_MIX : ORC_NEG); }

static void ifma_restated(double au, double ad, double bu, double bd, double cu, double cd, double* up, double* dn) {
    const int ca = sign_class(ad, au), cb = sign_class(bd, bu);
    volatile double x, y;
    if (ca == ORC_MIX && cb == ORC_MIX) {
        fesetround(FE_UPWARD);     /* larger of au*bu and ad*bd (:97-100) */
        x = fma(au, bu, cu);
        y = fma(ad, bd, cu);
        *up = (y < x) ? x : y;
        fesetround(FE_DOWNWARD);   /* smaller of au*bd and ad*bu (:102-105) */
        x = fma(au, bd, cd);
        y = fma(ad, bu, cd);
        *dn = (y > x) ? x : y;
        fesetround(FE_TONEAREST);
        return;
    }
    /* endpoint pairs {a, b} for the upper and the lower end, per sign class
       (rows: a; columns: b; 0 = lower end, 1 = upper end) */
    static const int up_a[3][3] = {{1, 1, 0}, {1, -1, 0}, {1, 0, 0}};
    static const int up_b[3][3] = {{1, 1, 1}, {1, -1, 0}, {0, 0, 0}};
    static const int dn_a[3][3] = {{0, 1, 1}, {0, -1, 1}, {0, 0, 1}};
    static const int dn_b[3][3] = {{0, 0, 0}, {1, -1, 0}, {1, 1, 1}};
    const double a_end[2] = {ad, au}, b_end[2] = {bd, bu};
    fesetround(FE_UPWARD);
    x = fma(a_end[up_a[ca][cb]], b_end[up_b[ca][cb]], cu);
    fesetround(FE_DOWNWARD);
    y = fma(a_end[dn_a[ca][cb]], b_end[dn_b[ca][cb]], cd);
    fesetround(FE_TONEAREST);
    *up = x;
    *dn = y;
}

/* pidblas_fma::mulmm, vcp/pidblas_fma.hpp:45-72: C = IA * IB in inf-sup
   form, every element the chain of interval FMAs over the inner index in
   ascending order from [0, 0] (:62-71) */
int orc_pidblas_fma_mulmm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                          double* C, int ldc) {
    if (m < 0 || n < 0 || k < 0) return -1;
#pragma omp parallel for schedule(dynamic) num_threads(orc_get_threads())
    for (int col = 0; col < n; ++col) {
        for (int i = 0; i < m; ++i) {
            double cu = 0.0, cd = 0.0;
            for (int j = 0; j < k; ++j) {
                const double* a = A + 2 * (i + (size_t)lda * j);
                const double* b = B + 2 * (j + (size_t)ldb * col);
                ifma_restated(a[1], a[0], b[1], b[0], cu, cd, &cu, &cd);
            }
            C[2 * (i + (size_t)ldc * col)] = cd;
            C[2 * (i + (size_t)ldc * col) + 1] = cu;
        }
    }
    return 0;
}

/* imats::linearsolve R*A - I segment (vcp/imats.hpp:318-330) and
 * mats::norminf for type 'M' (vcp/mats.hpp:1583-1600) on intervals:
 * |[l,u]| = [mig, mag] and the row sum is an interval sum, so the upper
 * bound is the upward sequential sum of the magnitudes in column order. */
int orc_residual(int n, const double* R, int ldr, const double* IA, int lda, double* C, int ldc,
                 double* normub) {
    orc_pidblas_mul_m_im(n, n, n, R, ldr, IA, lda, C, ldc);        /* imats.hpp:319 */
    {
        const int old = fegetround();
        for (int i = 0; i < n; ++i) {                              /* imats.hpp:324-326 */
            double* c = C + 2 * ((size_t)i + (size_t)ldc * i);
            volatile double lo = c[0], hi = c[1];
            fesetround(FE_DOWNWARD);
            c[0] = lo - 1.0;
            fesetround(FE_UPWARD);
            c[1] = hi - 1.0;
        }
        fesetround(FE_UPWARD);
        double best = 0.0;
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) {
                const double lo = C[2 * ((size_t)i + (size_t)ldc * j)];
                const double hi = C[2 * ((size_t)i + (size_t)ldc * j) + 1];
                const double mag = fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi);
                s = (j == 0) ? mag : s + mag;
            }
            if (i == 0 || s > best) best = s;
        }
        fesetround(old);
        *normub = best;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* verified linear solve (imats::linearsolve, vcp/imats.hpp:292-362)         */
/* ------------------------------------------------------------------------ */
/* kv's mid() as restated here: the nearest-rounded (lo + hi) / 2, or
 * lo/2 + hi/2 where lo + hi overflows; point intervals give lo.  (kv's own
 * rounding choice is unpinned: kv is not vendored.) */
static double kv_mid(double lo, double hi) {
    if (lo == hi) return lo;
    volatile double s = lo + hi;
    if (isinf(s)) return 0.5 * lo + 0.5 * hi;
    return s * 0.5;
}

/* upper end of mats::norminf of an interval matrix (vcp/mats.hpp:1565-1604):
 * a column ('C') gives the largest magnitude, a matrix ('M') the largest
 * upward row sum of magnitudes (each sum from column 0 up); rows >= 2.
 * The maximum keeps the first row's value and replaces it only by larger
 * ones (so a NaN in row 0 stays, later NaNs are passed over). */
static double norminf_ub(int rows, int cols, const double* X, int ldx) {
    const int old = fegetround();
    fesetround(FE_UPWARD);
    double best = 0.0;
    for (int i = 0; i < rows; ++i) {
        volatile double s = 0.0;
        for (int j = 0; j < cols; ++j) {
            const double lo = X[2 * ((size_t)i + (size_t)ldx * j)], hi = X[2 * ((size_t)i + (size_t)ldx * j) + 1];
            const double mag = fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi);
            s = (j == 0) ? mag : s + mag;
        }
        if (i == 0 || s > best) best = s;
    }
    fesetround(old);
    return best;
}

static void interval_abs(size_t cnt, double* X) {   /* kv abs(), in place */
    for (size_t t = 0; t < cnt; ++t) {
        const double lo = X[2 * t], hi = X[2 * t + 1];
        if (lo >= 0.0) continue;
        if (hi <= 0.0) { X[2 * t] = -hi; X[2 * t + 1] = -lo; continue; }
        X[2 * t] = 0.0;
        X[2 * t + 1] = -lo > hi ? -lo : hi;
    }
}

int orc_lss(int n, int nrhs, const double* IA, int lda, const double* Ib, int ldb, const double* R, int ldr,
            double* X, int ldx, double* normub) {
    if (n < 2 || nrhs < 1 || lda < n || ldb < n || ldr < n || ldx < n) return -1;
    const size_t nr = (size_t)n * nrhs;
    double *mb = dalloc(nr), *mx = dalloc(nr), *Xi = dalloc(2 * nr), *Y = dalloc(2 * nr), *Tm = dalloc(2 * nr);
    double *W = dalloc(2 * nr), *C = dalloc(2 * (size_t)n * n);
    for (int j = 0; j < nrhs; ++j)                                  /* mb = mid(b), :299 */
        for (int i = 0; i < n; ++i)
            mb[i + (size_t)n * j] = kv_mid(Ib[2 * (i + (size_t)ldb * j)], Ib[2 * (i + (size_t)ldb * j) + 1]);
    orc_rgemm(n, nrhs, n, R, ldr, mb, n, mx, n, 0, 0);               /* mx = R mb, :302 */
    orc_pidblas_mul_im_m(n, nrhs, n, IA, lda, mx, n, Xi, n);         /* x = A mx, :306 */
    {                                                               /* x = x - b, :308 */
        const int old = fegetround();
        for (int j = 0; j < nrhs; ++j)
            for (int i = 0; i < n; ++i) {
                double* x = Xi + 2 * (i + (size_t)n * j);
                const double* b = Ib + 2 * (i + (size_t)ldb * j);
                volatile double xl = x[0], xu = x[1], bl = b[0], bu = b[1];
                fesetround(FE_DOWNWARD);
                x[0] = xl - bu;
                fesetround(FE_UPWARD);
                x[1] = xu - bl;
            }
        fesetround(old);
    }
    orc_pidblas_mul_m_im(n, nrhs, n, R, ldr, Xi, n, Y, n);          /* b = R x, :310 */
    interval_abs(nr, Y);                                            /* x = |b|, :312-315 */
    double g = 0.0;
    orc_residual(n, R, ldr, IA, lda, C, n, &g);                     /* R A - I, ||.||, :318-329 */
    *normub = g;
    int rc = 0;
    if (!(g < 1.0)) {                                               /* :330-333 */
        rc = 1;
        goto done;
    }
    {
        const double t = norminf_ub(n, nrhs, Y, n);                  /* T = ||x||, :335-336 */
        interval_abs((size_t)n * n, C);                             /* b = |RA - I|, :338 */
        double glo, ghi;
        {                                                           /* G = T / (1 - G), :341 */
            const int old = fegetround();
            volatile double one = 1.0, gv = g, tv = t;
            fesetround(FE_DOWNWARD);
            volatile double ddn = one - gv;
            fesetround(FE_UPWARD);
            volatile double dup = one - gv;
            fesetround(FE_DOWNWARD);
            glo = tv / dup;
            fesetround(FE_UPWARD);
            ghi = tv / ddn;
            fesetround(old);
        }
        for (size_t q = 0; q < nr; ++q) { Tm[2 * q] = glo; Tm[2 * q + 1] = ghi; }   /* :343-348 */
        orc_pidblas_mulmm(n, nrhs, n, C, n, Tm, n, W, n);            /* G = |RA - I| T, :350 */
        const int old = fegetround();
        for (int j = 0; j < nrhs; ++j)                              /* :352-361 */
            for (int i = 0; i < n; ++i) {
                const size_t q = i + (size_t)n * j;
                fesetround(FE_UPWARD);
                volatile double u = Y[2 * q + 1] + W[2 * q + 1];
                volatile double m = mx[q];
                fesetround(FE_DOWNWARD);
                X[2 * (i + (size_t)ldx * j)] = -u + m;
                fesetround(FE_UPWARD);
                X[2 * (i + (size_t)ldx * j) + 1] = u + m;
            }
        fesetround(old);
    }
done:
    free(mb); free(mx); free(Xi); free(Y); free(Tm); free(W); free(C);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Galerkin matrix: Legendre_Bases_Generator::uhphiphi (vcp/ldbase.hpp:955-1013) */
/* ------------------------------------------------------------------------ */
/* kv's interval product: the lower end is the smallest downward-rounded
 * endpoint product, the upper end the largest upward-rounded one. */
void orc_imul(double al, double au, double bl, double bu, double* lo, double* hi) {
    const int old = fegetround();
    volatile double a0 = al, a1 = au, b0 = bl, b1 = bu;
    fesetround(FE_DOWNWARD);
    double p[4] = {a0 * b0, a0 * b1, a1 * b0, a1 * b1};
    double l = p[0];
    for (int t = 1; t < 4; ++t) if (p[t] < l) l = p[t];
    fesetround(FE_UPWARD);
    double q[4] = {a0 * b0, a0 * b1, a1 * b0, a1 * b1};
    double h = q[0];
    for (int t = 1; t < 4; ++t) if (q[t] > h) h = q[t];
    fesetround(old);
    *lo = l;
    *hi = h;
}

int orc_uhphiphi(int m, int npts, const double* phi, int ldp, const double* w, const double* u, double* A,
                 int lda) {
    if (m < 0 || npts < 0 || ldp < (m > 1 ? m : 1) || lda < (m > 1 ? m : 1)) return -1;
    const int nt = orc_get_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int i = 0; i < m; ++i) {
        for (int j = i; j < m; ++j) {
            double lo = 0.0, hi = 0.0;
            for (int k = 0; k < npts; ++k) {
                double vl, vh, sl, sh, pl, ph;
                orc_imul(w[2 * k], w[2 * k + 1], u[2 * k], u[2 * k + 1], &vl, &vh);
                orc_imul(vl, vh, phi[2 * (i + (size_t)ldp * k)], phi[2 * (i + (size_t)ldp * k) + 1], &sl, &sh);
                orc_imul(sl, sh, phi[2 * (j + (size_t)ldp * k)], phi[2 * (j + (size_t)ldp * k) + 1], &pl, &ph);
                const int old = fegetround();
                volatile double x = lo, y = hi;
                fesetround(FE_DOWNWARD);
                lo = x + pl;
                fesetround(FE_UPWARD);
                hi = y + ph;
                fesetround(old);
            }
            A[2 * (i + (size_t)lda * j)] = lo;
            A[2 * (i + (size_t)lda * j) + 1] = hi;
            A[2 * (j + (size_t)lda * i)] = lo;
            A[2 * (j + (size_t)lda * i) + 1] = hi;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Exact arithmetic: superaccumulator                                       */
/* ------------------------------------------------------------------------ */
#define SA_EMIN (-2208) /* bit weight of digit 0's least significant bit   */
#define SA_N 140        /* 32-bit digits held in int64 slots (4480 bits)   */
typedef struct { int64_t d[SA_N]; } sacc;

static void dsplit(double x, int* neg, uint64_t* mant, int* ex) {
    uint64_t u;
    memcpy(&u, &x, 8);
    *neg = (int)(u >> 63);
    const int be = (int)((u >> 52) & 0x7ff);
    const uint64_t f = u & ((1ull << 52) - 1);
    if (be == 0) { *mant = f; *ex = -1074; }
    else { *mant = f | (1ull << 52); *ex = be - 1075; }
}

typedef struct { int neg; u128 mag; int e; } xprod; /* value = (-1)^neg mag 2^e */

static xprod xmul(double a, double b) {
    int na, nb, ea, eb;
    uint64_t ma, mb;
    dsplit(a, &na, &ma, &ea);
    dsplit(b, &nb, &mb, &eb);
    xprod p;
    p.mag = (u128)ma * mb;
    p.e = ea + eb;
    p.neg = (na ^ nb) && p.mag != 0;
    return p;
}
static xprod xdbl(double a) {
    int na, ea;
    uint64_t ma;
    dsplit(a, &na, &ma, &ea);
    xprod p;
    p.mag = ma;
    p.e = ea;
    p.neg = na && ma != 0;
    return p;
}
static int u128_bitlen(u128 x) {
    const uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
    if (hi) return 128 - __builtin_clzll(hi);
    if (lo) return 64 - __builtin_clzll(lo);
    return 0;
}
/* compare |x| and |y| */
static int xcmp_mag(xprod x, xprod y) {
    if (x.mag == 0 || y.mag == 0) return (x.mag != 0) - (y.mag != 0);
    const int bx = u128_bitlen(x.mag), by = u128_bitlen(y.mag);
    const int px = x.e + bx, py = y.e + by;
    if (px != py) return px > py ? 1 : -1;
    u128 a = x.mag, b = y.mag;
    if (bx < by) a <<= (by - bx);
    else if (by < bx) b <<= (bx - by);
    return (a > b) - (a < b);
}
static int xcmp(xprod x, xprod y) {
    const int sx = x.mag == 0 ? 0 : (x.neg ? -1 : 1);
    const int sy = y.mag == 0 ? 0 : (y.neg ? -1 : 1);
    if (sx != sy) return sx > sy ? 1 : -1;
    if (sx == 0) return 0;
    const int c = xcmp_mag(x, y);
    return sx > 0 ? c : -c;
}

static void sa_zero(sacc* s) { memset(s, 0, sizeof(*s)); }
static void sa_add(sacc* s, xprod p) {
    if (p.mag == 0) return;
    const int pos = p.e - SA_EMIN;
    const int li = pos >> 5, sh = pos & 31;
    for (int c = 0; c < 4; ++c) {
        const uint64_t piece = (uint64_t)(p.mag >> (32 * c)) & 0xffffffffull;
        if (!piece) continue;
        const uint64_t sp = piece << sh;
        const int64_t lo = (int64_t)(sp & 0xffffffffull), hi = (int64_t)(sp >> 32);
        if (p.neg) { s->d[li + c] -= lo; s->d[li + c + 1] -= hi; }
        else { s->d[li + c] += lo; s->d[li + c + 1] += hi; }
    }
}
static void sa_add_d(sacc* s, double x, int negate) {
    xprod p = xdbl(x);
    if (negate) p.neg = !p.neg && p.mag != 0;
    sa_add(s, p);
}
/* carry-normalise: digits 0..N-2 in [0, 2^32), top digit signed */
static void sa_norm(sacc* s) {
    for (int i = 0; i < SA_N - 1; ++i) {
        const int64_t carry = s->d[i] >> 32; /* arithmetic shift = floor */
        s->d[i] -= carry * ((int64_t)1 << 32);
        s->d[i + 1] += carry;
    }
}
static int sa_sign(const sacc* s0) {
    sacc s = *s0;
    sa_norm(&s);
    if (s.d[SA_N - 1] < 0) return -1;
    for (int i = SA_N - 1; i >= 0; --i)
        if (s.d[i] != 0) return 1;
    return 0;
}
/* magnitude digits (all in [0,2^32)) and sign */
static int sa_mag(const sacc* s0, uint32_t* mag) {
    sacc s = *s0;
    sa_norm(&s);
    int neg = s.d[SA_N - 1] < 0;
    if (neg) {
        for (int i = 0; i < SA_N; ++i) s.d[i] = -s.d[i];
        sa_norm(&s);
    }
    for (int i = 0; i < SA_N; ++i) mag[i] = (uint32_t)s.d[i];
    return neg;
}
static int mag_bit(const uint32_t* mag, int b) {
    if (b < 0) return 0;
    return (int)((mag[b >> 5] >> (b & 31)) & 1u);
}
static int mag_any_below(const uint32_t* mag, int b) { /* any bit < b set */
    if (b <= 0) return 0;
    const int full = b >> 5;
    for (int i = 0; i < full; ++i)
        if (mag[i]) return 1;
    const int r = b & 31;
    if (r && (mag[full] & ((1u << r) - 1u))) return 1;
    return 0;
}
/* round the exact value to double: dir 1 up, -1 down, 0 nearest-even */
static double sa_round(const sacc* s, int dir) {
    uint32_t mag[SA_N];
    const int neg = sa_mag(s, mag);
    int top = -1;
    for (int i = SA_N - 1; i >= 0; --i)
        if (mag[i]) { top = i * 32 + 31 - __builtin_clz(mag[i]); break; }
    if (top < 0) return 0.0;
    /* bit b has weight 2^(b + SA_EMIN) */
    int q = top - 52;
    if (q + SA_EMIN < -1074) q = -1074 - SA_EMIN;
    uint64_t mant = 0;
    for (int b = top; b >= q; --b) mant = (mant << 1) | (uint64_t)mag_bit(mag, b);
    const int rbit = mag_bit(mag, q - 1);
    const int sticky = mag_any_below(mag, q - 1);
    int up;
    if (dir == 0) up = rbit && (sticky || (mant & 1));
    else {
        const int away = (dir > 0) ? !neg : neg;
        up = away && (rbit || sticky);
    }
    mant += (uint64_t)up;
    const int ex = q + SA_EMIN;
    double v;
    if (ex + 64 - __builtin_clzll(mant) > 1024) {
        /* overflow */
        const int away = dir == 0 ? 1 : ((dir > 0) ? !neg : neg);
        v = away ? INFINITY : DBL_MAX;
    } else {
        v = ldexp((double)mant, ex);
    }
    return neg ? -v : v;
}

void orc_exact_dot(int k, const double* a, int astride, const double* b, int bstride,
                   double* rd, double* ru, double* rn) {
    sacc s;
    sa_zero(&s);
    for (int t = 0; t < k; ++t) sa_add(&s, xmul(a[(size_t)t * astride], b[(size_t)t * bstride]));
    if (rd) *rd = sa_round(&s, -1);
    if (ru) *ru = sa_round(&s, 1);
    if (rn) *rn = sa_round(&s, 0);
}

/* exact hull of sum_k [a_k] * [b_k]: lower = sum of per-k minimum products,
 * upper = sum of per-k maximum products (interval multiplication is exact
 * on the endpoints' products). */
static void hull_acc(int k, const double* IArow, int rowstride, const double* IBcol, int colstride,
                     sacc* lo, sacc* hi) {
    sa_zero(lo);
    sa_zero(hi);
    for (int t = 0; t < k; ++t) {
        const double a1 = IArow[2 * (size_t)t * rowstride], a2 = IArow[2 * (size_t)t * rowstride + 1];
        const double b1 = IBcol[2 * (size_t)t * colstride], b2 = IBcol[2 * (size_t)t * colstride + 1];
        xprod p[4] = {xmul(a1, b1), xmul(a1, b2), xmul(a2, b1), xmul(a2, b2)};
        int imin = 0, imax = 0;
        for (int u = 1; u < 4; ++u) {
            if (xcmp(p[u], p[imin]) < 0) imin = u;
            if (xcmp(p[u], p[imax]) > 0) imax = u;
        }
        sa_add(lo, p[imin]);
        sa_add(hi, p[imax]);
    }
}

void orc_exact_hull_entry(int k, const double* IArow, int rowstride, const double* IBcol,
                          int colstride, double c_lower, double c_upper,
                          int* ok_lower, int* ok_upper, double* hull_lo, double* hull_hi) {
    sacc lo, hi;
    hull_acc(k, IArow, rowstride, IBcol, colstride, &lo, &hi);
    if (hull_lo) *hull_lo = sa_round(&lo, -1);
    if (hull_hi) *hull_hi = sa_round(&hi, 1);
    sacc t = lo;
    sa_add_d(&t, c_lower, 1); /* lo* - C_lower */
    if (ok_lower) *ok_lower = sa_sign(&t) >= 0;
    t = hi;
    sa_add_d(&t, c_upper, 1); /* hi* - C_upper */
    if (ok_upper) *ok_upper = sa_sign(&t) <= 0;
}

void orc_check_containment(int k, const double* IA, int lda, const double* IB, int ldb,
                           const double* C, int ldc, int nsamp, const int* ii, const int* jj,
                           int* ok, double* hull_lo, double* hull_hi) {
    const int nt = orc_get_threads();
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4)
    for (int s = 0; s < nsamp; ++s) {
        const int i = ii[s], j = jj[s];
        int okl = 0, oku = 0;
        double hl = 0, hh = 0;
        const double* c = C + 2 * ((size_t)i + (size_t)ldc * j);
        orc_exact_hull_entry(k, IA + 2 * (size_t)i, lda, IB + 2 * (size_t)ldb * j, 1, c[0], c[1],
                             &okl, &oku, &hl, &hh);
        ok[s] = okl && oku;
        if (hull_lo) hull_lo[s] = hl;
        if (hull_hi) hull_hi[s] = hh;
    }
}

void orc_exact_dd_dot(int k, const double* a, int astride, const double* b, int bstride,
                      double ch, double cl, double* ex_hi, double* ex_lo, double* err_ru) {
    sacc s;
    sa_zero(&s);
    for (int t = 0; t < k; ++t) {
        const double ah = a[2 * (size_t)t * astride], al = a[2 * (size_t)t * astride + 1];
        const double bh = b[2 * (size_t)t * bstride], bl = b[2 * (size_t)t * bstride + 1];
        sa_add(&s, xmul(ah, bh));
        sa_add(&s, xmul(ah, bl));
        sa_add(&s, xmul(al, bh));
        sa_add(&s, xmul(al, bl));
    }
    const double h = sa_round(&s, 0);
    sacc r = s;
    sa_add_d(&r, h, 1);
    const double l = sa_round(&r, 0);
    if (ex_hi) *ex_hi = h;
    if (ex_lo) *ex_lo = l;
    if (err_ru) {
        sacc e = s;
        sa_add_d(&e, ch, 1);
        sa_add_d(&e, cl, 1);
        const double up = sa_round(&e, 1), dn = sa_round(&e, -1);
        const double a1 = fabs(up), a2 = fabs(dn);
        *err_ru = a1 > a2 ? a1 : a2;
    }
}

/* ------------------------------------------------------------------------ */
/* Ozaki scheme II interval GEMM restatement                                */
/* ------------------------------------------------------------------------ */
#define OZ_CHUNK 256 /* k-chunk of the upward 1-norm sums */

static int gcd_i(int a, int b) { while (b) { int t = a % b; a = b; b = t; } return a; }

int orc_ozaki_moduli(int nmod, int* out) {
    int cnt = 0;
    for (int m = 256; m >= 2 && cnt < nmod; --m) {
        int ok = 1;
        for (int t = 0; t < cnt; ++t)
            if (gcd_i(m, out[t]) != 1) { ok = 0; break; }
        if (ok) out[cnt++] = m;
    }
    return cnt;
}

int orc_ozaki_beta(int k, int nmod) {
    int mods[64];
    if (nmod > 64 || orc_ozaki_moduli(nmod, mods) != nmod) return -1;
    double lm = 0.0;
    for (int t = 0; t < nmod; ++t) lm += log2((double)mods[t]);
    int lk = 0;
    while ((1ll << lk) < (long long)(k > 1 ? k : 1)) ++lk; /* ceil(log2 k) */
    /* k 2^(2 beta + 1) < M  with a safety margin for log2 rounding */
    const int beta = (int)floor((lm - 1.0 - (double)lk - 1e-6) / 2.0);
    return beta > 62 ? 62 : beta;
}

/* frexp exponent of a positive value: 2^(e-1) <= x < 2^e */
static int fexp(double x) { int e; frexp(x, &e); return e; }

/* x * 2^e under the caller's rounding mode, in steps whose scale factors
 * stay normal so only the final underflow (if any) rounds */
static double mulpow2(double x, int e) {
    while (e > 1000) { x = x * ldexp(1.0, 1000); e -= 1000; }
    while (e < -1000) { x = x * ldexp(1.0, -1000); e += 1000; }
    return x * ldexp(1.0, e);
}

/* nearest integer (ties up) of x, computed without the rounding mode:
 * x - floor(x) is exact for the magnitudes used here */
static int64_t rnint_floor(double x) {
    const double f = floor(x);
    const double r = x - f;
    return (int64_t)f + (r >= 0.5 ? 1 : 0);
}

/* two upward 7-bit digits of x >= 0 against 2^g (x < 2^g), under FE_UPWARD:
 * t = RU(127 x 2^-g), h = floor(t), l = ceil(RU(127 (t - h))); then
 * x <= (h + l/127) 2^g / 127 with h, l in [0, 127]. */
static void digits2_ru(double x, int g, int* h, int* l) {
    const double t = mulpow2(x, -g) * 127.0;
    const double hf = floor(t);
    const double r = (t - hf) * 127.0;
    *h = (int)hf;
    *l = (int)ceil(r);
}

/* The same bound of x written with a non-positive low digit:
 *   x <= (H + L/127) 2^g / 127,  H = h + [l > 0] in [0, 127],  L = l - 127 [l > 0] in [-126, 0].
 * H + L/127 = h + l/127, and H <= 127 because t <= 127 forces l = 0 when
 * h = 127.  For a factor x1 <= (h1 + l1/127) s1 (l1 >= 0) and x2 held this way,
 *   x1 x2 <= s1 s2 (h1 H + (h1 L + l1 H)/127 + l1 L/127^2)
 * and the last term is never positive, so
 *   sum_k x1 x2 <= s1 s2 (sum h1 H + (sum h1 L + l1 H)/127).            */
static void digits2_ru_neglo(double x, int g, int* H, int* L) {
    int h, l;
    digits2_ru(x, g, &h, &l);
    *H = h + (l > 0);
    *L = l > 0 ? l - 127 : 0;
}

/* X = hi 2^64 + lo (lo unsigned; |X| < 2^190) times 2^s, rounded toward -inf
 * (dir<0) or +inf (dir>0), exactly.  The interval form's scaled midpoints
 * reach |A'|, |B'| = 2^62, so a dot product of k of their products needs up
 * to 124 + log2 k + 1 bits: more than one i128 holds. */
static double conv_x192(i128 hi, uint64_t lo, int s, int dir) {
    const int neg = hi < 0;
    u128 mh;
    uint64_t ml;
    if (!neg) { mh = (u128)hi; ml = lo; }
    else if (lo == 0) { mh = (u128)(-hi); ml = 0; }
    else { mh = (u128)(-(hi + 1)); ml = (uint64_t)0 - lo; } /* 2^64 - lo */
    uint32_t x[6];
    for (int i = 0; i < 2; ++i) x[i] = (uint32_t)(ml >> (32 * i));
    for (int i = 0; i < 4; ++i) x[2 + i] = (uint32_t)(mh >> (32 * i));
    int top = -1;
    for (int i = 5; i >= 0 && top < 0; --i)
        if (x[i]) top = 32 * i + 31 - __builtin_clz(x[i]);
    if (top < 0) return 0.0;
#define X192_BIT(b) ((b) < 0 || (b) >= 192 ? 0u : (x[(b) >> 5] >> ((b) & 31)) & 1u)
    int q = top - 52;
    if (q < -1074 - s) q = -1074 - s;
    if (q < 0) q = 0; /* exact: the value lies on the double grid */
    uint64_t mant = 0;
    for (int b = top; b >= q; --b) mant = (mant << 1) | X192_BIT(b);
    int rest = 0;
    for (int b = 0; b < q && !rest; ++b) rest = (int)X192_BIT(b);
#undef X192_BIT
    const int away = dir > 0 ? !neg : neg;
    if (rest && away) mant += 1;
    const int ex = q + s;
    double v;
    if (mant != 0 && ex + 64 - __builtin_clzll(mant) > 1024) v = away ? INFINITY : DBL_MAX;
    else v = ldexp((double)mant, ex);
    return neg ? -v : v;
}

enum { OZ_MIDZERO = 1, OZ_SUMZERO = 2, OZ_RADZERO = 4 };
typedef struct {
    int beta;
    int *e, *g, *h;      /* mid exponent, |m|+r digit scale, r digit scale */
    double *n1;          /* upward 1-norm of |mid| */
    int *flags;          /* OZ_*ZERO: mid / |mid|+rad / rad identically zero */
} oz_side;

/* statistics of the rows of a mid/rad operand: element c of row r is at
 * r*rs + c*cs (rows of A: rs=1, cs=m; columns of B: rs=k, cs=1).  Upward
 * 1-norm order: lanes == 0 -- sequential inside k-chunks of OZ_CHUNK, then
 * over the chunks in order (rows of A); lanes == 64 -- partial l sums the
 * elements l, l+64, ... in order, then the 64 partials in order (columns of
 * B, one GPU wave per column). */
static void oz_stats(int rows, int len, const double* M, const double* R, size_t rs, size_t cs,
                     int beta, oz_side* S, int lanes) {
    S->beta = beta;
    S->e = (int*)malloc(sizeof(int) * rows);
    S->g = (int*)malloc(sizeof(int) * rows);
    S->h = (int*)malloc(sizeof(int) * rows);
    S->flags = (int*)malloc(sizeof(int) * rows);
    S->n1 = dalloc(rows);
    const int old = fegetround();
    fesetround(FE_UPWARD);
    for (int r = 0; r < rows; ++r) {
        double mu = 0.0, smax = 0.0, rmax = 0.0, tot = 0.0;
        const int step = lanes > 0 ? 1 : OZ_CHUNK;
        const int nblk = lanes > 0 ? lanes : (len + OZ_CHUNK - 1) / OZ_CHUNK;
        for (int b = 0; b < nblk; ++b) {
            double part = 0.0;
            const int c0 = lanes > 0 ? b : b * step;
            const int c1 = lanes > 0 ? len : ((c0 + OZ_CHUNK < len) ? c0 + OZ_CHUNK : len);
            const int inc = lanes > 0 ? lanes : 1;
            for (int c = c0; c < c1; c += inc) {
                const size_t idx = (size_t)r * rs + (size_t)c * cs;
                const double am = fabs(M[idx]), ar = R[idx];
                const double as = am + ar;
                if (am > mu) mu = am;
                if (as > smax) smax = as;
                if (ar > rmax) rmax = ar;
                part = part + am;
            }
            tot = tot + part;
        }
        S->flags[r] = (mu == 0.0 ? OZ_MIDZERO : 0) | (smax == 0.0 ? OZ_SUMZERO : 0) | (rmax == 0.0 ? OZ_RADZERO : 0);
        S->e[r] = mu > 0.0 ? fexp(mu) : 0;
        S->g[r] = smax > 0.0 ? fexp(smax) : 0;
        S->h[r] = rmax > 0.0 ? fexp(rmax) : 0;
        S->n1[r] = tot;
    }
    fesetround(old);
}
static void oz_free(oz_side* S) { free(S->e); free(S->g); free(S->h); free(S->flags); free(S->n1); }

void orc_ozaki_scaled_rows(int m, int k, const double* A, int lda, int beta, int64_t* Ap, int* e_out) {
    double *mA = dalloc((size_t)m * k), *rA = dalloc((size_t)m * k);
    midrad_mat(m, k, A, lda, mA, rA);
    oz_side S;
    oz_stats(m, k, mA, rA, 1, (size_t)m, beta, &S, 0);
    const int old = fegetround();
    fesetround(FE_UPWARD);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < m; ++i)
            Ap[i + (size_t)m * j] = (S.flags[i] & OZ_MIDZERO) ? 0 : rnint_floor(mulpow2(mA[i + (size_t)m * j], beta - S.e[i]));
    fesetround(old);
    for (int i = 0; i < m; ++i) e_out[i] = S.e[i];
    oz_free(&S);
    free(mA); free(rA);
}
void orc_ozaki_scaled_cols(int k, int n, const double* B, int ldb, int beta, int64_t* Bp, int* f_out) {
    double *mB = dalloc((size_t)k * n), *rB = dalloc((size_t)k * n);
    midrad_mat(k, n, B, ldb, mB, rB);
    oz_side S;
    oz_stats(n, k, mB, rB, (size_t)k, 1, beta, &S, 64);
    const int old = fegetround();
    fesetround(FE_UPWARD);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < k; ++i)
            Bp[i + (size_t)k * j] = (S.flags[j] & OZ_MIDZERO) ? 0 : rnint_floor(mulpow2(mB[i + (size_t)k * j], beta - S.e[j]));
    fesetround(old);
    for (int j = 0; j < n; ++j) f_out[j] = S.e[j];
    oz_free(&S);
    free(mB); free(rB);
}

/* The product IA*IB of pidblas::mulmm (vcp/pidblas.hpp:239-294) through the
 * Ozaki-II route: exact integer product of the scaled midpoints (the value
 * the CRT reconstruction recovers), one directed conversion each way, the
 * scaling-error bound E and the upward two-digit bound of the radius term
 * (|mA|+rA) rB + rA |mB| (:277-283). */
static int oz_igemm_core(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                         double* C, int ldc, int nmod, int radius) {
    if (radius < ORC_OZ_RADIUS_FULL || radius > ORC_OZ_RADIUS_CEIL) return -3;
    const int beta = orc_ozaki_beta(k, nmod);
    if (beta < 1) return -2;
    const size_t mk = (size_t)m * k, kn = (size_t)k * n;
    double *mA = dalloc(mk), *rA = dalloc(mk), *mB = dalloc(kn), *rB = dalloc(kn);
    midrad_mat(m, k, A, lda, mA, rA);
    midrad_mat(k, n, B, ldb, mB, rB);
    oz_side SA, SB;
    oz_stats(m, k, mA, rA, 1, (size_t)m, beta, &SA, 0);
    oz_stats(n, k, mB, rB, (size_t)k, 1, beta, &SB, 64);
    int64_t* Ap = (int64_t*)malloc(sizeof(int64_t) * mk); /* row i contiguous */
    int64_t* Bp = (int64_t*)malloc(sizeof(int64_t) * kn); /* col j contiguous */
    /* digit planes (h, l) of |mA|+rA, rA (rows of A) and their rounded-up
     * single digits h + [l > 0]; (H, L), the same bound with a non-positive
     * low digit, of rB, |mB| (cols of B), whose H is the rounded-up digit */
    int8_t *Sh = (int8_t*)malloc(mk), *Sl = (int8_t*)malloc(mk), *Rh = (int8_t*)malloc(mk), *Rl = (int8_t*)malloc(mk);
    int8_t *Sc = (int8_t*)malloc(mk), *Rc = (int8_t*)malloc(mk);
    int8_t *BRh = (int8_t*)malloc(kn), *BRl = (int8_t*)malloc(kn), *BMh = (int8_t*)malloc(kn), *BMl = (int8_t*)malloc(kn);
    double c127, c1, c1dn, c127dn;
    const int old = fegetround();
    {
        volatile double one = 1.0, d = 16129.0, d1 = 127.0;
        fesetround(FE_DOWNWARD);
        c1dn = one / d1;
        c127dn = one / d;
        fesetround(FE_UPWARD);
        c127 = one / d;
        c1 = one / d1;
    }
    for (int i = 0; i < m; ++i)
        for (int c = 0; c < k; ++c) {
            const double am = mA[i + (size_t)m * c], ar = rA[i + (size_t)m * c];
            const double as = fabs(am) + ar;
            const size_t o = (size_t)i * k + c;
            int h, l;
            Ap[o] = (SA.flags[i] & OZ_MIDZERO) ? 0 : rnint_floor(mulpow2(am, beta - SA.e[i]));
            digits2_ru(as, SA.g[i], &h, &l); Sh[o] = (int8_t)h; Sl[o] = (int8_t)l; Sc[o] = (int8_t)(h + (l > 0));
            digits2_ru(ar, SA.h[i], &h, &l); Rh[o] = (int8_t)h; Rl[o] = (int8_t)l; Rc[o] = (int8_t)(h + (l > 0));
        }
    for (int j = 0; j < n; ++j) {
        const int zc = (SB.flags[j] & OZ_MIDZERO) != 0;
        for (int c = 0; c < k; ++c) {
            const double bm = mB[c + (size_t)k * j], br = rB[c + (size_t)k * j];
            const size_t o = (size_t)j * k + c;
            int h, l;
            Bp[o] = zc ? 0 : rnint_floor(mulpow2(bm, beta - SB.e[j]));
            digits2_ru_neglo(br, SB.h[j], &h, &l); BRh[o] = (int8_t)h; BRl[o] = (int8_t)l;
            digits2_ru_neglo(fabs(bm), zc ? 0 : SB.e[j], &h, &l); BMh[o] = (int8_t)h; BMl[o] = (int8_t)l;
        }
    }
    fesetround(old);
    const int nt = orc_get_threads();
#pragma omp parallel num_threads(nt)
    {
        const int o2 = fegetround();
        fesetround(FE_UPWARD);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < n; ++j) {
            for (int i = 0; i < m; ++i) {
                i128 Xh = 0; /* sum of the products' high halves (arithmetic >> 64) */
                u128 Xl = 0; /* sum of their low 64 bits */
                int64_t z1h = 0, z1m = 0, z1l = 0, z2h = 0, z2m = 0, z2l = 0;
                int64_t z1lh = 0, z2lh = 0, z1c = 0, z2c = 0; /* half: lo x H; ceil: H x H */
                const size_t ai = (size_t)i * k, bj = (size_t)j * k;
                for (int c = 0; c < k; ++c) {
                    const i128 pr = (i128)Ap[ai + c] * Bp[bj + c];
                    Xh += pr >> 64;
                    Xl += (uint64_t)pr;
                    z1h += (int64_t)Sh[ai + c] * BRh[bj + c];
                    z1m += (int64_t)Sh[ai + c] * BRl[bj + c] + (int64_t)Sl[ai + c] * BRh[bj + c];
                    z2h += (int64_t)Rh[ai + c] * BMh[bj + c];
                    z2m += (int64_t)Rh[ai + c] * BMl[bj + c] + (int64_t)Rl[ai + c] * BMh[bj + c];
                    z1l += (int64_t)Sl[ai + c] * BRl[bj + c];
                    z2l += (int64_t)Rl[ai + c] * BMl[bj + c];
                    z1lh += (int64_t)Sl[ai + c] * BRh[bj + c];
                    z2lh += (int64_t)Rl[ai + c] * BMh[bj + c];
                    z1c += (int64_t)Sc[ai + c] * BRh[bj + c];
                    z2c += (int64_t)Rc[ai + c] * BMh[bj + c];
                }
                const int fa = SA.flags[i], fb = SB.flags[j];
                const int sc = SA.e[i] + SB.e[j] - 2 * beta;
                Xh += (i128)(Xl >> 64);
                const double lo_mid = conv_x192(Xh, (uint64_t)Xl, sc, -1);
                const double hi_mid = conv_x192(Xh, (uint64_t)Xl, sc, 1);
                double t1 = 0.0, t2 = 0.0;
                /* Per term, with x1 <= (h1 + l1/127) s1 <= c1 s1 (c1 = h1 + [l1 > 0]) and
                 * x2 <= (H + L/127) s2 <= H s2, s = 2^g/127:
                 *   full    sum x1 x2 <= s1 s2 (zh + zm/127 + zl/127^2), zl <= 0;
                 *   folded  s1 s2 (zh + zm/127)   (zl left out);
                 *   half    s1 s2 (zh + zlh/127), zlh = sum l1 H;
                 *   ceil    s1 s2 zc,             zc = sum c1 H.
                 * A quotient with a numerator of either sign is bounded above by the
                 * numerator times RU or RD of the reciprocal. */
                double u1, u2;
                if (radius == ORC_OZ_RADIUS_CEIL) {
                    u1 = (double)z1c;
                    u2 = (double)z2c;
                } else {
                    const int64_t m1 = radius == ORC_OZ_RADIUS_HALF ? z1lh : z1m;
                    const int64_t m2 = radius == ORC_OZ_RADIUS_HALF ? z2lh : z2m;
                    u1 = (double)z1h + (double)m1 * (m1 >= 0 ? c1 : c1dn);
                    u2 = (double)z2h + (double)m2 * (m2 >= 0 ? c1 : c1dn);
                    if (radius == ORC_OZ_RADIUS_FULL) {
                        u1 = u1 + (double)z1l * (z1l >= 0 ? c127 : c127dn);
                        u2 = u2 + (double)z2l * (z2l >= 0 ? c127 : c127dn);
                    }
                }
                if (!((fa & OZ_SUMZERO) || (fb & OZ_RADZERO))) t1 = mulpow2(u1, SA.g[i] + SB.h[j]) * c127;
                if (!((fa & OZ_RADZERO) || (fb & OZ_MIDZERO))) t2 = mulpow2(u2, SA.h[i] + SB.e[j]) * c127;
                const double rt = t1 + t2;
                double E = 0.0;
                if (!(fa & OZ_MIDZERO) && !(fb & OZ_MIDZERO)) {
                    const double ka = mulpow2((double)k, SA.e[i] - beta - 1);
                    const double e1 = mulpow2(SA.n1[i] + ka, SB.e[j] - beta - 1);
                    const double e2 = mulpow2(SB.n1[j], SA.e[i] - beta - 1);
                    E = e1 + e2;
                }
                const double rad = rt + E;
                double* cc = C + 2 * ((size_t)i + (size_t)ldc * j);
                cc[1] = hi_mid + rad;
                cc[0] = -(rad - lo_mid);
            }
        }
        fesetround(o2);
    }
    free(Ap); free(Bp); free(Sh); free(Sl); free(Rh); free(Rl); free(Sc); free(Rc); free(BRh); free(BRl); free(BMh);
    free(BMl);
    oz_free(&SA); oz_free(&SB);
    free(mA); free(rA); free(mB); free(rB);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* pddblas (Ozaki scheme I, FP64 slices, dd accumulation)                   */
/* ------------------------------------------------------------------------ */
static int ceil_log2_i(int k) { int r = 0, t = 1; while (t < k) { t *= 2; ++r; } return r; }
int orc_dd_split_shift(int k) { return (51 + ceil_log2_i(k) + 1) / 2; }          /* :66-68 */
int orc_dd_split_bits(int k) { int b = 52 - orc_dd_split_shift(k); return b < 1 ? 1 : b; } /* :70-74 */
int orc_dd_max_slices(int k) { return 110 / orc_dd_split_bits(k) + 2; }          /* :76-78 */

/* double-double arithmetic in round-to-nearest (Knuth TwoSum, Dekker
 * FastTwoSum); kv::dd's own operator bits are unpinned. */
static void two_sum(double a, double b, double* s, double* e) {
    const double x = a + b;
    const double bb = x - a;
    *e = (a - (x - bb)) + (b - bb);
    *s = x;
}
static void fast_two_sum(double a, double b, double* s, double* e) {
    const double x = a + b;
    *e = b - (x - a);
    *s = x;
}
static void dd_add_d(double* h, double* l, double b) {
    double s, e;
    two_sum(*h, b, &s, &e);
    e += *l;
    fast_two_sum(s, e, h, l);
}

/* split of a dd array; rowwise=1: one sigma per row of an (r x c) array,
 * rowwise=0: one sigma per column.  pddblas.hpp:82-116 / :120-154. */
static int dd_split(const double* A /* AoS r x c, ld = r */, int r, int c, int rowwise, int kdim,
                    double** D /* out slices, each r x c */) {
    const int s = orc_dd_split_shift(kdim);
    const int smax = orc_dd_max_slices(kdim);
    double* T = dalloc(2 * (size_t)r * c);
    memcpy(T, A, sizeof(double) * 2 * (size_t)r * c);
    int cnt = 0;
    const int outer = rowwise ? r : c, inner = rowwise ? c : r;
    for (int p = 0; p < smax; ++p) {
        double* S = dalloc((size_t)r * c);
        memset(S, 0, sizeof(double) * (size_t)r * c);
        int nonzero = 0;
        for (int o = 0; o < outer; ++o) {
            double mu = 0.0;
            for (int q = 0; q < inner; ++q) {
                const size_t idx = rowwise ? (size_t)o + (size_t)r * q : (size_t)q + (size_t)r * o;
                const double a = fabs(T[2 * idx]);
                if (a > mu) mu = a;
            }
            if (mu == 0.0) continue;
            int e;
            frexp(mu, &e);
            const double sigma = ldexp(1.0, e + s);
            for (int q = 0; q < inner; ++q) {
                const size_t idx = rowwise ? (size_t)o + (size_t)r * q : (size_t)q + (size_t)r * o;
                const double qv = (T[2 * idx] + sigma) - sigma;
                if (qv != 0.0) {
                    S[idx] = qv;
                    dd_add_d(&T[2 * idx], &T[2 * idx + 1], -qv);
                    nonzero = 1;
                }
            }
        }
        if (!nonzero) { free(S); break; }
        D[cnt++] = S;
    }
    free(T);
    return cnt;
}

/* ozaki_acc, pddblas.hpp:160-193: slice pairs by increasing p+q, exact
 * products accumulated into C in dd. */
static void dd_ozaki_acc(double** DA, int na, int transA, double** DB, int nb, int m, int n, int k,
                         double* C /* AoS m x n */) {
    const int limit = orc_dd_max_slices(k);
    double* P = dalloc((size_t)m * n);
    double* At = transA ? dalloc((size_t)m * k) : NULL;
    for (int t = 0; t < limit; ++t) {
        for (int p = 0; p <= t; ++p) {
            const int q = t - p;
            if (p >= na || q >= nb) continue;
            const double* Aop = DA[p];
            int lda = m;
            if (transA) { transpose(k, m, DA[p], k, At); Aop = At; lda = m; }
            orc_rgemm(m, n, k, Aop, lda, DB[q], k, P, m, 0, 0);
            for (size_t idx = 0; idx < (size_t)m * n; ++idx)
                if (P[idx] != 0.0) dd_add_d(&C[2 * idx], &C[2 * idx + 1], P[idx]);
        }
    }
    free(P);
    free(At);
}

int orc_pddblas_mulmm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                      double* C, int ldc, int* nslices_a, int* nslices_b) {
    const int smax = orc_dd_max_slices(k);
    double** DA = (double**)malloc(sizeof(double*) * smax);
    double** DB = (double**)malloc(sizeof(double*) * smax);
    double* Ac = dalloc(2 * (size_t)m * k);
    double* Bc = dalloc(2 * (size_t)k * n);
    for (int j = 0; j < k; ++j) memcpy(Ac + 2 * (size_t)m * j, A + 2 * (size_t)lda * j, sizeof(double) * 2 * m);
    for (int j = 0; j < n; ++j) memcpy(Bc + 2 * (size_t)k * j, B + 2 * (size_t)ldb * j, sizeof(double) * 2 * k);
    const int na = dd_split(Ac, m, k, 1, k, DA);   /* split_row, :366 */
    const int nb = dd_split(Bc, k, n, 0, k, DB);   /* split_col, :367 */
    double* Cv = dalloc(2 * (size_t)m * n);
    memset(Cv, 0, sizeof(double) * 2 * (size_t)m * n);
    dd_ozaki_acc(DA, na, 0, DB, nb, m, n, k, Cv); /* :369 */
    for (int j = 0; j < n; ++j) memcpy(C + 2 * (size_t)ldc * j, Cv + 2 * (size_t)m * j, sizeof(double) * 2 * m);
    if (nslices_a) *nslices_a = na;
    if (nslices_b) *nslices_b = nb;
    for (int t = 0; t < na; ++t) free(DA[t]);
    for (int t = 0; t < nb; ++t) free(DB[t]);
    free(DA); free(DB); free(Ac); free(Bc); free(Cv);
    return 0;
}

int orc_pddblas_mulltmm(int k, int n, const double* A, int lda, double* C, int ldc) {
    const int smax = orc_dd_max_slices(k);
    double** DA = (double**)malloc(sizeof(double*) * smax);
    double* Ac = dalloc(2 * (size_t)k * n);
    for (int j = 0; j < n; ++j) memcpy(Ac + 2 * (size_t)k * j, A + 2 * (size_t)lda * j, sizeof(double) * 2 * k);
    const int na = dd_split(Ac, k, n, 0, k, DA);   /* split_col, :388 */
    double* Cv = dalloc(2 * (size_t)n * n);
    memset(Cv, 0, sizeof(double) * 2 * (size_t)n * n);
    dd_ozaki_acc(DA, na, 1, DA, na, n, n, k, Cv); /* :390 */
    for (int i = 0; i < n; ++i)                    /* :399-404 */
        for (int j = i + 1; j < n; ++j) {
            Cv[2 * ((size_t)j + (size_t)n * i)] = Cv[2 * ((size_t)i + (size_t)n * j)];
            Cv[2 * ((size_t)j + (size_t)n * i) + 1] = Cv[2 * ((size_t)i + (size_t)n * j) + 1];
        }
    for (int j = 0; j < n; ++j) memcpy(C + 2 * (size_t)ldc * j, Cv + 2 * (size_t)n * j, sizeof(double) * 2 * n);
    for (int t = 0; t < na; ++t) free(DA[t]);
    free(DA); free(Ac); free(Cv);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Ozaki scheme II double-double product (restatement of the GPU pipeline)  */
/* ------------------------------------------------------------------------ */
#define DL 9 /* 32-bit digits of the exact products (|X| < 2^256) */

int orc_ozaki_dd_beta(int k, int nmod) {
    int mods[64];
    if (nmod > 64 || orc_ozaki_moduli(nmod, mods) != nmod) return -1;
    double lm = 0.0;
    for (int t = 0; t < nmod; ++t) lm += log2((double)mods[t]);
    int lk = 0;
    while ((1ll << lk) < (long long)(k > 1 ? k : 1)) ++lk;
    const int beta = (int)floor((lm - 2.0 - (double)lk - 1e-6) / 2.0);
    return beta > 115 ? 115 : beta;
}

/* nearest integer (ties away) of an exact double, as i128 */
static i128 rn_int128(double y) {
    const double a = fabs(y);
    if (!(a >= 0.5)) return 0;
    int e;
    const double f = frexp(a, &e);
    const uint64_t m = (uint64_t)ldexp(f, 53);
    const int p = e - 53;
    u128 mag = p >= 0 ? ((u128)m << p) : (u128)((m + (1ull << (-p - 1))) >> (-p));
    return y < 0 ? -(i128)mag : (i128)mag;
}

/* e of a dd row: max (|hi| + |lo|) rounded up < 2^e; returns 0 for a zero row */
static int dd_row_exp(const double* X, size_t rs, size_t cs, int r, int len, int* zero) {
    const int old = fegetround();
    fesetround(FE_UPWARD);
    double mu = 0.0;
    for (int c = 0; c < len; ++c) {
        const size_t idx = (size_t)r * rs + (size_t)c * cs;
        volatile double h = X[2 * idx], l = X[2 * idx + 1];
        const double a = fabs(h) + fabs(l);
        if (a > mu) mu = a;
    }
    fesetround(old);
    *zero = mu == 0.0;
    return mu > 0.0 ? fexp(mu) : 0;
}

/* digits (32-bit, little-endian) of |x| for a signed 256-bit carry-save acc */
static int acc_to_mag(int64_t* d, uint32_t* mag) {
    for (int i = 0; i < DL - 1; ++i) {
        const int64_t carry = d[i] >> 32;
        d[i] -= carry * ((int64_t)1 << 32);
        d[i + 1] += carry;
    }
    const int neg = d[DL - 1] < 0;
    if (neg) {
        for (int i = 0; i < DL; ++i) d[i] = -d[i];
        for (int i = 0; i < DL - 1; ++i) {
            const int64_t carry = d[i] >> 32;
            d[i] -= carry * ((int64_t)1 << 32);
            d[i + 1] += carry;
        }
    }
    for (int i = 0; i < DL; ++i) mag[i] = (uint32_t)d[i];
    return neg;
}

static int dg_bit(const uint32_t* x, int b) { return b < 0 ? 0 : (int)((x[b >> 5] >> (b & 31)) & 1u); }
static int dg_any_below(const uint32_t* x, int b) {
    for (int i = 0; i < b; ++i)
        if (dg_bit(x, i)) return 1;
    return 0;
}

/* nearest double (ties to even) of |x| 2^s; *q = lowest kept bit, *up = rounded up */
static double dg_rn(const uint32_t* x, int s, int* qo, int* upo) {
    int top = -1;
    for (int i = DL - 1; i >= 0 && top < 0; --i)
        if (x[i]) top = 32 * i + 31 - __builtin_clz(x[i]);
    *qo = 0;
    *upo = 0;
    if (top < 0) return 0.0;
    int q = top - 52;
    if (q < -1074 - s) q = -1074 - s;
    if (q <= 0) {
        uint64_t m = 0;
        for (int b = top; b >= 0; --b) m = (m << 1) | (uint64_t)dg_bit(x, b);
        return ldexp((double)m, s);
    }
    uint64_t m = 0;
    for (int b = top; b >= q; --b) m = (m << 1) | (uint64_t)dg_bit(x, b);
    const int rb = dg_bit(x, q - 1), st = dg_any_below(x, q - 1);
    const int up = rb && (st || (m & 1));
    m += (uint64_t)up;
    *qo = q;
    *upo = up;
    if (m != 0 && q + s + 64 - __builtin_clzll(m) > 1024) return INFINITY;
    return ldexp((double)m, q + s);
}

/* C = A*B (or A^T A when B == NULL, A k x n) for dd matrices through the
 * exact scaled-integer product the GPU CRT pipeline recovers. */
static int oz_ddgemm_core(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                          double* C, int ldc, int nmod) {
    const int beta = orc_ozaki_dd_beta(k, nmod);
    if (beta < 8) return -2;
    const int tn = (B == NULL);
    if (tn) { B = A; ldb = lda; }
    int *ea = (int*)malloc(sizeof(int) * m), *za = (int*)malloc(sizeof(int) * m);
    int *eb = (int*)malloc(sizeof(int) * n), *zb = (int*)malloc(sizeof(int) * n);
    i128 *Ap = (i128*)malloc(sizeof(i128) * (size_t)m * k), *Bp = (i128*)malloc(sizeof(i128) * (size_t)k * n);
    for (int i = 0; i < m; ++i) {
        if (tn) ea[i] = dd_row_exp(A, (size_t)lda, 1, i, k, &za[i]);
        else ea[i] = dd_row_exp(A, 1, (size_t)lda, i, k, &za[i]);
        for (int c = 0; c < k; ++c) {
            const double* p = tn ? A + 2 * ((size_t)c + (size_t)lda * i) : A + 2 * ((size_t)i + (size_t)lda * c);
            Ap[(size_t)i * k + c] = za[i] ? 0 : rn_int128(ldexp(p[0], beta - ea[i])) + rn_int128(ldexp(p[1], beta - ea[i]));
        }
    }
    for (int j = 0; j < n; ++j) {
        eb[j] = dd_row_exp(B, (size_t)ldb, 1, j, k, &zb[j]);
        for (int c = 0; c < k; ++c) {
            const double* p = B + 2 * ((size_t)c + (size_t)ldb * j);
            Bp[(size_t)j * k + c] = zb[j] ? 0 : rn_int128(ldexp(p[0], beta - eb[j])) + rn_int128(ldexp(p[1], beta - eb[j]));
        }
    }
    const int nt = orc_get_threads();
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < m; ++i) {
            double* cc = C + 2 * ((size_t)i + (size_t)ldc * j);
            if (za[i] || zb[j]) { cc[0] = 0.0; cc[1] = 0.0; continue; }
            int64_t d[DL];
            memset(d, 0, sizeof(d));
            for (int c = 0; c < k; ++c) {
                const i128 a = Ap[(size_t)i * k + c], b = Bp[(size_t)j * k + c];
                if (a == 0 || b == 0) continue;
                const int neg = (a < 0) != (b < 0);
                const u128 ua = a < 0 ? (u128)(-a) : (u128)a, ub = b < 0 ? (u128)(-b) : (u128)b;
                uint32_t da[4], db[4];
                for (int u = 0; u < 4; ++u) { da[u] = (uint32_t)(ua >> (32 * u)); db[u] = (uint32_t)(ub >> (32 * u)); }
                for (int u = 0; u < 4; ++u)
                    for (int v = 0; v < 4; ++v) {
                        const uint64_t pr = (uint64_t)da[u] * db[v];
                        const int64_t lo = (int64_t)(pr & 0xffffffffull), hi = (int64_t)(pr >> 32);
                        if (neg) { d[u + v] -= lo; d[u + v + 1] -= hi; }
                        else { d[u + v] += lo; d[u + v + 1] += hi; }
                    }
            }
            uint32_t mag[DL];
            const int neg = acc_to_mag(d, mag);
            const int sc = ea[i] + eb[j] - 2 * beta;
            int q, up;
            const double hi = dg_rn(mag, sc, &q, &up);
            uint32_t rem[DL];
            memset(rem, 0, sizeof(rem));
            int rneg = 0;
            if (q > 0) {
                for (int b = 0; b < q; ++b)
                    if (dg_bit(mag, b)) rem[b >> 5] |= 1u << (b & 31);
                if (up) { /* rem = 2^q - rem */
                    rneg = 1;
                    int64_t borrow = 0;
                    for (int l = 0; l < DL; ++l) {
                        const int64_t twoq = (q >= 32 * l && q < 32 * l + 32) ? ((int64_t)1 << (q - 32 * l)) : 0;
                        int64_t dd = twoq - (int64_t)rem[l] - borrow;
                        borrow = dd < 0;
                        rem[l] = (uint32_t)(dd + (borrow << 32));
                    }
                }
            }
            int q2, up2;
            double lo = dg_rn(rem, sc, &q2, &up2);
            if (rneg) lo = -lo;
            cc[0] = neg ? -hi : hi;
            cc[1] = neg ? -lo : lo;
        }
    }
    free(ea); free(za); free(eb); free(zb); free(Ap); free(Bp);
    return 0;
}

/* ---- non-finite operands of the Ozaki forms ----------------------------
   A row of A or a column of B holding a +-inf or NaN end is split as zeros;
   every product entry it takes part in is then [-inf, +inf] ([NaN, NaN] if
   a NaN is involved) for the interval form, and (NaN, NaN) for the dd form.
   The exact product of such operands is not a finite interval; this is the
   widest enclosure, or NaN. */
enum { OZ_NONFINITE = 8, OZ_HASNAN = 16 };
static int nf_bits(double x, double y) {
    const int nan = isnan(x) || isnan(y), inf = isinf(x) || isinf(y);
    return ((nan || inf) ? OZ_NONFINITE : 0) | (nan ? OZ_HASNAN : 0);
}

/* flags of `count` vectors of `len` pairs; pair c of vector v at X + 2 (v vs + c cs) */
static int nf_scan(int count, int len, const double* X, size_t vs, size_t cs, int* f) {
    int any = 0;
    for (int v = 0; v < count; ++v) {
        int b = 0;
        for (int c = 0; c < len; ++c) {
            const double* x = X + 2 * ((size_t)v * vs + (size_t)c * cs);
            b |= nf_bits(x[0], x[1]);
        }
        f[v] = b;
        any |= b;
    }
    return any;
}

/* copy of a rows x cols pair matrix (leading dimension ld) with the flagged
   rows (by_row) or columns zeroed, leading dimension rows */
static double* nf_clean(int rows, int cols, const double* X, int ld, const int* f, int by_row) {
    double* Y = dalloc((size_t)2 * (rows > 0 ? rows : 1) * (cols > 0 ? cols : 1));
    for (int c = 0; c < cols; ++c)
        for (int r = 0; r < rows; ++r) {
            const int bad = by_row ? f[r] : f[c];
            const double* x = X + 2 * ((size_t)r + (size_t)ld * c);
            double* y = Y + 2 * ((size_t)r + (size_t)rows * c);
            y[0] = bad ? 0.0 : x[0];
            y[1] = bad ? 0.0 : x[1];
        }
    return Y;
}

int orc_ozaki_igemm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                    double* C, int ldc, int nmod, int radius) {
    int* fa = (int*)calloc(m > 0 ? m : 1, sizeof(int));
    int* fb = (int*)calloc(n > 0 ? n : 1, sizeof(int));
    const int any = nf_scan(m, k, A, 1, (size_t)lda, fa) | nf_scan(n, k, B, (size_t)ldb, 1, fb);
    int rc;
    if (!any) {
        rc = oz_igemm_core(m, n, k, A, lda, B, ldb, C, ldc, nmod, radius);
    } else {
        double* A2 = nf_clean(m, k, A, lda, fa, 1);
        double* B2 = nf_clean(k, n, B, ldb, fb, 0);
        rc = oz_igemm_core(m, n, k, A2, m > 0 ? m : 1, B2, k > 0 ? k : 1, C, ldc, nmod, radius);
        for (int j = 0; j < n && rc == 0; ++j)
            for (int i = 0; i < m; ++i) {
                const int f = fa[i] | fb[j];
                if (!(f & OZ_NONFINITE)) continue;
                double* c = C + 2 * ((size_t)i + (size_t)ldc * j);
                c[0] = (f & OZ_HASNAN) ? NAN : -INFINITY;
                c[1] = (f & OZ_HASNAN) ? NAN : INFINITY;
            }
        free(A2);
        free(B2);
    }
    free(fa);
    free(fb);
    return rc;
}

int orc_ozaki_ddgemm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                     double* C, int ldc, int nmod) {
    const int tn = (B == NULL);
    int* fa = (int*)calloc(m > 0 ? m : 1, sizeof(int));
    int* fb = (int*)calloc(n > 0 ? n : 1, sizeof(int));
    /* tn: C = A^T A with A k x m: operand rows are the columns of A */
    int any = tn ? nf_scan(m, k, A, (size_t)lda, 1, fa) : nf_scan(m, k, A, 1, (size_t)lda, fa);
    any |= tn ? nf_scan(n, k, A, (size_t)lda, 1, fb) : nf_scan(n, k, B, (size_t)ldb, 1, fb);
    int rc;
    if (!any) {
        rc = oz_ddgemm_core(m, n, k, A, lda, B, ldb, C, ldc, nmod);
    } else {
        double* A2 = tn ? nf_clean(k, m, A, lda, fa, 0) : nf_clean(m, k, A, lda, fa, 1);
        double* B2 = tn ? NULL : nf_clean(k, n, B, ldb, fb, 0);
        rc = oz_ddgemm_core(m, n, k, A2, tn ? (k > 0 ? k : 1) : (m > 0 ? m : 1), B2, k > 0 ? k : 1, C, ldc, nmod);
        for (int j = 0; j < n && rc == 0; ++j)
            for (int i = 0; i < m; ++i)
                if ((fa[i] | fb[j]) & OZ_NONFINITE) {
                    double* c = C + 2 * ((size_t)i + (size_t)ldc * j);
                    c[0] = c[1] = NAN;
                }
        free(A2);
        free(B2);
    }
    free(fa);
    free(fb);
    return rc;
}

/* R = B - A X for dd matrices: the residual r = b - A x of pddblas::
 * linearsolve's refinement (vcp/pddblas.hpp:459-466) with the Ozaki-II
 * product above, then a round-to-nearest double-double subtraction per
 * element: TwoSum of the high parts and of the low parts, the low error
 * added, FastTwoSum, the other error added, FastTwoSum. */
int orc_dd_residual(int m, int n, int k, const double* A, int lda, const double* X, int ldx, const double* B,
                    int ldb, double* R, int ldr, int nmod) {
    if (!X) return -1;
    const int rc = orc_ozaki_ddgemm(m, n, k, A, lda, X, ldx, R, ldr, nmod);
    if (rc < 0) return rc;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            const double* b = B + 2 * ((size_t)i + (size_t)ldb * j);
            double* r = R + 2 * ((size_t)i + (size_t)ldr * j);
            double s, e, t, f;
            two_sum(b[0], -r[0], &s, &e);
            two_sum(b[1], -r[1], &t, &f);
            e = e + t;
            fast_two_sum(s, e, &s, &e);
            e = e + f;
            fast_two_sum(s, e, &r[0], &r[1]);
        }
    return rc;
}
