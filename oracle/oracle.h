/*
 * oracle.h -- CPU restatement of VCP's verified dense matrix product.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the vcp_amd
 * package, the libvcphip C ABI) may include, link, load or call this code.
 * It is used by tests/ (as the checker), by __graft_entry__.smoke() (as the
 * checker) and by bench.py's cpu_baseline leg (as the timed CPU baseline).
 *
 * Every function states the reference file:line it restates.  Matrices are
 * column-major like vcp::mats (vcp/mats.hpp:411-417).  Intervals are AoS
 * {inf, sup} pairs (vcp/vcp_fio.hpp:269-272); double-double values are AoS
 * {hi, lo} pairs (vcp/pddblas.hpp:97,213).  Leading dimensions count
 * elements (an interval or dd element is two doubles).
 *
 * Rounding-mode encoding follows vblas rmatmul (vblas/README.md:62-67):
 * 1 = toward +inf, 0 = nearest, -1 = toward -inf.
 */
#ifndef VCP_ORACLE_H
#define VCP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- directed-rounding point GEMM (the rmatmul contract) ---------------- */
/* C(m x n) = A(m x k) * B(k x n); each C element is the k-ascending chain of
 * rounded FMAs starting from 0 (accumulate=0) or from C's value
 * (accumulate=1, the udmatmul form).  vblas/README.md:167-170,
 * vblas/rmatmul_nosimd.hpp:76-115, vblas/udmatmul.hpp:398-441. */
void orc_rgemm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
               double* C, int ldc, int mode, int accumulate);
/* C(n x n) = A^T A of A (k x n): upper triangle by the chain, mirrored.
 * vcp/pdblas.hpp:119-157 (dsyrk 'U','T' + mirror). */
void orc_rsyrk(int k, int n, const double* A, int lda, double* C, int ldc, int mode);

/* ---- kv::midrad restatement ----------------------------------------------- */
/* [m - r, m + r] contains [inf, sup]; point intervals give r = 0.
 * Used by vcp/pidblas.hpp:43-60.  (kv's own rounding choice is unpinned.) */
void orc_midrad(double inf, double sup, double* m, double* r);

/* ---- pidblas restatements (op order of the reference) ------------------- */
int orc_pidblas_mulmm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                      double* C, int ldc);                       /* pidblas.hpp:239-294 */
int orc_pidblas_mul_im_m(int m, int n, int k, const double* IA, int lda, const double* B, int ldb,
                         double* C, int ldc);                    /* pidblas.hpp:90-140  */
int orc_pidblas_mul_m_im(int m, int n, int k, const double* B, int ldb, const double* IA, int lda,
                         double* C, int ldc);                    /* pidblas.hpp:142-196 */
int orc_pidblas_mulltmm(int k, int n, const double* IA, int lda, double* C, int ldc); /* :296-358 */
int orc_pidblas_vmulmm(int k, int n, const double* A, int lda, double* C, int ldc);   /* :198-236 */

/* ---- R*A - I and ||.||_inf (imats.hpp:319-336, mats.hpp:1565-1604) ------- */
/* C = R*IA - I (interval, n x n) and the upper bound of ||C||_inf. */
/* pidblas_fma::mulmm (inf-sup interval FMA chains), vcp/pidblas_fma.hpp:45-72 */
int orc_pidblas_fma_mulmm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                          double* C, int ldc);
int orc_residual(int n, const double* R, int ldr, const double* IA, int lda, double* C, int ldc,
                 double* normub);

/* ---- verified linear solve (imats::linearsolve, vcp/imats.hpp:292-362) --- */
/* After the reference's LAPACK step: R is the approximate inverse of
 * mid(IA).  X (n x nrhs intervals) = mx + [-u, u], u = |y| + |RA - I| G,
 * with mx = R mid(b) (nearest), y = R (IA mx - b), G = ||y|| / (1 - g),
 * g = ||R IA - I||_inf (written to *normub).  Returns 0, 1 when g >= 1 or
 * is NaN (the reference throws, :330-333; X is untouched), -1 for n < 2. */
int orc_lss(int n, int nrhs, const double* IA, int lda, const double* Ib, int ldb, const double* R, int ldr,
            double* X, int ldx, double* normub);

/* ---- Galerkin matrix (Legendre_Bases_Generator::uhphiphi, vcp/ldbase.hpp:955-1013) */
/* kv's interval product [min RD(a b), max RU(a b)] over the endpoint pairs */
void orc_imul(double al, double au, double bl, double bu, double* lo, double* hi);
/* A(i, j) = sum_k ((w_k u_k) phi(i, k)) phi(j, k), interval arithmetic, i <= j
 * from [0, 0] in k order, mirrored */
int orc_uhphiphi(int m, int npts, const double* phi, int ldp, const double* w, const double* u, double* A,
                 int lda);

/* ---- exact evaluation (superaccumulator; no rounding anywhere) ---------- */
/* For entry (i,j) of IA*IB: the exact hull [lo*, hi*] of { sum_k a_k b_k }.
 * Writes sign(lo* - C_lower) and sign(C_upper - hi*) (each >= 0 means the
 * output bound is valid) and the hull bounds rounded outward to double. */
void orc_exact_hull_entry(int k, const double* IArow, int rowstride, const double* IBcol,
                          int colstride, double c_lower, double c_upper,
                          int* ok_lower, int* ok_upper, double* hull_lo, double* hull_hi);
/* Exact point dot product sum_k a_k b_k rounded down / up / nearest. */
void orc_exact_dot(int k, const double* a, int astride, const double* b, int bstride,
                   double* rd, double* ru, double* rn);
/* Exact dd dot product sum_k (ah+al)(bh+bl): nearest dd (hi, lo), and the
 * exact error of a given dd value c = ch + cl, returned as a double
 * |exact - c| rounded up. */
void orc_exact_dd_dot(int k, const double* a, int astride, const double* b, int bstride,
                      double ch, double cl, double* ex_hi, double* ex_lo, double* err_ru);
/* Same sampled checks, run over a list of entries (OpenMP). */
void orc_check_containment(int k, const double* IA, int lda, const double* IB, int ldb,
                           const double* C, int ldc, int nsamp, const int* ii, const int* jj,
                           int* ok, double* hull_lo, double* hull_hi);

/* ---- Ozaki scheme II interval GEMM restatement --------------------------- */
/* Moduli: greedy pairwise-coprime integers from 256 downward. */
int orc_ozaki_moduli(int nmod, int* out);
/* Scale exponent beta for k and nmod so that k*2^(2 beta + 1) < M. */
int orc_ozaki_beta(int k, int nmod);
/* Interval GEMM through the exact scaled-integer product.  Mirrors the GPU
 * pipeline bit for bit (the CRT result equals the integer product).
 * radius: ORC_OZ_RADIUS_FULL (the exact sum of the products of the upward
 * two-digit bounds of the radius-term factors), ORC_OZ_RADIUS_FOLDED (the
 * lo x lo part, never positive with B's non-positive low digits, left out),
 * ORC_OZ_RADIUS_HALF (A's two digits against B's rounded-up digit) or
 * ORC_OZ_RADIUS_CEIL (rounded-up single digits of both factors). */
enum { ORC_OZ_RADIUS_FULL = 1, ORC_OZ_RADIUS_FOLDED = 2, ORC_OZ_RADIUS_HALF = 3, ORC_OZ_RADIUS_CEIL = 4 };
int orc_ozaki_igemm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                    double* C, int ldc, int nmod, int radius);
/* Integer slice work of the same pipeline, exposed for bit-exact checks:
 * the scaled integer operand A' (row scaling) of row-stats exponent e. */
void orc_ozaki_scaled_rows(int m, int k, const double* A, int lda, int beta, int64_t* Ap /* m x k col-major */,
                           int* e_out);
void orc_ozaki_scaled_cols(int k, int n, const double* B, int ldb, int beta, int64_t* Bp /* k x n col-major */,
                           int* f_out);

/* dd product through the Ozaki-II route (the GPU's pddhip pipeline): scale
 * exponent beta for k and nmod (k 2^(2 beta + 2) < M), and C = A*B, or
 * C = A^T A when B is NULL (A is then k x n, m = n). */
int orc_ozaki_dd_beta(int k, int nmod);
int orc_ozaki_ddgemm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                     double* C, int ldc, int nmod);
/* R = B - A X (dd): pddblas::linearsolve's residual (vcp/pddblas.hpp:459-466)
 * with the product above and a TwoSum-based double-double subtraction. */
int orc_dd_residual(int m, int n, int k, const double* A, int lda, const double* X, int ldx, const double* B,
                    int ldb, double* R, int ldr, int nmod);

/* ---- pddblas restatements (vcp/pddblas.hpp) ------------------------------ */
int orc_dd_split_shift(int k);   /* pddblas.hpp:66-68 */
int orc_dd_split_bits(int k);    /* pddblas.hpp:70-74 */
int orc_dd_max_slices(int k);    /* pddblas.hpp:76-78 */
/* C = A*B (dd, AoS {hi,lo}), Ozaki scheme I with FP64 slices:
 * split_row / split_col / ozaki_acc, pddblas.hpp:82-193, 318-376. */
int orc_pddblas_mulmm(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                      double* C, int ldc, int* nslices_a, int* nslices_b);
/* C = A^T A (dd), pddblas.hpp:379-409. */
int orc_pddblas_mulltmm(int k, int n, const double* A, int lda, double* C, int ldc);

/* threads the oracle's OpenMP regions use (0 = library default) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
