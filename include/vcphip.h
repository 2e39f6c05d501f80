/*
 * vcphip.h -- C ABI of libvcphip, the MI355X (gfx950) verified dense matrix
 * product behind the vcp::pidhip / vcp::pddhip policies.
 *
 * Plain pointers and sizes only.  Matrices are column-major as in vcp::mats
 * (vcp/mats.hpp:411-417).  An interval element is the AoS pair {inf, sup}
 * (kv::interval<double>, the order of vcp/vcp_fio.hpp:269-272); a
 * double-double element is the AoS pair {hi, lo} (kv::dd, vcp/pddblas.hpp:97).
 * Leading dimensions count ELEMENTS (an interval or dd element is two
 * doubles).  Every result is an enclosure: lower bounds are rounded toward
 * -inf and upper bounds toward +inf on the GPU; the caller's floating-point
 * environment is never touched (the reference restores round-to-nearest
 * with vcp::hwround_guard, vcp/hwround_guard.hpp:39-48; this library never
 * changes the host rounding mode at all).
 *
 * Entry points without a suffix take HOST pointers and block until the
 * result is in host memory.  The *_dev variants take DEVICE pointers and are
 * asynchronous on the context's stream, except the solvers whose comments
 * say they synchronise (they need a norm, a pivot check or a stopping test
 * on the host).  Device memory comes from the context's cached workspace;
 * the rocSOLVER / rocBLAS calls of the inverses and dd refinements may also
 * allocate their handle's own.  A batch of inverses spreads its groups of
 * equal padded order over the context stream and three more streams of its
 * own (default priority), forked from and joined back to the context stream,
 * so the batch stays ordered on that stream.  Host-pointer copies above 8 MiB are staged
 * through two pinned host buffers of 128 MiB that the context allocates on
 * first use and holds until it is destroyed.  A context is used by one host
 * thread at a time.
 *
 * Every function returns VCPHIP_OK (0) or a negative VCPHIP_ERR_* code;
 * vcphip_strerror() names it.  vcp/error.hpp:42-114 is the exception
 * hierarchy the C++ policy headers map these codes onto.
 */
#ifndef VCPHIP_H
#define VCPHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vcphip_ctx vcphip_ctx;

/* status codes (mapped to vcp::dimension_error, invalid_argument,
 * backend_error, verification_error ... by include/vcp/pidhip.hpp) */
enum {
    VCPHIP_OK = 0,
    VCPHIP_ERR_DIMENSION = -1,    /* vcp::dimension_error  (pdblas.hpp:113-115)  */
    VCPHIP_ERR_ARGUMENT = -2,     /* vcp::invalid_argument                       */
    VCPHIP_ERR_DEVICE = -3,       /* vcp::backend_error("hip", ...)              */
    VCPHIP_ERR_OOM = -4,          /* vcp::backend_error("hip", ...)              */
    VCPHIP_ERR_ROUNDING = -5,     /* directed rounding not in effect             */
    VCPHIP_ERR_VERIFICATION = -6, /* vcp::verification_error (imats.hpp:333-336) */
    VCPHIP_ERR_UNSUPPORTED = -7,  /* size or form outside what the kernels cover */
    VCPHIP_ERR_RANGE = -8,        /* exponent range outside the supported one    */
    VCPHIP_ERR_IO = -9,           /* vcp::io_error (vcp/vcp_portable_fio.hpp)    */
    VCPHIP_ERR_SINGULAR = -10,    /* vcp::lapack_error: zero pivot (pdblas.hpp:
                                     191-203) or not positive definite (:215-217) */
    VCPHIP_ERR_SCRATCH = -11      /* the product needs more scratch than the
                                     no-allocation entry point holds          */
};

/* forms of the interval x interval product */
enum {
    VCPHIP_FORM_PLAIN = 0,  /* FP64 directed-rounding chains (pidblas op order) */
    VCPHIP_FORM_OZAKI = 1,  /* Ozaki scheme II on int8 MFMA + rigorous bounds   */
    VCPHIP_FORM_INFSUP = 2  /* inf-sup chains of interval FMAs: the product of
                               vcp::pidblas_fma::mulmm (vcp/pidblas_fma.hpp:45-72,
                               vcp/ifma.hpp:44-137); tighter, about 3x the work */
};

/* context parameters */
enum {
    VCPHIP_PARAM_OZ_NMOD = 1,      /* moduli of the Ozaki form (0 = tuned/default) */
    VCPHIP_PARAM_OZ_VARIANT = 2,   /* int8 GEMM tile variant 6-18 (0 = tuned/default) */
    VCPHIP_PARAM_STAGE_TIMING = 3, /* 1: record per-stage times with HIP events  */
    VCPHIP_PARAM_DD_NMOD = 4,      /* moduli of the dd product (0 = default)      */
    VCPHIP_PARAM_OZ_RADIUS = 5,    /* radius-digit scheme of the Ozaki form:
                                      VCPHIP_OZ_RADIUS_* (0 = tuned, else full for
                                      k < 4096 and folded from k = 4096) */
    VCPHIP_PARAM_BATCH_TILES = 6,  /* tile shape of the batched launches for every
                                      problem: VCPHIP_BATCH_TILES_* (0 = tuned per
                                      size class, else 64 x 64) */
    VCPHIP_PARAM_PLANE_ROWS = 7    /* output rows per tile of the plain single
                                      products: 64 or 32 (0 = 32 up to 4 64-row
                                      tiles per CU for interval x interval, 3 for
                                      the other forms, else 64); the bits of
                                      every output are the same in both */
};

/* tile shapes of the batched persistent launches (vcphip_igemm_batched*,
 * vcphip_residual_batched_dev, vcphip_lss_batched*), per problem:
 *   64    64 x 64 output tiles only;
 *   EDGE  64 x 64 tiles, and 32 x 64 ones for a last row block of at most 32
 *         rows;
 *   32    32 x 64 tiles only.
 * Each output's FMA chains are the same in every shape, so are its bits; the
 * shapes trade padded work (rows rounded up to 64 or to 32) against the
 * larger tile's efficiency.  Each shape runs as a persistent launch of its
 * own, 64 x 64 first. */
enum {
    VCPHIP_BATCH_TILES_64 = 1,
    VCPHIP_BATCH_TILES_EDGE = 2,
    VCPHIP_BATCH_TILES_32 = 3
};

/* radius-digit schemes of the Ozaki form.  Each factor of the radius term
 * (|mid A| + rad A) rad B + rad A |mid B| (vcp/pidblas.hpp:277-283) is bounded
 * above by 7-bit digits against its row / column maximum, and the sum over k
 * of the products of those bounds is formed on int8 MFMA:
 *   FULL    two digits per factor; hi x hi, hi x lo + lo x hi and lo x lo per
 *           term: the exact sum of the digit-bound products (8 int8 plane
 *           products);
 *   FOLDED  the same with B's low digit non-positive, so that the lo x lo sum
 *           is never positive and is left out (6 plane products);
 *   HALF    A's two digits against B's rounded-up single digit (4);
 *   CEIL    rounded-up single digits of both factors (2).
 * Every scheme gives a rigorous enclosure; they trade radius for int8 work
 * (each cheaper one is wider where small entries sit beside large ones in a
 * row or column, since a digit rounds up against the row / column maximum). */
enum {
    VCPHIP_OZ_RADIUS_FULL = 1,
    VCPHIP_OZ_RADIUS_FOLDED = 2,
    VCPHIP_OZ_RADIUS_HALF = 3,
    VCPHIP_OZ_RADIUS_CEIL = 4
};

/* ---- context ------------------------------------------------------------ */
/* One context per (host thread, device); it owns a stream and a workspace.
 * Replaces the implicit per-thread state of the reference (MXCSR rounding
 * mode + BLAS thread pool), vcp/pidblas.hpp:116-134. */
int vcphip_ctx_create(int device, vcphip_ctx** out);
int vcphip_ctx_destroy(vcphip_ctx* ctx);
/* run *_dev work on a caller stream (a hipStream_t); NULL restores the
 * context's own stream */
int vcphip_ctx_set_stream(vcphip_ctx* ctx, void* stream);
int vcphip_ctx_synchronize(vcphip_ctx* ctx);
int vcphip_ctx_set_param(vcphip_ctx* ctx, int key, long long value);
int vcphip_ctx_get_param(vcphip_ctx* ctx, int key, long long* value);
/* average time (ms) per call of each stage, over the calls made since
 * STAGE_TIMING was last set (events are read at vcphip_ctx_synchronize);
 * stage names are written as a ';'-separated list into names (if given). */
int vcphip_ctx_stage_times(vcphip_ctx* ctx, double* ms, int max, int* count, char* names, int names_len);
const char* vcphip_strerror(int code);
int vcphip_version(void);
/* number of moduli and scale exponent the Ozaki form uses for (k, nmod) */
int vcphip_ozaki_config(int k, int nmod, int* beta, int* moduli, int max_moduli);

/* ---- interval x interval: C = IA * IB ----------------------------------- */
/* Replaces vcp::pidblas::mulmm (vcp/pidblas.hpp:239-294), i.e. the four
 * dgemm_ calls at :266,271,279,281 under kv::hwround and the n^2 fix-ups
 * (forms PLAIN and OZAKI), and vcp::pidblas_fma::mulmm (form INFSUP). */
int vcphip_igemm(vcphip_ctx* ctx, int form, int m, int n, int k,
                 const double* IA, int lda, const double* IB, int ldb, double* C, int ldc);
int vcphip_igemm_dev(vcphip_ctx* ctx, int form, int m, int n, int k,
                     const double* IA, int lda, const double* IB, int ldb, double* C, int ldc);

/* ---- interval x interval, Ozaki form, allocating nothing -----------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same product as vcphip_igemm_dev(ctx,
 * VCPHIP_FORM_OZAKI, ...), bit for bit, with the same moduli count, radius
 * scheme and tile variant (parameters, tuned table, defaults), on the
 * context's stream; its workspace (int8 planes, residues, digit sums,
 * row / column statistics) lives in one device array of
 * VCPHIP_NOALLOC_SCRATCH_BYTES per GPU, reserved when the library's code is
 * loaded on that GPU and held until the process ends.  It calls no
 * hipMalloc, hipFree, memset or event routine.  That covers every shape
 * with m, n, k <= 1024 at up to 20 moduli; a product whose padded workspace
 * does not fit returns VCPHIP_ERR_SCRATCH before anything is enqueued.
 * Calls on one GPU share its array: two that overlap in time on that GPU
 * (different streams or host threads) race and corrupt each other's
 * results; serialise them.  Device pointers only. */
#define VCPHIP_NOALLOC_SCRATCH_BYTES ((size_t)100 << 20)
int vcphip_igemm_ozaki_noalloc_dev(vcphip_ctx* ctx, int m, int n, int k, const double* IA, int lda,
                                   const double* IB, int ldb, double* C, int ldc);

/* ---- interval x point: C = IA * B ---------------------------------------- */
/* Replaces vcp::pidblas::mul_im_m (vcp/pidblas.hpp:90-140). */
int vcphip_igemm_ip(vcphip_ctx* ctx, int m, int n, int k,
                    const double* IA, int lda, const double* B, int ldb, double* C, int ldc);
int vcphip_igemm_ip_dev(vcphip_ctx* ctx, int m, int n, int k,
                        const double* IA, int lda, const double* B, int ldb, double* C, int ldc);

/* ---- point x interval: C = B * IA ---------------------------------------- */
/* Replaces vcp::pidblas::mul_m_im (vcp/pidblas.hpp:142-196); B is m x k. */
int vcphip_igemm_pi(vcphip_ctx* ctx, int m, int n, int k,
                    const double* B, int ldb, const double* IA, int lda, double* C, int ldc);
int vcphip_igemm_pi_dev(vcphip_ctx* ctx, int m, int n, int k,
                        const double* B, int ldb, const double* IA, int lda, double* C, int ldc);

/* ---- interval A^T A: C = IA^T IA (IA is k x n) --------------------------- */
/* Replaces vcp::pidblas::mulltmm (vcp/pidblas.hpp:296-358). */
int vcphip_igemm_tn(vcphip_ctx* ctx, int k, int n, const double* IA, int lda, double* C, int ldc);
int vcphip_igemm_tn_dev(vcphip_ctx* ctx, int k, int n, const double* IA, int lda, double* C, int ldc);

/* ---- verified point A^T A: C = [RD(A^T A), RU(A^T A)] -------------------- */
/* Replaces vcp::pidblas::vmulmm (vcp/pidblas.hpp:198-236). */
int vcphip_vgemm_tn(vcphip_ctx* ctx, int k, int n, const double* A, int lda, double* C, int ldc);
int vcphip_vgemm_tn_dev(vcphip_ctx* ctx, int k, int n, const double* A, int lda, double* C, int ldc);

/* ---- point GEMM rounded up and down in one pass -------------------------- */
/* Cup = RU(A*B), Cdn = RD(A*B) as k-ascending rounded FMA chains.  Replaces
 * rmatmul(m,k,n,A,B,C,+1) and (...,-1) (vblas/rmatmul.hpp:100-127; note the
 * reference names the inner dimension n) and, with accumulate = 1, the
 * accumulating udmatmul (vblas/udmatmul.hpp:738-760); also the pdblas::mulmm
 * dgemm_ under roundup()/rounddown() (vcp/pdblas.hpp:55-117). */
int vcphip_gemm_rurd(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda,
                     const double* B, int ldb, double* Cup, double* Cdn, int ldc, int accumulate);
int vcphip_gemm_rurd_dev(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda,
                         const double* B, int ldb, double* Cup, double* Cdn, int ldc, int accumulate);

/* ---- double-double GEMM: C = A * B --------------------------------------- */
/* Replaces vcp::pddblas::mulmm (vcp/pddblas.hpp:318-376), the Ozaki-scheme
 * kv::dd product (split_row/split_col/ozaki_acc, :82-193), computed here by
 * an Ozaki scheme II on int8 MFMA; the result is the nearest dd of a
 * product accurate to ~2^-106 relative to the row/column scales. */
int vcphip_ddgemm(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda,
                  const double* B, int ldb, double* C, int ldc);
int vcphip_ddgemm_dev(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda,
                      const double* B, int ldb, double* C, int ldc);
/* C = A^T A for dd A (k x n), exactly symmetric.  Replaces
 * vcp::pddblas::mulltmm (vcp/pddblas.hpp:379-409). */
int vcphip_ddgemm_tn(vcphip_ctx* ctx, int k, int n, const double* A, int lda, double* C, int ldc);

/* C = A^T B for dd matrices (A k x m, B k x n, C m x n): the X^T (A X) and
 * X^T X products of pddblas's Ogita-Aishima refinement (oa_step,
 * vcp/pddblas.hpp:221-314), whose ozaki_acc calls take a transposed left
 * operand; the same Ozaki-II dd product as vcphip_ddgemm. */
int vcphip_ddgemm_ta(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                     double* C, int ldc);
int vcphip_ddgemm_ta_dev(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                         double* C, int ldc);
/* R = B - A*X for dd matrices (A m x k, X k x n, B and R m x n): the
 * residual r = b - A*x of pddblas::linearsolve's refinement
 * (vcp/pddblas.hpp:459-466), whose products ozaki_acc forms; here A*X is the
 * Ozaki-II dd product of vcphip_ddgemm, then a double-double subtraction
 * (TwoSum-based, round-to-nearest) per element.  pddhip::linearsolve and
 * pddhip::inv (dd inverse: A*X = I) run the reference's refinement loop
 * with it. */
int vcphip_dd_residual(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda, const double* X, int ldx,
                       const double* B, int ldb, double* R, int ldr);
int vcphip_dd_residual_dev(vcphip_ctx* ctx, int m, int n, int k, const double* A, int lda, const double* X, int ldx,
                           const double* B, int ldb, double* R, int ldr);
/* x = A \ b for dd matrices, every step on the device: pddblas::linearsolve
 * (vcp/pddblas.hpp:412-474).  LU with partial pivoting of the high parts of
 * A (rocSOLVER dgetrf for LAPACK's), then up to itmax rounds of d = LU \ hi(r)
 * (dgetrs), x += dd(d), stopping when max|d| is 0, at most 0.25e-32 max|x.hi|
 * or at least half the previous round's, else r = b - A x as
 * vcphip_dd_residual forms it (the first round takes r = b, x = 0).  A is
 * n x n, b and x n x nrhs, dd pairs {hi, lo}; *iters (may be NULL) receives
 * the rounds done.  The reference's itmax is 10; pddblas::inv is this with
 * b = I.  Returns VCPHIP_ERR_SINGULAR when a pivot is exactly zero (the
 * reference's lapack_error("dgetrf")).  Both forms synchronise the stream. */
int vcphip_dd_solve(vcphip_ctx* ctx, int n, int nrhs, const double* A, int lda, const double* b, int ldb, double* x,
                    int ldx, int itmax, int* iters);
int vcphip_dd_solve_dev(vcphip_ctx* ctx, int n, int nrhs, const double* A, int lda, const double* b, int ldb,
                        double* x, int ldx, int itmax, int* iters);
/* Upper triangular U with U^T U = A to dd accuracy, every step on the
 * device: pddblas::Cholesky (vcp/pddblas.hpp:488-563).  The reference's
 * symmetry check (both parts of A(i, j) and A(j, i) equal, else
 * VCPHIP_ERR_ARGUMENT for its domain_error); dpotrf of double(A) (rocSOLVER;
 * VCPHIP_ERR_SINGULAR when A is not positive definite, its lapack_error),
 * the strictly lower part zeroed; then itep rounds (the reference's 2) of
 * E = A - U^T U (dd product), G = Ud^-T E Ud^-1 (rocBLAS dtrsm, Ud = U's
 * high parts), W = triu(G, 1) + diag(G) / 2 and U = U + W U (dd product,
 * double-double addition).  A and U are n x n dd pairs {hi, lo}. */
int vcphip_dd_cholesky(vcphip_ctx* ctx, int n, const double* A, int lda, double* U, int ldu, int itep);
int vcphip_dd_cholesky_dev(vcphip_ctx* ctx, int n, const double* A, int lda, double* U, int ldu, int itep);
/* Eigenpairs of a symmetric dd matrix to dd accuracy, every step on the
 * device: pddblas::eigsym (vcp/pddblas.hpp:566-606) and, with B symmetric
 * positive definite, the generalized problem A x = lam B x of
 * pddblas::eigsymge (:611-659).  The reference's symmetry checks
 * (VCPHIP_ERR_ARGUMENT), then dsyev / dsygv of the high parts (rocSOLVER
 * dsyevd / dsygvd; VCPHIP_ERR_SINGULAR when they fail, the reference's
 * lapack_error), then max(itep, 2) Ogita-Aishima steps (oa_step, :219-314):
 * A X, X^T (A X), X^T X or X^T (B X) as dd products, lam_i = S_ii / G_ii, the
 * separation bound delta, the correction E and X += X E in double-double
 * arithmetic.  lam receives n dd pairs (ascending order of the start), X the
 * n x n eigenvectors (columns).  The reference's itep is 4 for eigsym and 1
 * for eigsymge. */
int vcphip_dd_eigsym(vcphip_ctx* ctx, int n, const double* A, int lda, double* lam, double* X, int ldx, int itep);
int vcphip_dd_eigsym_dev(vcphip_ctx* ctx, int n, const double* A, int lda, double* lam, double* X, int ldx, int itep);
int vcphip_dd_eigsymge(vcphip_ctx* ctx, int n, const double* A, int lda, const double* B, int ldb, double* lam,
                       double* X, int ldx, int itep);
int vcphip_dd_eigsymge_dev(vcphip_ctx* ctx, int n, const double* A, int lda, const double* B, int ldb, double* lam,
                           double* X, int ldx, int itep);

/* ---- batched verified residual: C_i = R_i * IA_i - I ---------------------- */
/* For each problem i (square, size n[i]): C_i = R_i*IA_i - I as intervals and
 * normub[i] >= ||C_i||_inf.  Replaces the R*A - I segment of
 * imats::linearsolve (vcp/imats.hpp:318-336: mul_m_im, the diagonal
 * subtraction and mats::norminf, vcp/mats.hpp:1583-1600).  Pointer arrays
 * are host arrays of device pointers for the _dev form. */
int vcphip_residual_batched(vcphip_ctx* ctx, int count, const int* n,
                            const double* const* R, const int* ldr,
                            const double* const* IA, const int* lda,
                            double* const* C, const int* ldc, double* normub);
int vcphip_residual_batched_dev(vcphip_ctx* ctx, int count, const int* n,
                                const double* const* R, const int* ldr,
                                const double* const* IA, const int* lda,
                                double* const* C, const int* ldc, double* normub_dev);

/* ---- batched interval products: C_p = IA_p * IB_p ------------------------ */
/* Many interval products of uneven (m_p, n_p, k_p) (the Legendre-basis
 * Galerkin blocks of a verification run) in ONE persistent launch: the plain
 * form of pidblas::mulmm (vcp/pidblas.hpp:239-294) per problem, so each C_p
 * is bit-identical to vcphip_igemm(VCPHIP_FORM_PLAIN).  Workgroups stride
 * over a tile list sorted by problem size, largest first.  The _dev form
 * captures the launch into a hipGraph on first use and replays it while the
 * problem set (pointers, sizes, leading dimensions, stream) is unchanged.
 * Pointer arrays are host arrays (of device pointers for the _dev form);
 * all operands AoS {inf, sup}, column-major.  The _dev form is asynchronous;
 * the host form synchronises. */
int vcphip_igemm_batched(vcphip_ctx* ctx, int count, const int* m, const int* n, const int* k,
                         const double* const* IA, const int* lda, const double* const* IB, const int* ldb,
                         double* const* C, const int* ldc);
int vcphip_igemm_batched_dev(vcphip_ctx* ctx, int count, const int* m, const int* n, const int* k,
                             const double* const* IA, const int* lda, const double* const* IB, const int* ldb,
                             double* const* C, const int* ldc);

/* ---- approximate inverse of the midpoint matrix ---------------------------- */
/* R = inv(mid(IA)), the step of imats::linearsolve before the verified ones
 * (vcp/imats.hpp:298-301: imid, then pdblas::inv, vcp/pdblas.hpp:180-204):
 * each entry's midpoint as kv's mid() (round to nearest), then LU with
 * partial pivoting and the inverse from it (rocSOLVER dgetrf, then dgetrs
 * against the identity: the device counterpart of the reference's LAPACK
 * dgetrf + dgetri, which on the device is the faster route to an inverse of
 * the same accuracy).  R is only approximate; nothing verified depends on
 * its accuracy.  IA is n x n
 * intervals (AoS {inf, sup}), R n x n doubles.  Returns VCPHIP_ERR_SINGULAR
 * when a pivot is exactly zero (LAPACK's info > 0; R is then unspecified).
 * Both forms synchronise the context's stream before returning. */
int vcphip_inv_mid(vcphip_ctx* ctx, int n, const double* IA, int lda, double* R, int ldr);
int vcphip_inv_mid_dev(vcphip_ctx* ctx, int n, const double* IA, int lda, double* R, int ldr);

/* ---- verified linear solve: X contains A^-1 b ----------------------------- */
/* Replaces imats::linearsolve (vcp/imats.hpp:292-362).  R is a point
 * approximate inverse of mid(IA); R = NULL has the library form it on the
 * device as vcphip_inv_mid does (the reference's step at :298-301), else it
 * is the caller's (ldr is then ignored when R is NULL).  With mx = R*mid(Ib)
 * (nearest rounding), y = R*(IA*mx - Ib) and g >= ||R*IA - I||_inf, it
 * writes X = mx + [-u, u] with u = |y| + |R*IA - I| * (||y||_inf / (1 - g)),
 * every step in the reference's order and rounding, so X contains the
 * solution of A x = b for every A in IA and b in Ib.  IA is n x n, Ib and X
 * are n x nrhs intervals (AoS {inf, sup}), R is n x n.  *normub (may be NULL)
 * receives g.  Returns VCPHIP_ERR_VERIFICATION when g >= 1 or is NaN
 * (imats.hpp:330-333; X is then unspecified), VCPHIP_ERR_SINGULAR when the
 * library's own R met a zero pivot (X unspecified) and VCPHIP_ERR_DIMENSION
 * for n < 2 (1 x 1 systems are the reference's scalar case, left to the host).
 * Both forms synchronise the context's stream before returning. */
int vcphip_lss(vcphip_ctx* ctx, int n, int nrhs, const double* IA, int lda, const double* Ib, int ldb,
               const double* R, int ldr, double* X, int ldx, double* normub);
int vcphip_lss_dev(vcphip_ctx* ctx, int n, int nrhs, const double* IA, int lda, const double* Ib, int ldb,
                   const double* R, int ldr, double* X, int ldx, double* normub);
/* Many systems of uneven size (the Galerkin blocks of a verification run):
 * per problem p the vcphip_lss result, bit-identical to it.  R_p*IA_p - I
 * and its norm run in the batched residual's persistent launch (its graph
 * is reused while the problem set is unchanged); every other step is one
 * launch over all problems.  Pointer arrays are host arrays (of device
 * pointers for the _dev form); normub is a host array of count doubles.
 * R (and ldr) may be NULL, or R[p] NULL for some p: those R_p = inv(mid(IA_p))
 * are formed on the device first, one rocSOLVER batched getrf + getrs per
 * group of equal order (orders bordered by the identity to a multiple of 64).  Returns VCPHIP_ERR_SINGULAR when any such R_p met a
 * zero pivot, else VCPHIP_ERR_VERIFICATION when any normub[p] >= 1 (the
 * other problems' X are valid; a failed problem's X is NaN). */
int vcphip_lss_batched(vcphip_ctx* ctx, int count, const int* n, const int* nrhs,
                       const double* const* IA, const int* lda, const double* const* Ib, const int* ldb,
                       const double* const* R, const int* ldr, double* const* X, const int* ldx,
                       double* normub);
int vcphip_lss_batched_dev(vcphip_ctx* ctx, int count, const int* n, const int* nrhs,
                           const double* const* IA, const int* lda, const double* const* Ib, const int* ldb,
                           const double* const* R, const int* ldr, double* const* X, const int* ldx,
                           double* normub);

/* ---- autotuner ------------------------------------------------------------
 * No counterpart in the reference (its BLAS path has nothing to tune).  The
 * choices live in the context and are keyed by product kind and size class
 * (ceil(log2) of m, n and k); a product of a tuned class uses them unless
 * VCPHIP_PARAM_OZ_VARIANT / VCPHIP_PARAM_OZ_NMOD are set.  No choice changes
 * the containment guarantee: tile variants give bit-identical results, and
 * every moduli count yields a rigorous enclosure. */

/* Times every int8 MFMA tile variant of the interval Ozaki product's GEMM
 * stage at (m, n, k) and records the fastest; *variant and *ms (per launch)
 * receive the choice. */
int vcphip_tune_igemm_tile(vcphip_ctx* ctx, int m, int n, int k, int* variant, double* ms);
/* The same for the double-double product (vcphip_ddgemm). */
int vcphip_tune_ddgemm_tile(vcphip_ctx* ctx, int m, int n, int k, int* variant, double* ms);
/* What an interval Ozaki product of shape (m, n, k) in this context uses
 * (parameters, then the tuned table, then the defaults): moduli count,
 * radius-digit scheme, int8 GEMM tile variant and the number of int8 plane
 * products of its GEMM stage (each 2 m n k int8 operations, padded).  Any
 * output may be NULL. */
int vcphip_ozaki_plan(vcphip_ctx* ctx, int m, int n, int k, int* nmod, int* radius, int* variant,
                      int* products);
/* Slice count and radius-digit scheme of the interval Ozaki form for these
 * operands (device pointers): the cheapest pair (moduli N in [12, 20] with
 * VCPHIP_OZ_RADIUS_CEIL at N + 2 int8 plane products, _HALF at N + 4,
 * _FOLDED at N + 6 or _FULL at N + 8) whose radii, on an evenly spaced sample
 * of sample_rows rows of IA (0: max(1024, m/8)) against all of IB, are at most
 * max_ratio times the plain form's; at equal cost the smaller ratio wins; 20
 * moduli with the full scheme if none is.
 * *nmod, *radius and *ratio (the sampled maximum radius ratio) receive the
 * choice (any of them may be NULL). */
int vcphip_tune_igemm_moduli_dev(vcphip_ctx* ctx, int m, int n, int k, const double* IA, int lda,
                                 const double* IB, int ldb, double max_ratio, int sample_rows,
                                 int* nmod, int* radius, double* ratio);
/* Tunes the batched residual for this problem set (arguments as
 * vcphip_residual_batched_dev, which it also runs): first the tile shape of
 * each size class of the set (VCPHIP_BATCH_TILES_*; classes are ceil(log2 n),
 * taken in order of their work, each tried with the three shapes while the
 * others keep theirs, the whole set's tile launches timed), then 1-3
 * persistent workgroups per CU.  Keeps the fastest; the next call captures
 * its graph with them.  *per_cu, *ms (the set's tile launches) and *padded
 * (the fraction of the tiles' multiply-adds that is padding, under the
 * choice) receive the result; any may be NULL. */
int vcphip_tune_residual_dev(vcphip_ctx* ctx, int count, const int* n, const double* const* R,
                             const int* ldr, const double* const* IA, const int* lda,
                             double* const* C, const int* ldc, double* normub_dev, int* per_cu,
                             double* ms, double* padded);
/* The same shape choice per size class (ceil(log2) of a problem's rows) for
 * the batched interval products (arguments as vcphip_igemm_batched_dev, which
 * it also runs).  *ms: the set's launch with the choice; *padded as above. */
int vcphip_tune_igemm_batched_dev(vcphip_ctx* ctx, int count, const int* m, const int* n, const int* k,
                                  const double* const* IA, const int* lda, const double* const* IB,
                                  const int* ldb, double* const* C, const int* ldc, double* ms,
                                  double* padded);
/* Writes / reads the context's tuning table (text: a "vcphip-tune 2" line,
 * then "igemm|ddgemm lm ln lk variant nmod radius" and "residual per_cu"
 * lines; version 1 tables, without the radius field, still load). */
int vcphip_tune_save(vcphip_ctx* ctx, const char* path);
int vcphip_tune_load(vcphip_ctx* ctx, const char* path);

/* ---- Galerkin matrix producer ---------------------------------------------- */
/* A(i, j) = sum over points k of ((w_k * u_k) * phi(i, k)) * phi(j, k) in
 * interval arithmetic (each product and sum separately rounded outward, k
 * ascending from [0, 0]), for i <= j, mirrored to A(j, i): the matrix
 * Legendre_Bases_Generator::uhphiphi builds (vcp/ldbase.hpp:955-1013) from
 * its basis values phi_point (m x npts), quadrature weights weight_point and
 * the coefficient function values uh_p_phi_point (1 x npts each).  phi and A
 * are column-major interval matrices, w and u interval vectors. */
int vcphip_uhphiphi(vcphip_ctx* ctx, int m, int npts, const double* phi, int ldp, const double* w, const double* u,
                    double* A, int lda);
int vcphip_uhphiphi_dev(vcphip_ctx* ctx, int m, int npts, const double* phi, int ldp, const double* w,
                        const double* u, double* A, int lda);

/* ---- matrix files ----------------------------------------------------------
 * The portable ".vcpmat" format (vcp/vcp_portable_fio.hpp: magic "VCPMAT2",
 * version 2, column-major, little-endian, CRC-64 over header and payload)
 * and the native ".matrix_*" format (vcp/vcp_fio.hpp: type tag, int32 sizes,
 * row-major elements).  path is the full file name (the reference appends
 * ".vcpmat" / ".matrix_d" etc. to the name it is given).  Element types and
 * their bytes: */
enum {
    VCPHIP_MAT_INT32 = 1,       /* int32                                  4 B */
    VCPHIP_MAT_DOUBLE = 2,      /* binary64                               8 B */
    VCPHIP_MAT_INTERVAL = 3,    /* interval<double> {lower, upper}       16 B */
    VCPHIP_MAT_DD = 4,          /* kv::dd {a1, a2}                       16 B */
    VCPHIP_MAT_INTERVAL_DD = 5  /* interval<kv::dd> {l.a1, l.a2, u.a1, u.a2} 32 B */
};
/* Type, size and format (*portable = 1 for .vcpmat; may be NULL) of a file.
 * Either format is recognised by its first bytes; a bad magic, version,
 * layout, size, payload length, CRC or trailing data gives VCPHIP_ERR_IO. */
int vcphip_mat_info(const char* path, int* type, int* rows, int* cols, int* portable);
/* Reads the file into a column-major host array (leading dimension ld, in
 * elements).  The file's type must be `type` (VCPHIP_ERR_IO otherwise, the
 * reference's data type mismatch) and its size rows x cols. */
int vcphip_mat_read(const char* path, int type, int rows, int cols, void* data, int ld);
/* Writes a column-major host array in the portable (portable != 0) or the
 * native format, through path + ".tmp" and a rename, as the reference does. */
int vcphip_mat_write(const char* path, int type, int rows, int cols, const void* data, int ld, int portable);
/* CRC-64 of the portable format (polynomial 0x42F0E1EBA9EA3693, MSB first,
 * initial value 0, no final xor). */
int vcphip_crc64(const void* data, size_t bytes, uint64_t* crc);

#ifdef __cplusplus
}
#endif
#endif /* VCPHIP_H */
