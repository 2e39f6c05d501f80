// Launchers of the libvcphip kernels (host-callable, C++ linkage inside the
// library).  Every launcher returns 0 or a negative VCPHIP_ERR_* code and only
// enqueues work on the given stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vcphip.h"

enum { VCPHIP_DR_II = 0, VCPHIP_DR_IP = 1, VCPHIP_DR_PI = 2, VCPHIP_DR_PP = 3 };

// plain FP64 directed-rounding GEMM family (kern_drgemm.hip)
int vcphip_launch_isgemm(int m, int n, int k, const double* A, long lda, const double* B, long ldb, double* C,
                         long ldc, hipStream_t s);
// the single products in two steps: planes of A (column-major, mpad x kpad;
// mid and rad of an interval A, or A itself) and of B (row-major, kpad x
// npad; mid and rad, or B), zero-padded to mpad, npad = ceil64, kpad =
// ceil16, then the product from them (k > 0) in tiles of `rows` (64 or 32)
// x 64 outputs.  A1 / B1 are unused for a point operand.
int vcphip_launch_planes(int kind, int m, int n, int k, const double* A, long lda, const double* B, long ldb,
                         double* A0, double* A1, double* B0, double* B1, hipStream_t s);
int vcphip_launch_drgemm_planes(int kind, int m, int n, int k, const double* A0, const double* A1, const double* B0,
                                const double* B1, double* C, long ldc, double* C2, int accumulate, int rows,
                                hipStream_t s);
// A^T A (interval or point, A k x n) from the planes of A (made here into
// P0 / P1, each kpad x npad doubles; P1 unused for a point A); k > 0
int vcphip_launch_drsyrk_planes(int interval, int k, int n, const double* A, long lda, double* P0, double* P1,
                                double* C, long ldc, hipStream_t s);
int vcphip_launch_drgemm(int kind, int m, int n, int k, const double* A, long lda, const double* B,
                         long ldb, double* C, long ldc, double* C2, int accumulate, hipStream_t s);
// C = A^T A forms (pidblas::mulltmm / vmulmm), kern_drgemm.hip
int vcphip_launch_drsyrk(int interval, int k, int n, const double* A, long lda, double* C, long ldc,
                         hipStream_t s);

// batched verified residual R_p * IA_p - I (kern_drgemm.hip)
struct VcphipResidProblem {
  const double* R;   // point n x n, column-major
  const double* A;   // interval n x n, AoS {inf, sup}
  double* C;         // interval n x n
  long ldr, lda, ldc;
  int n;
  int row_base;      // first global row of this problem in the row-sum buffer
};
// one output tile of a batched persistent launch: problem, first row and
// column.  A launch's tile array holds its 64 x 64 tiles first, then its
// 32 x 64 ones (rows x columns); each shape runs as a persistent launch of
// its own, 64 x 64 first.
struct VcphipResidTile {
  int problem, row0, col0, pad;
};
#define VCPHIP_RESID_TILE 64
// batched interval products C_p = IA_p * IB_p (the plain form)
struct VcphipIgemmProblem {
  const double* A;   // interval m x k, AoS {inf, sup}, column-major
  const double* B;   // interval k x n
  double* C;         // interval m x n
  long lda, ldb, ldc;
  int m, n, k, pad;
};
// tiles[0, n64) are 64 x 64 tiles, tiles[n64, n64 + n32) 32 x 64 ones; each
// shape's persistent grid takes its part's tiles from a work queue: queue[0]
// and queue[1] (device counters, zeroed by a kernel these launchers enqueue
// first; one pair per stream of launches)
int vcphip_launch_igemm_tiles(int n64, int n32, const VcphipResidTile* tiles, const VcphipIgemmProblem* probs,
                              int grid64, int grid32, unsigned* queue, hipStream_t s);
int vcphip_launch_resid(int n64, int n32, const VcphipResidTile* tiles, const VcphipResidProblem* probs, int nrows,
                        const int* row_problem, double* rowsum, int count, double* normub, int grid64, int grid32,
                        unsigned* queue, hipStream_t s);
// the tile launches alone (the tuner times them)
int vcphip_launch_resid_tiles(int n64, int n32, const VcphipResidTile* tiles, const VcphipResidProblem* probs,
                              int grid64, int grid32, unsigned* queue, hipStream_t s);
// the number of kernels vcphip_launch_resid / vcphip_launch_igemm_tiles
// enqueue for these sizes
inline int vcphip_igemm_tile_kernels(int n64, int n32) { return (n64 > 0 || n32 > 0) + (n64 > 0) + (n32 > 0); }
inline int vcphip_resid_kernels(int n64, int n32, int nrows, int count) {
  return vcphip_igemm_tile_kernels(n64, n32) + (nrows > 0) + (count > 0);
}
// p[0..n) = v by a kernel
int vcphip_fill_f64(double* p, long n, double v, hipStream_t s);

// ---- Ozaki scheme II (kern_ozaki.hip) ------------------------------------
#define VCPHIP_OZ_MAXMOD 32   // moduli (<= 20 for intervals, up to 32 for double-double)
#define VCPHIP_OZ_MAXMOD_IV 20
// radius-digit GEMM outputs (int32 planes of Z) of each scheme, for both
// radius terms together: full 6 (hi x hi, hi x lo + lo x hi, lo x lo per
// term: 8 plane products), folded 4 (no lo x lo: 6), half 4 (A's hi and lo
// against B's rounded-up digit: 4), ceil 2 (rounded-up digits: 2)
#define VCPHIP_OZ_MAXRAD 6
#define VCPHIP_OZ_CHUNK 256   // k-chunk of the upward 1-norm partial sums
#define VCPHIP_DD_CHUNK 64    // k-chunk of the dd row maxima (a maximum: any split gives the same result)
#define VCPHIP_OZ_PADMN 256   // m and n are padded to multiples of this
#define VCPHIP_OZ_PADK 128    // k is padded to a multiple of this
#define VCPHIP_OZ_NLIMB 8     // 32-bit limbs of M (M < 2^256)
#define VCPHIP_OZ_CRTCOL 48   // CRT columns: N weight bytes, 6 quotient bytes, the residue sum (3 x 16)

struct VcphipOzParams {
  int nmod;                    // number of moduli N
  int beta;                    // scaled integers satisfy |A'| <= 2^beta (+1 for dd)
  int mod[VCPHIP_OZ_MAXMOD];   // pairwise coprime moduli (mod[0] = 256)
  // CRT on the matrix cores, with w_t = (M/m_t) ((M/m_t)^-1 mod m_t) < M and
  // F_t = floor(2^48 w_t / M): crt_a[c][g] packs the int8 weights of
  // residues t = 4g .. 4g + 3 (low byte first, 0 for t >= N) of column c:
  // byte c of F_t minus 128 for c < 6, 1 for c = 6 (the residue sum), and
  // byte c - 7 of w_t minus 128 for c = 7 .. N + 6; crt_k[c] =
  // 128 sum_t (byte) - 128^2 N for the byte columns (c != 6)
  alignas(16) uint32_t crt_a[VCPHIP_OZ_CRTCOL][VCPHIP_OZ_MAXMOD / 4];
  int32_t crt_k[VCPHIP_OZ_CRTCOL];
  // residues of dd magnitudes < 2^120 from their 16 bytes: the byte
  // weights 256^i mod m_t, four to a word ([t][q]: i = 4q .. 4q + 3, low
  // byte first); dd_kc[t] = 128 sum_i (256^i mod m_t) - 128^2 16 (the
  // centred-product constant); r_magic[t] = ceil(2^29 / m_t), with which
  // floor(y / m_t) = (8 y r_magic[t]) >> 32 exactly for y < 2^21, m_t <= 256
  alignas(16) uint32_t dd_wb[VCPHIP_OZ_MAXMOD][4];
  int32_t dd_kc[VCPHIP_OZ_MAXMOD];
  uint32_t r_magic[VCPHIP_OZ_MAXMOD];
  // the GEMM epilogue's reduction mod m_t (k <= 8192): ep_bias[t] =
  // ceil(2^11 / m_t) m_t 2^16 and ep_c16[t] = (2^16 mod m_t) - 2^16
  uint32_t ep_bias[VCPHIP_OZ_MAXMOD];
  int32_t ep_c16[VCPHIP_OZ_MAXMOD];
  // residues of |x| <= 2^62 from the eight bytes of x + 2^63 (interval
  // split): the byte weights 256^i mod m_t, four to a word; r_bias =
  // ((floor(m/2) - 2^63) mod m_t) - floor(m/2) (two's complement)
  uint32_t r_wb[2][VCPHIP_OZ_MAXMOD];
  uint32_t r_bias[VCPHIP_OZ_MAXMOD];
  uint32_t Mlimb[VCPHIP_OZ_NLIMB];  // M
  double c1;                   // RU(1/127)
  double c1dn;                 // RD(1/127)
  double c127;                 // RU(1/127^2)
  double c127dn;               // RD(1/127^2)
  // radius-digit GEMM outputs of the interval form: rscheme (a
  // VCPHIP_OZ_RADIUS_* value; 0 for the dd form), rnz outputs; output r sums
  // rnp[r] passes (1 or 2) of A digit plane rpa[r][p] times B digit plane
  // rpb[r][p] (digit planes counted from the first one after the residues)
  int rscheme, rnz;
  int rnp[VCPHIP_OZ_MAXRAD], rpa[VCPHIP_OZ_MAXRAD][2], rpb[VCPHIP_OZ_MAXRAD][2];
  // the digit planes those passes read (bit d: digit plane d of A / of B);
  // the split kernels write only these
  int dmask_a, dmask_b;
};

// radius-digit planes that follow the N residue planes (7-bit upward digits
// per factor of the radius term, against the row / column maximum):
//   A (h, l >= 0):  |mid|+rad hi, lo, rad hi, lo, then the rounded-up single
//                   digits H = h + [l > 0] of |mid|+rad and of rad;
//   B (H, L <= 0):  rad H, L, |mid| H, L (the same bound h + l/127 written
//                   with a non-positive low digit; H alone is the rounded-up
//                   single digit)
#define VCPHIP_OZ_NDIGIT_A 6
#define VCPHIP_OZ_NDIGIT_B 4
// untuned products with k at least this use the folded scheme
#define VCPHIP_OZ_FOLDED_MINK 4096
#define VCPHIP_OZ_MAXK 65536  // 2 k 127^2 < 2^31 for the mixed digit products

// per-row (A) / per-column (B) statistics of a midpoint-radius operand
// NONFINITE: an end of some element of the row / column is +-inf or NaN
// (the row / column is then split as zeros and every product entry it takes
// part in is [-inf, +inf], or [NaN, NaN] with HASNAN)
enum { VCPHIP_OZ_MIDZERO = 1, VCPHIP_OZ_SUMZERO = 2, VCPHIP_OZ_RADZERO = 4, VCPHIP_OZ_NONFINITE = 8,
       VCPHIP_OZ_HASNAN = 16 };
struct VcphipOzSide {
  int* e;       // mid exponent: max|mid| < 2^e
  int* g;       // digit scale of |mid|+rad
  int* h;       // digit scale of rad
  int* flags;   // VCPHIP_OZ_*ZERO: mid / |mid|+rad / rad identically zero
  double* n1;   // upward 1-norm of |mid|
  double* part; // partials [4][nchunk][rows]
};

int vcphip_oz_stats(const double* X, long rs, long cs, int rows, int len, VcphipOzSide side,
                    hipStream_t s);
// planes: N residue planes followed by the radius-digit planes
// (VCPHIP_OZ_NDIGIT_A for A, _B for B), each [rows_pad][kpad] int8 with k
// contiguous
int vcphip_oz_split_rows(const double* A, long lda, int m, int k, int mpad, int kpad,
                         const VcphipOzParams& P, VcphipOzSide side, int8_t* planes, hipStream_t s);
int vcphip_oz_split_cols(const double* B, long ldb, int k, int n, int npad, int kpad,
                         const VcphipOzParams& P, VcphipOzSide side, int8_t* planes, hipStream_t s);
// nmod modular products (u8 residues U) plus nrad radius products (int32 Z)
int vcphip_oz_gemm(int mpad, int npad, int kpad, int nmod, int nrad, const int8_t* Aplanes,
                   const int8_t* Bplanes, const VcphipOzParams& P, uint8_t* U, int32_t* Z, int variant,
                   hipStream_t s);
// the radius outputs in Z follow P.rscheme / P.rnz
int vcphip_oz_recon(int m, int n, int k, int mpad, int npad, const VcphipOzParams& P, VcphipOzSide sa,
                    VcphipOzSide sb, const uint8_t* U, const int32_t* Z, double* C, long ldc,
                    hipStream_t s);

// ---- double-double product, Ozaki scheme II (kern_dd.hip) -------------------
// rows (rs, cs) of a dd operand: e[r] with max |hi|+|lo| < 2^e; flags bit0: zero
int vcphip_dd_stats(const double* X, long rs, long cs, int rows, int len, int* e, int* flags, int* part,
                    hipStream_t s);
int vcphip_dd_split_rows(const double* A, long lda, int m, int k, int mpad, int kpad, const VcphipOzParams& P,
                         const int* e, const int* flags, int8_t* planes, hipStream_t s);
int vcphip_dd_split_cols(const double* B, long ldb, int k, int n, int npad, int kpad, const VcphipOzParams& P,
                         const int* e, const int* flags, int8_t* planes, hipStream_t s);
int vcphip_dd_recon(int m, int n, int mpad, int npad, const VcphipOzParams& P, const int* ea, const int* fa,
                    const int* eb, const int* fb, const uint8_t* U, double* C, long ldc, hipStream_t s);

// autotuner helpers (kern_tune.hip)
// R = B - R for dd matrices (the residual step of the dd linear solve)
int vcphip_dd_sub_from(int m, int n, const double* B, long ldb, double* R, long ldr, hipStream_t s);
// dd solve (pddblas::linearsolve): hi parts, and x += dd(d) with max |d|, max |x.hi|
constexpr int VCPHIP_DD_ADD_BLOCKS = 1024;
int vcphip_dd_hi(int m, int n, const double* X, long ldx, double* D, hipStream_t s);
// D(i, j) = hi(X(rows[i], j)) (rows: a permutation of 0 .. m-1 on the device)
int vcphip_dd_hi_rows(int m, int n, const double* X, long ldx, const int* rows, double* D, hipStream_t s);
int vcphip_dd_add_d(int m, int n, const double* d, double* X, long ldx, double* part, double* out, hipStream_t s);
// dd Cholesky refinement (pddblas::Cholesky): U from dpotrf's upper part, W
// from G, U += F (optionally zeroing U's lower part), the symmetry check
int vcphip_dd_from_upper(int n, const double* D, long ldd, double* U, long ldu, hipStream_t s);
int vcphip_dd_w_from_g(int n, const double* G, double* W, hipStream_t s);
int vcphip_dd_add_to(int n, const double* F, long ldf, double* U, long ldu, int zero_lower, hipStream_t s);
int vcphip_dd_symmetric(int n, const double* A, long lda, int* flag, hipStream_t s);
// dd eigensolvers (pddblas::eigsym / eigsymge): X = dd(D), and the elementwise
// part of one Ogita-Aishima step (lam, row norms, delta, the correction E)
int vcphip_dd_from_d(int m, int n, const double* D, long ldd, double* X, long ldx, hipStream_t s);
int vcphip_oa_elementwise(int n, const double* S, long lds, const double* G, long ldg, double* lam, double* rows,
                          double* delta, double* E, hipStream_t s);
int vcphip_gather_rows(const double* A, long lda, int m, int k, int s, double* out, hipStream_t st);
int vcphip_ratio_max(const double* X, const double* Y, long total, double* part, int nblocks, hipStream_t st);
int vcphip_fill_random(void* p, size_t bytes, uint32_t seed, hipStream_t st);

// verified linear solve: elementwise and scalar steps (kern_lss.hip)
int vcphip_lss_launch_mid(int n, int ncol, const double* b, long ldb, double* mb, long ldm, hipStream_t s);
// approximate inverse: midpoint matrix bordered by the identity to order npad,
// and the copy of its leading n x n block out (kern_lss.hip)
int vcphip_inv_launch_pad(int n, int npad, const double* IA, long lda, double* P, hipStream_t s);
int vcphip_inv_launch_copy(int n, const double* P, long ldp, double* R, long ldr, hipStream_t s);
int vcphip_lss_launch_gemv_rn(int m, int ncol, int k, const double* A, long lda, const double* B, long ldb, double* C,
                              long ldc, hipStream_t s);
int vcphip_lss_launch_sub(int n, int ncol, double* X, long ldx, const double* b, long ldb, hipStream_t s);
int vcphip_lss_launch_abs(int rows, int cols, double* X, long ldx, hipStream_t s);
int vcphip_lss_launch_diag(int n, double* C, long ldc, hipStream_t s);
int vcphip_lss_launch_norm(int rows, int cols, const double* X, long ldx, double* rowbuf, double* out,
                           hipStream_t s);
int vcphip_lss_launch_fill_g(int n, int ncol, const double* norms, double* Tm, long ldt, double* Gout,
                             hipStream_t s);
int vcphip_lss_launch_final(int n, int ncol, const double* Y, const double* W, const double* mx, double* X, long ldx,
                            hipStream_t s);

// batched verified solve: one problem's operands and its workspace
struct VcphipLssProblem {
  const double* A;    // n x n intervals
  const double* b;    // n x nrhs intervals
  const double* R;    // n x n
  double* X;          // n x nrhs intervals (output)
  long lda, ldb, ldr, ldx;
  int n, nrhs;
  double* mb;         // n x nrhs, packed: mid(b)
  double* mx;         // n x nrhs: R mid(b)
  double* xi;         // n x nrhs intervals: IA mx - b
  double* y;          // n x nrhs intervals: |R (IA mx - b)|
  const double* C;    // n x n intervals, ld n: R IA - I (from the batched residual)
  const double* normub;   // g >= ||R IA - I||_inf (from the batched residual)
  double* norms;      // [1] T = ||y||_inf, [2] [3] G
};
// problem p's (row, column) items are [pb[p], pb[p + 1]), pb[count] = total.
// rhs: mid(b), R mid(b), IA mx - b, |R (...)| and ||y||_inf (independent of
// R IA - I); bound: G and the enclosure (needs R IA - I and its norm).
int vcphip_lss_launch_batched_rhs(int count, long items, const long* pb, const VcphipLssProblem* probs,
                                  hipStream_t s);
int vcphip_lss_launch_batched_bound(int count, long items, const long* pb, const VcphipLssProblem* probs,
                                    hipStream_t s);

// Galerkin matrix A(i, j) = sum_k ((w_k u_k) phi(i, k)) phi(j, k) in interval
// arithmetic, symmetric (kern_galerkin.hip)
int vcphip_launch_uhphiphi(int m, int npts, const double* phi, long ldp, const double* w, const double* u, double* A,
                           long lda, hipStream_t s);
