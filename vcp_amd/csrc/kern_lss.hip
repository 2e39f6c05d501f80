// Verified linear solve around R*A - I (imats::linearsolve,
// vcp/imats.hpp:292-362): the elementwise and scalar steps between the
// products.  The products themselves are the k_drgemm forms (IP for A*mx,
// PI for R*x and R*A, II for |RA - I| * T) and the nearest-rounded point
// product R*mid(b) below; the host drives the sequence
// (vcphip_lss_dev in vcphip_host.hip).
//
// Kernels that need directed rounding run in round-toward-+inf (see
// vcphip_device.h) and form RD(x) = -RU(-x).  k_lss_mid, k_lss_gemv_rn and
// k_lss_fill_g run in the default round-to-nearest mode: the midpoint and
// R*mid(b) are nearest-rounded in the reference (no rounding mode is set
// around them), and the scalar G = T / (1 - ||RA - I||) gets its directed
// roundings from exact error terms of nearest-rounded operations.
#include "vcphip_device.h"
#include "vcphip_kernels.h"

namespace {

constexpr int NT = 256;

// kv's mid(): the nearest-rounded average, (lo + hi) / 2, or lo/2 + hi/2
// where lo + hi overflows; point intervals give lo.  mb has leading
// dimension ldm.
__global__ __launch_bounds__(NT) void k_lss_mid(int n, int ncol, const double* __restrict__ b, long ldb,
                                                double* __restrict__ mb, long ldm) {
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  if (idx >= (long)n * ncol) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  const double2 v = *reinterpret_cast<const double2*>(b + 2 * (i + ldb * j));
  double m;
  if (v.x == v.y) {
    m = v.x;
  } else {
    const double s = v.x + v.y;
    m = __builtin_isinf(s) ? 0.5 * v.x + 0.5 * v.y : s * 0.5;
  }
  mb[i + ldm * j] = m;
}

// The midpoint matrix M of IA (n x n, as k_lss_mid) bordered by the identity
// to order npad: P = [[M, 0], [0, I]], whose inverse is [[inv(M), 0], [0, I]].
// In LU with partial pivoting the border rows keep zeros in M's columns
// (their multipliers are 0), so no pivot is taken from them and a singular M
// still meets a zero pivot.  n = 0 writes the identity (IA is not read).
__global__ __launch_bounds__(NT) void k_inv_pad(int n, int npad, const double* __restrict__ IA, long lda,
                                                double* __restrict__ P) {
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  if (idx >= (long)npad * npad) return;
  const int i = (int)(idx % npad), j = (int)(idx / npad);
  double m;
  if (i < n && j < n) {
    const double2 v = *reinterpret_cast<const double2*>(IA + 2 * (i + lda * j));
    if (v.x == v.y) {
      m = v.x;
    } else {
      const double s = v.x + v.y;
      m = __builtin_isinf(s) ? 0.5 * v.x + 0.5 * v.y : s * 0.5;
    }
  } else {
    m = (i == j) ? 1.0 : 0.0;
  }
  P[idx] = m;
}

// R(0:n, 0:n) = P(0:n, 0:n)
__global__ __launch_bounds__(NT) void k_copy_block(int n, const double* __restrict__ P, long ldp,
                                                   double* __restrict__ R, long ldr) {
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  R[i + ldr * j] = P[i + ldp * j];
}

// C = A * B in round-to-nearest, each element the k-ascending FMA chain
// from 0 (the rmatmul contract with nearest rounding, vblas/README.md:
// 167-170).  One thread per output row and column; lanes of a wave take
// consecutive rows, so the column-major A is read coalesced.
__global__ __launch_bounds__(NT) void k_lss_gemv_rn(int m, int k, const double* __restrict__ A, long lda,
                                                    const double* __restrict__ B, long ldb,
                                                    double* __restrict__ C, long ldc) {
  const int i = blockIdx.x * NT + threadIdx.x, j = blockIdx.y;
  if (i >= m) return;
  const double* a = A + i;
  const double* bj = B + ldb * (long)j;
  double acc = 0.0;
  // a thread per row is few waves per SIMD: 32 k of loads in flight per
  // thread, then the chain in k order
  constexpr int U = 32;
  int kk = 0;
  for (; kk + U <= k; kk += U) {
    double av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      av[u] = a[lda * (kk + u)];
      bv[u] = bj[kk + u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc = __builtin_fma(av[u], bv[u], acc);
  }
  for (; kk < k; ++kk) acc = __builtin_fma(a[lda * kk], bj[kk], acc);
  C[i + ldc * (long)j] = acc;
}

// X = X - b (interval subtraction, imats.hpp:308): [RD(xl - bu), RU(xu - bl)]
__global__ __launch_bounds__(NT) void k_lss_sub(int n, int ncol, double* __restrict__ X, long ldx,
                                                const double* __restrict__ b, long ldb) {
  vcphip_mode_round_up();
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  if (idx >= (long)n * ncol) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  double2* x = reinterpret_cast<double2*>(X + 2 * (i + ldx * j));
  const double2 bv = *reinterpret_cast<const double2*>(b + 2 * (i + ldb * j));
  const double2 xv = *x;
  *x = make_double2(-(bv.y - xv.x), xv.y - bv.x);
}

// kv's abs() of an interval, in place (imats.hpp:315, :338): exact
__global__ __launch_bounds__(NT) void k_lss_abs(int rows, int cols, double* __restrict__ X, long ldx) {
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  if (idx >= (long)rows * cols) return;
  const int i = (int)(idx % rows), j = (int)(idx / rows);
  double2* x = reinterpret_cast<double2*>(X + 2 * (i + ldx * j));
  const double2 v = *x;
  double lo, hi;
  if (v.x >= 0.0) {
    lo = v.x;
    hi = v.y;
  } else if (v.y <= 0.0) {
    lo = -v.y;
    hi = -v.x;
  } else {
    lo = 0.0;
    hi = -v.x > v.y ? -v.x : v.y;
  }
  *x = make_double2(lo, hi);
}

// C = C - I on the diagonal (imats.hpp:324-326): [RD(lo - 1), RU(hi - 1)]
__global__ __launch_bounds__(NT) void k_lss_diag(int n, double* __restrict__ C, long ldc) {
  vcphip_mode_round_up();
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  double2* c = reinterpret_cast<double2*>(C + 2 * (i + ldc * (long)i));
  const double2 v = *c;
  *c = make_double2(-(1.0 - v.x), v.y - 1.0);
}

// upper end of the row sums of |X| (mats::norminf, vcp/mats.hpp:1587-1592):
// s = mag(x_i0), then s = RU(s + mag(x_ij)) for j = 1, 2, ...
__global__ __launch_bounds__(NT) void k_lss_rowsum(int rows, int cols, const double* __restrict__ X, long ldx,
                                                   double* __restrict__ out) {
  vcphip_mode_round_up();
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  auto add = [&](int j, double2 v) {
    const double a = __builtin_fabs(v.x), b = __builtin_fabs(v.y);
    const double mag = a > b ? a : b;
    s = (j == 0) ? mag : s + mag;
  };
  // upward, in column order; 32 columns of loads in flight per thread
  constexpr int U = 32;
  int j = 0;
  for (; j + U <= cols; j += U) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const double2*>(X + 2 * (i + ldx * (long)(j + u)));
#pragma unroll
    for (int u = 0; u < U; ++u) add(j + u, v[u]);
  }
  for (; j < cols; ++j) add(j, *reinterpret_cast<const double2*>(X + 2 * (i + ldx * (long)j)));
  out[i] = s;
}

// maximum of in[0..rows) into *out with the reference's rule (mats.hpp:
// 1593-1598): best = in[0], then best = in[i] where in[i] > best.  So the
// result is NaN exactly when in[0] is, and otherwise the largest non-NaN.
__global__ __launch_bounds__(NT) void k_lss_max(int rows, const double* __restrict__ in, double* __restrict__ out) {
  __shared__ double red[NT];
  double best = -__builtin_inf();
  for (int i = threadIdx.x; i < rows; i += NT) {
    const double v = in[i];
    best = v > best ? v : best;
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = red[threadIdx.x + s];
      red[threadIdx.x] = o > red[threadIdx.x] ? o : red[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = __builtin_isnan(in[0]) ? in[0] : red[0];
}

// ---- directed roundings from nearest-rounded operations --------------------
// next double above / below x for finite x >= 0 (x > 0 for below)
VCPHIP_DEVINL double next_up_nn(double x) { return __longlong_as_double(__double_as_longlong(x) + 1); }
VCPHIP_DEVINL double next_dn_pos(double x) { return __longlong_as_double(__double_as_longlong(x) - 1); }

// RD and RU of t / d for t >= 0 and 2^-960 <= d (here d is 1 - g with
// 0 <= g < 1, so d >= 2^-53).  q = RN(t / d); when t >= 2^-960 the
// remainder t - q d is a double and the FMA gives it exactly, and its sign
// says on which side of q the quotient lies.  Smaller t is scaled by 2^128
// first; scaling the directed results back by 2^-128 is then corrected to
// the same direction (the scaled-back product is exact unless it is
// subnormal, and multiplying it by 2^128 again is always exact).
VCPHIP_DEVINL void div_dir_core(double t, double d, double& dn, double& up) {
  const double q = t / d;
  const double r = __builtin_fma(-q, d, t);
  up = r > 0.0 ? next_up_nn(q) : q;
  dn = (r < 0.0 && q > 0.0) ? next_dn_pos(q) : q;
}

VCPHIP_DEVINL void div_dir(double t, double d, double& dn, double& up) {
  const double tiny = vcphip_pow2(-960);
  if (!(t < tiny)) {   // also t = inf or NaN: q carries it through
    div_dir_core(t, d, dn, up);
    return;
  }
  double sdn, sup;
  div_dir_core(t * vcphip_pow2(128), d, sdn, sup);
  const double s = vcphip_pow2(-128), S = vcphip_pow2(128);
  up = sup * s;
  if (up * S < sup) up = next_up_nn(up);
  dn = sdn * s;
  if (dn * S > sdn) dn = next_dn_pos(dn);
}

// 1 - g rounded down and up (TwoSum: the error of s = RN(1 - g) is exact)
VCPHIP_DEVINL void one_minus_dir(double g, double& dn, double& up) {
  const double a = 1.0, b = -g;
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  up = e > 0.0 ? next_up_nn(s) : s;
  dn = e < 0.0 ? next_dn_pos(s) : s;
}

// G = T / (1 - g) as an interval (imats.hpp:341; g is ||RA - I|| with its
// lower end set to its upper end, T likewise), written to every element of
// the n x ncol interval matrix Tm (imats.hpp:343-348) and to Gout.  When
// g >= 1 (verification failed) or g is NaN, G is [NaN, NaN].
__global__ __launch_bounds__(NT) void k_lss_fill_g(int n, int ncol, const double* __restrict__ norms,
                                                   double* __restrict__ Tm, long ldt, double* __restrict__ Gout) {
  const double g = norms[0], t = norms[1];
  double glo, ghi;
  if (g < 1.0) {
    double ddn, dup;
    one_minus_dir(g, ddn, dup);
    double lo_dn, lo_up, hi_dn, hi_up;
    div_dir(t, dup, lo_dn, lo_up);   // lower end: RD(t / RU(1 - g))
    div_dir(t, ddn, hi_dn, hi_up);   // upper end: RU(t / RD(1 - g))
    glo = lo_dn;
    ghi = hi_up;
  } else {
    glo = ghi = __builtin_nan("");
  }
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  if (idx == 0) {
    Gout[0] = glo;
    Gout[1] = ghi;
  }
  if (idx >= (long)n * ncol) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  *reinterpret_cast<double2*>(Tm + 2 * (i + ldt * j)) = make_double2(glo, ghi);
}

// X = mx + [-u, u] with u = RU(|y|u + Wu) (imats.hpp:352-361):
// [RD(-u + mx), RU(u + mx)]
__global__ __launch_bounds__(NT) void k_lss_final(int n, int ncol, const double* __restrict__ Y,
                                                  const double* __restrict__ W, const double* __restrict__ mx,
                                                  double* __restrict__ X, long ldx) {
  vcphip_mode_round_up();
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  if (idx >= (long)n * ncol) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  const double yu = Y[2 * idx + 1], wu = W[2 * idx + 1], m = mx[idx];
  const double u = yu + wu;
  *reinterpret_cast<double2*>(X + 2 * (i + ldx * j)) = make_double2(-(u - m), u + m);
}

inline dim3 grid_of(long count) { return dim3((unsigned)((count + NT - 1) / NT)); }

// ===========================================================================
// Batched form: many systems of uneven size, each stage one launch over the
// (problem, row, column) items of all problems.  The products are evaluated
// one output per thread with exactly the chains and epilogue of the k_drgemm
// tiles (k ascending, midrad on load, the same fused operations in the same
// order), so every result is bit-identical to the single-problem form.
// ===========================================================================

// problem p's items are [pb[p], pb[p+1]) of the flat item space
VCPHIP_DEVINL int find_problem(const long* __restrict__ pb, int count, long item) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pb[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// one k of the chains of one output of the k_drgemm forms: the operands as
// loaded (an interval as its AoS pair, a point in .x), midpoint-radius taken
// here, every chain advanced in the tile kernels' order.  Round-up mode.
template <bool AI, bool BI>
VCPHIP_DEVINL void dr_dot_step(double2 av, double2 bv, double& up, double& dn, double& r1, double& r2) {
  constexpr bool II = AI && BI;
  double am, ar = 0.0, bm, br = 0.0;
  if (AI) vcphip_midrad_ru(av.x, av.y, am, ar);
  else am = av.x;
  if (BI) vcphip_midrad_ru(bv.x, bv.y, bm, br);
  else bm = bv.x;
  up = __builtin_fma(am, bm, up);
  dn = __builtin_fma(-am, bm, dn);
  if (II) {
    const double as = __builtin_fabs(am) + ar;
    r1 = __builtin_fma(as, br, r1);
    r2 = __builtin_fma(ar, __builtin_fabs(bm), r2);
  } else if (AI) {
    r1 = __builtin_fma(ar, __builtin_fabs(bm), r1);
  } else if (BI) {
    r1 = __builtin_fma(__builtin_fabs(am), br, r1);
  }
}

// one output of the k_drgemm forms: A row r (stride lda over k), B column c
// (stride 1 over k); AI / BI: operand is an interval (AoS pairs).
// Round-up mode.  Returns {lower, upper}.
template <bool AI, bool BI>
VCPHIP_DEVINL double2 dr_dot(int k, const double* __restrict__ A, long lda, const double* __restrict__ B) {
  constexpr bool II = AI && BI;
  double up = 0.0, dn = 0.0, r1 = 0.0, r2 = 0.0;
  // Each thread's chain is sequential in k and these launches hold few waves
  // per SIMD, so their time is load latency: the operands of U consecutive k
  // are loaded first (U loads of each operand in flight per thread), then
  // chained in order.
  constexpr int U = (AI || BI) ? 16 : 32;
  int k0 = 0;
  for (; k0 + U <= k; k0 += U) {
    double2 av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (AI) av[u] = *reinterpret_cast<const double2*>(A + 2 * lda * (k0 + u));
      else av[u].x = A[lda * (k0 + u)];
      if (BI) bv[u] = *reinterpret_cast<const double2*>(B + 2 * (k0 + u));
      else bv[u].x = B[k0 + u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) dr_dot_step<AI, BI>(av[u], bv[u], up, dn, r1, r2);
  }
  for (int kk = k0; kk < k; ++kk) {
    double2 av, bv;
    if (AI) av = *reinterpret_cast<const double2*>(A + 2 * lda * kk);
    else av.x = A[lda * kk];
    if (BI) bv = *reinterpret_cast<const double2*>(B + 2 * kk);
    else bv.x = B[kk];
    dr_dot_step<AI, BI>(av, bv, up, dn, r1, r2);
  }
  const double t = II ? r1 + r2 : r1;
  const double hi = up + t;
  const double nlo = dn + t;
  return make_double2(-nlo, hi);
}

// mb = mid(b) and mx = R mb need all of mb: two launches.
__global__ __launch_bounds__(NT) void k_lssb_mid(int count, const long* __restrict__ pb,
                                                 const VcphipLssProblem* __restrict__ probs) {
  const long item = (long)blockIdx.x * NT + threadIdx.x;
  if (item >= pb[count]) return;
  const int p = find_problem(pb, count, item);
  const VcphipLssProblem P = probs[p];
  const long q = item - pb[p];
  const int i = (int)(q % P.n), j = (int)(q / P.n);
  const double2 v = *reinterpret_cast<const double2*>(P.b + 2 * (i + P.ldb * (long)j));
  double m;
  if (v.x == v.y) {
    m = v.x;
  } else {
    const double s = v.x + v.y;
    m = __builtin_isinf(s) ? 0.5 * v.x + 0.5 * v.y : s * 0.5;
  }
  P.mb[q] = m;
}

__global__ __launch_bounds__(NT) void k_lssb_mx(int count, const long* __restrict__ pb,
                                                const VcphipLssProblem* __restrict__ probs) {
  const long item = (long)blockIdx.x * NT + threadIdx.x;
  if (item >= pb[count]) return;
  const int p = find_problem(pb, count, item);
  const VcphipLssProblem P = probs[p];
  const long q = item - pb[p];
  const int i = (int)(q % P.n), j = (int)(q / P.n);
  const double* a = P.R + i;
  const double* bj = P.mb + (long)P.n * j;
  double acc = 0.0;
  // 32 k of loads in flight per thread, then the chain in k order
  constexpr int U = 32;
  int k0 = 0;
  for (; k0 + U <= P.n; k0 += U) {
    double av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      av[u] = a[P.ldr * (long)(k0 + u)];
      bv[u] = bj[k0 + u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc = __builtin_fma(av[u], bv[u], acc);
  }
  for (int kk = k0; kk < P.n; ++kk) acc = __builtin_fma(a[P.ldr * (long)kk], bj[kk], acc);
  P.mx[q] = acc;
}

// x = IA mx - b (IP product, then the interval subtraction)
__global__ __launch_bounds__(NT) void k_lssb_ax(int count, const long* __restrict__ pb,
                                                const VcphipLssProblem* __restrict__ probs) {
  vcphip_mode_round_up();
  const long item = (long)blockIdx.x * NT + threadIdx.x;
  if (item >= pb[count]) return;
  const int p = find_problem(pb, count, item);
  const VcphipLssProblem P = probs[p];
  const long q = item - pb[p];
  const int i = (int)(q % P.n), j = (int)(q / P.n);
  const double2 x = dr_dot<true, false>(P.n, P.A + 2 * i, P.lda, P.mx + (long)P.n * j);
  const double2 bv = *reinterpret_cast<const double2*>(P.b + 2 * (i + P.ldb * (long)j));
  *reinterpret_cast<double2*>(P.xi + 2 * q) = make_double2(-(bv.y - x.x), x.y - bv.x);
}

// y = |R x| (PI product, then kv's abs)
__global__ __launch_bounds__(NT) void k_lssb_ry(int count, const long* __restrict__ pb,
                                                const VcphipLssProblem* __restrict__ probs) {
  vcphip_mode_round_up();
  const long item = (long)blockIdx.x * NT + threadIdx.x;
  if (item >= pb[count]) return;
  const int p = find_problem(pb, count, item);
  const VcphipLssProblem P = probs[p];
  const long q = item - pb[p];
  const int i = (int)(q % P.n), j = (int)(q / P.n);
  double2 v = dr_dot<false, true>(P.n, P.R + i, P.ldr, P.xi + 2 * (long)P.n * j);
  if (v.x < 0.0) {
    if (v.y <= 0.0) v = make_double2(-v.y, -v.x);
    else v = make_double2(0.0, -v.x > v.y ? -v.x : v.y);
  }
  *reinterpret_cast<double2*>(P.y + 2 * q) = v;
}

// T = ||y||_inf per problem (one workgroup each): upward row sums over the
// columns, then the maximum with the reference's rule (as k_lss_max)
__global__ __launch_bounds__(NT) void k_lssb_ynorm(const VcphipLssProblem* __restrict__ probs) {
  vcphip_mode_round_up();
  __shared__ double red[NT];
  __shared__ double first;
  const VcphipLssProblem P = probs[blockIdx.x];
  double best = -__builtin_inf();
  for (int i = threadIdx.x; i < P.n; i += NT) {
    double s = 0.0;
    for (int j = 0; j < P.nrhs; ++j) {
      const double2 v = *reinterpret_cast<const double2*>(P.y + 2 * (i + (long)P.n * j));
      const double a = __builtin_fabs(v.x), b = __builtin_fabs(v.y);
      const double mag = a > b ? a : b;
      s = (j == 0) ? mag : s + mag;
    }
    if (i == 0) first = s;
    best = s > best ? s : best;
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      const double o = red[threadIdx.x + st];
      red[threadIdx.x] = o > red[threadIdx.x] ? o : red[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) P.norms[1] = __builtin_isnan(first) ? first : red[0];
}

// G = T / (1 - g) per problem (one thread each, round-to-nearest mode: the
// exact-error method of k_lss_fill_g), into norms[2], norms[3]
__global__ __launch_bounds__(NT) void k_lssb_g(int count, const VcphipLssProblem* __restrict__ probs) {
  const int p = blockIdx.x * NT + threadIdx.x;
  if (p >= count) return;
  const VcphipLssProblem P = probs[p];
  const double g = P.normub[0], t = P.norms[1];
  double glo, ghi;
  if (g < 1.0) {
    double ddn, dup, a, b;
    one_minus_dir(g, ddn, dup);
    div_dir(t, dup, glo, a);
    div_dir(t, ddn, b, ghi);
  } else {
    glo = ghi = __builtin_nan("");
  }
  P.norms[2] = glo;
  P.norms[3] = ghi;
}

// W = |R IA - I| G (the II product with the constant column G; kv's abs()
// taken on load) and X = mx + [-u, u], u = RU(|y|u + Wu).  Round-up mode.
__global__ __launch_bounds__(NT) void k_lssb_final(int count, const long* __restrict__ pb,
                                                   const VcphipLssProblem* __restrict__ probs) {
  vcphip_mode_round_up();
  const long item = (long)blockIdx.x * NT + threadIdx.x;
  if (item >= pb[count]) return;
  const int p = find_problem(pb, count, item);
  const VcphipLssProblem P = probs[p];
  const long q = item - pb[p];
  const int i = (int)(q % P.n), j = (int)(q / P.n);
  const double2 G = make_double2(P.norms[2], P.norms[3]);
  double bm, br;
  vcphip_midrad_ru(G.x, G.y, bm, br);
  double up = 0.0, r1 = 0.0, r2 = 0.0;   // the lower chain of W is not used
  const double* crow = P.C + 2 * i;
  auto step = [&](double2 v) {
    if (v.x < 0.0) {
      if (v.y <= 0.0) v = make_double2(-v.y, -v.x);
      else v = make_double2(0.0, -v.x > v.y ? -v.x : v.y);
    }
    double am, ar;
    vcphip_midrad_ru(v.x, v.y, am, ar);
    up = __builtin_fma(am, bm, up);
    const double as = __builtin_fabs(am) + ar;
    r1 = __builtin_fma(as, br, r1);
    r2 = __builtin_fma(ar, __builtin_fabs(bm), r2);
  };
  // a thread per row: 32 columns of loads in flight, then the chain in order
  constexpr int U = 32;
  int kk = 0;
  for (; kk + U <= P.n; kk += U) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const double2*>(crow + 2 * (long)P.n * (kk + u));
#pragma unroll
    for (int u = 0; u < U; ++u) step(v[u]);
  }
  for (; kk < P.n; ++kk) step(*reinterpret_cast<const double2*>(crow + 2 * (long)P.n * kk));
  const double wu = up + (r1 + r2);
  const double yu = P.y[2 * q + 1], m = P.mx[q];
  const double u = yu + wu;
  *reinterpret_cast<double2*>(P.X + 2 * (i + P.ldx * (long)j)) = make_double2(-(u - m), u + m);
}

}  // namespace

int vcphip_lss_launch_mid(int n, int ncol, const double* b, long ldb, double* mb, long ldm, hipStream_t s) {
  hipLaunchKernelGGL(k_lss_mid, grid_of((long)n * ncol), dim3(NT), 0, s, n, ncol, b, ldb, mb, ldm);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_inv_launch_pad(int n, int npad, const double* IA, long lda, double* P, hipStream_t s) {
  hipLaunchKernelGGL(k_inv_pad, grid_of((long)npad * npad), dim3(NT), 0, s, n, npad, IA, lda, P);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_inv_launch_copy(int n, const double* P, long ldp, double* R, long ldr, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_block, grid_of((long)n * n), dim3(NT), 0, s, n, P, ldp, R, ldr);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_gemv_rn(int m, int ncol, int k, const double* A, long lda, const double* B, long ldb, double* C,
                              long ldc, hipStream_t s) {
  hipLaunchKernelGGL(k_lss_gemv_rn, dim3((m + NT - 1) / NT, ncol), dim3(NT), 0, s, m, k, A, lda, B, ldb, C, ldc);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_sub(int n, int ncol, double* X, long ldx, const double* b, long ldb, hipStream_t s) {
  hipLaunchKernelGGL(k_lss_sub, grid_of((long)n * ncol), dim3(NT), 0, s, n, ncol, X, ldx, b, ldb);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_abs(int rows, int cols, double* X, long ldx, hipStream_t s) {
  hipLaunchKernelGGL(k_lss_abs, grid_of((long)rows * cols), dim3(NT), 0, s, rows, cols, X, ldx);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_diag(int n, double* C, long ldc, hipStream_t s) {
  hipLaunchKernelGGL(k_lss_diag, grid_of(n), dim3(NT), 0, s, n, C, ldc);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_norm(int rows, int cols, const double* X, long ldx, double* rowbuf, double* out,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_lss_rowsum, grid_of(rows), dim3(NT), 0, s, rows, cols, X, ldx, rowbuf);
  hipLaunchKernelGGL(k_lss_max, dim3(1), dim3(NT), 0, s, rows, rowbuf, out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_fill_g(int n, int ncol, const double* norms, double* Tm, long ldt, double* Gout,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_lss_fill_g, grid_of((long)n * ncol), dim3(NT), 0, s, n, ncol, norms, Tm, ldt, Gout);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_final(int n, int ncol, const double* Y, const double* W, const double* mx, double* X, long ldx,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_lss_final, grid_of((long)n * ncol), dim3(NT), 0, s, n, ncol, Y, W, mx, X, ldx);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_batched_rhs(int count, long items, const long* pb, const VcphipLssProblem* probs,
                                  hipStream_t s) {
  const dim3 g = grid_of(items);
  hipLaunchKernelGGL(k_lssb_mid, g, dim3(NT), 0, s, count, pb, probs);
  hipLaunchKernelGGL(k_lssb_mx, g, dim3(NT), 0, s, count, pb, probs);
  hipLaunchKernelGGL(k_lssb_ax, g, dim3(NT), 0, s, count, pb, probs);
  hipLaunchKernelGGL(k_lssb_ry, g, dim3(NT), 0, s, count, pb, probs);
  hipLaunchKernelGGL(k_lssb_ynorm, dim3(count), dim3(NT), 0, s, probs);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_lss_launch_batched_bound(int count, long items, const long* pb, const VcphipLssProblem* probs,
                                    hipStream_t s) {
  hipLaunchKernelGGL(k_lssb_g, grid_of(count), dim3(NT), 0, s, count, probs);
  hipLaunchKernelGGL(k_lssb_final, grid_of(items), dim3(NT), 0, s, count, pb, probs);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
