// Double-double matrix product on int8 MFMA (gfx950): the vcp::pddblas
// product (vcp/pddblas.hpp:318-376, :379-409) by an Ozaki scheme II.
//
// Each dd entry a = hi + lo of row i is scaled to the integer
//   A' = RN(hi 2^(beta - e_i)) + RN(lo 2^(beta - e_i)),  |A'| <= 2^beta + 1,
// where max_k |hi| + |lo| < 2^e_i (one exponent per row of A / column of B,
// the sigma of split_row / split_col, pddblas.hpp:82-154).  A'B' is exact
// through N int8 modular GEMMs (shared with the interval path) and the
// Chinese remainder theorem; the integer is rounded to the nearest double
// (hi) and its remainder to the nearest double (lo).  The CPU oracle
// restates every step and the GPU result is bit-identical to it.
#include "vcphip_device.h"
#include "vcphip_kernels.h"
#include "vcphip_crt.h"
#include "vcphip_ddconv.h"

namespace {

typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// stats: per row (or column) the exponent of max |hi| + |lo| (round-up sum)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dd_stats_part(const double* __restrict__ X, long rs, long cs, int rows,
                                                       int len, int nchunk, int* __restrict__ part) {
  vcphip_mode_round_up();
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= rows) return;
  const int c0 = c * VCPHIP_DD_CHUNK, c1 = min(c0 + VCPHIP_DD_CHUNK, len);
  double mu = 0.0;
  bool nonfinite = false;
  auto take = [&](const double2 v) {
    nonfinite |= !(__builtin_isfinite(v.x) && __builtin_isfinite(v.y));
    const double a = __builtin_fabs(v.x) + __builtin_fabs(v.y);
    mu = a > mu ? a : mu;
  };
  const double* x = X + 2 * (long)r * rs;
  int cc = c0;
  for (; cc + 3 < c1; cc += 4) {   // four loads in flight per lane
    double2 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = *reinterpret_cast<const double2*>(x + 2 * (long)(cc + q) * cs);
#pragma unroll
    for (int q = 0; q < 4; ++q) take(v[q]);
  }
  for (; cc < c1; ++cc) take(*reinterpret_cast<const double2*>(x + 2 * (long)cc * cs));
  // INT32_MAX marks a non-finite element (it wins the maximum over chunks)
  part[(long)c * rows + r] = nonfinite ? INT32_MAX : (mu > 0.0 ? vcphip_fexp(mu) : INT32_MIN);
}

__global__ __launch_bounds__(256) void k_dd_stats_final(int rows, int nchunk, const int* __restrict__ part,
                                                        int* __restrict__ e, int* __restrict__ flags) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  int best = INT32_MIN;
  for (int c = 0; c < nchunk; ++c) best = max(best, part[(long)c * rows + r]);
  // flags: 1 = zero row / column, 2 = a non-finite element
  e[r] = (best == INT32_MIN || best == INT32_MAX) ? 0 : best;
  flags[r] = best == INT32_MIN ? 1 : (best == INT32_MAX ? 2 : 0);
}

// the same statistics for lines whose elements are contiguous (columns of B,
// rows of a transposed A): one wave per line, its lanes over the elements
// (coalesced), the maximum taken across the wave (it does not depend on the
// order) and written directly.  fexp is monotone, so the exponent of the
// maximum equals the maximum of the chunk exponents above.
__global__ __launch_bounds__(256) void k_dd_stats_lines(const double* __restrict__ X, long rs, int rows, int len,
                                                        int* __restrict__ e, int* __restrict__ flags) {
  vcphip_mode_round_up();
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;   // uniform per wave
  const double* x = X + 2 * (long)r * rs;
  double mu = 0.0;
  int nonfinite = 0;
  auto take = [&](const double2 v) {
    nonfinite |= !(__builtin_isfinite(v.x) && __builtin_isfinite(v.y));
    const double a = __builtin_fabs(v.x) + __builtin_fabs(v.y);
    mu = a > mu ? a : mu;
  };
  int cc = lane;
  for (; cc + 192 < len; cc += 256) {   // four loads in flight per lane
    double2 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = *reinterpret_cast<const double2*>(x + 2 * (long)(cc + 64 * q));
#pragma unroll
    for (int q = 0; q < 4; ++q) take(v[q]);
  }
  for (; cc < len; cc += 64) take(*reinterpret_cast<const double2*>(x + 2 * (long)cc));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(mu, off);
    mu = o > mu ? o : mu;
    nonfinite |= __shfl_xor(nonfinite, off);
  }
  if (lane == 0) {
    // flags: 1 = zero line, 2 = a non-finite element
    e[r] = (nonfinite || !(mu > 0.0)) ? 0 : vcphip_fexp(mu);
    flags[r] = nonfinite ? 2 : (mu > 0.0 ? 0 : 1);
  }
}

// ---------------------------------------------------------------------------
// split: 128-bit scaled integers -> symmetric int8 residues
// ---------------------------------------------------------------------------
// nearest integer (ties away) of the exact double y, as sign + magnitude
VCPHIP_DEVINL void dbl_to_int_rn(double y, bool& neg, u128& mag) {
  neg = y < 0.0;
  const double a = __builtin_fabs(y);
  if (!(a >= 0.5)) {
    mag = 0;
    return;
  }
  const int e = __builtin_amdgcn_frexp_exp(a);                              // a = f 2^e, f in [0.5, 1)
  const unsigned long long m = (unsigned long long)__builtin_ldexp(__builtin_amdgcn_frexp_mant(a), 53);
  const int p = e - 53;                                                     // a = m 2^p, p >= -53
  if (p >= 0) mag = (u128)m << p;
  else mag = (u128)((m + (1ull << (-p - 1))) >> (-p));
}

// A' = RN(hi 2^s) + RN(lo 2^s) as sign + magnitude
VCPHIP_DEVINL void dd_scaled(double hi, double lo, int s, bool& neg, u128& mag) {
  bool nh, nl;
  u128 mh, ml;
  dbl_to_int_rn(__builtin_ldexp(hi, s), nh, mh);
  dbl_to_int_rn(__builtin_ldexp(lo, s), nl, ml);
  if (nh == nl) {
    neg = nh;
    mag = mh + ml;
  } else if (mh >= ml) {
    neg = nh;
    mag = mh - ml;
  } else {
    neg = nl;
    mag = ml - mh;
  }
  if (mag == 0) neg = false;
}

// the magnitude (< 2^120) as its four 32-bit words, low first
VCPHIP_DEVINL void words4(u128 mag, unsigned (&p)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) p[i] = (unsigned)(mag >> (32 * i));
}

// The centred residues of the thread's four elements (magnitudes c[u] < 2^120
// as four words, signs neg[u]) for every modulus: emit(t, word) receives the
// word whose byte u is the residue of (neg[u] ? -mag : mag) modulo m_t in
// [-floor(m_t/2), m_t - 1 - floor(m_t/2)] (for m_t = 256 the low byte of
// +-mag).
//   y_t = sum_i b_i (256^i mod m_t) over the 16 bytes b_i of mag (< 2^20) is
// an int8 matrix product of centred operands: sum_i (b_i - 128)(w_ti - 128) +
// 128 sum_i b_i + dd_kc[t].  One v_mfma_i32_32x32x32_i8 takes 64 elements
// (one per lane) against 16 moduli: output row r is modulus 4 (r / 8) + r % 4
// of element group (r / 4) % 2, so lane register v holds modulus 16 ch + v of
// the lane's own element; the B operand is the lane's own 16 centred bytes
// (k-slots of lane half h carry element group h) and the A operand of lane
// half h is zero in the other group's rows.  Every lane must run it (the
// split kernels have no early exit).
//   q = floor((y + floor(m/2)) / m) = vcphip_quot24(8 (y + floor(m/2)), m) is
// exact, so c = y - q m lies in [-floor(m/2), m - 1 - floor(m/2)]
// (v_mad_i32_i24), and -c for a negative element is (c ^ -1) + 1 (v_xad_u32).  Moduli are taken 16 at a time, the
// four elements' products of a group at once.
template <int N, typename Emit>
VCPHIP_DEVINL void dd_residue_words(const unsigned (&c)[4][4], const bool (&neg)[4], const VcphipOzParams& P,
                                    Emit emit) {
  constexpr int CH = (N + 15) / 16;
  const unsigned lane = threadIdx.x & 63u, h = lane >> 5, r = lane & 31u;
  const bool mine = ((r >> 2) & 1u) == h;
  const unsigned cl = 4 * (r >> 3) + (r & 3u);
  vcphip_v4i b[4];
  int sb128[4];
  unsigned sx[4], s1[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    unsigned sb = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      b[u][e] = (int)(c[u][e] ^ 0x80808080u);
      sb = __builtin_amdgcn_udot4(c[u][e], 0x01010101u, sb, false);
    }
    sb128[u] = (int)(sb << 7);
    sx[u] = neg[u] ? ~0u : 0u;
    s1[u] = neg[u] ? 1u : 0u;
  }
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) {
    const unsigned tl = 16 * ch + cl;
    const bool use = mine && tl < (unsigned)N;
    const vcphip_v4i w = *reinterpret_cast<const vcphip_v4i*>(&P.dd_wb[use ? tl : 0][0]);
    vcphip_v4i a;
#pragma unroll
    for (int e = 0; e < 4; ++e) a[e] = use ? (int)((unsigned)w[e] ^ 0x80808080u) : 0;
    // element by element (one accumulator live), each residue inserted as
    // byte u of its modulus' word
    unsigned word[16];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const vcphip_v16i acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b[u], vcphip_v16i{}, 0, 0, 0);
      const unsigned sel = 0x03020100u | (0x04u << (8 * u));   // word's bytes, byte u from the residue
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int t = 16 * ch + v;
        if (t >= N) break;
        const unsigned m = (unsigned)P.mod[t] & 0x1FFu, h8 = 8u * (m >> 1);
        const unsigned y = (unsigned)(acc[v] + sb128[u] + P.dd_kc[t]);   // < 2^20
        const unsigned q = vcphip_quot24((y << 3) + h8, P.r_magic[t]);
        const unsigned res = vcphip_xad(vcphip_mad_i24(q, -(int)m, y), sx[u], s1[u]);
        word[v] = u == 0 ? res : __builtin_amdgcn_perm(res, word[v], (sel & ~(0xFFu << (8 * u))) | (0x04u << (8 * u)));
      }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v)
      if (16 * ch + v < N) emit(16 * ch + v, word[v]);
  }
}

constexpr int SR_ROWS = 16, SR_K = 64, SR_LD = 68;   // one 1 KB plane block per workgroup and plane

// rows of A (m x k, column-major dd) -> N residue planes of mpad x kpad.
// Thread: one row and a quad of 4 consecutive k (16 rows x 16 quads); the 4
// residues of a plane go to LDS as one word, then each plane's 1 KB block is
// written with one coalesced word per thread.
template <int N>
__global__ __launch_bounds__(256) void k_dd_split_rows(const double* __restrict__ A, long lda, int m, int k, int mpad,
                                                       int kpad, VcphipOzParams P, const int* __restrict__ ea,
                                                       const int* __restrict__ fa, int8_t* __restrict__ planes) {
  __shared__ __attribute__((aligned(16))) unsigned char s_tile[N * SR_ROWS * SR_LD];
  const int row0 = blockIdx.x * SR_ROWS, k0 = blockIdx.y * SR_K;
  const int tid = threadIdx.x;
  const int lr = tid % SR_ROWS, lk0 = 4 * (tid / SR_ROWS);
  const int r = row0 + lr;
  const bool rowok = r < m && fa[min(r, m - 1)] == 0;
  const int sc = rowok ? P.beta - ea[r] : 0;
  double2 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int kk = k0 + lk0 + u;
    v[u] = (rowok && kk < k) ? *reinterpret_cast<const double2*>(A + 2 * (r + lda * kk)) : make_double2(0.0, 0.0);
  }
  unsigned c[4][4];
  bool neg[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    u128 mag;
    dd_scaled(v[u].x, v[u].y, sc, neg[u], mag);
    words4(mag, c[u]);
  }
  dd_residue_words<N>(c, neg, P, [&](int t, unsigned word) {
    *reinterpret_cast<unsigned*>(&s_tile[(t * SR_ROWS + lr) * SR_LD + lk0]) = word;
  });
  __syncthreads();
  const int orow = tid / 16, oc = (tid % 16) * 4;
  const long plane_stride = (long)mpad * kpad;
  int8_t* dst = planes + vcphip_plane_off(row0 + orow, k0 + oc, kpad);
#pragma unroll 4
  for (int p = 0; p < N; ++p)
    *reinterpret_cast<unsigned*>(dst + p * plane_stride) =
        *reinterpret_cast<const unsigned*>(&s_tile[(p * SR_ROWS + orow) * SR_LD + oc]);
}

// columns of B (k x n, column-major dd) -> N residue planes of npad x kpad.
// Thread: a quad of 4 consecutive k of one column; 16 consecutive threads
// cover 64 k of a column (1 KB of contiguous input), 256 threads one 1 KB
// plane block (16 columns x 64 k), written as 4 contiguous 256-byte runs.
template <int N>
__global__ __launch_bounds__(256) void k_dd_split_cols(const double* __restrict__ B, long ldb, int k, int n, int npad,
                                                       int kpad, VcphipOzParams P, const int* __restrict__ eb,
                                                       const int* __restrict__ fb, int8_t* __restrict__ planes) {
  const int kblk = kpad / 64, unit = blockIdx.x, t = threadIdx.x;
  const int j = (unit / kblk) * 16 + (t >> 4), k0 = (unit % kblk) * 64 + 4 * (t & 15);
  const bool colok = j < n && fb[min(j, n - 1)] == 0;
  const int sc = colok ? P.beta - eb[j] : 0;
  double2 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int kk = k0 + u;
    v[u] = (colok && kk < k) ? *reinterpret_cast<const double2*>(B + 2 * (kk + ldb * (long)j)) : make_double2(0.0, 0.0);
  }
  unsigned c[4][4];
  bool neg[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    u128 mag;
    dd_scaled(v[u].x, v[u].y, sc, neg[u], mag);
    words4(mag, c[u]);
  }
  const long plane_stride = (long)npad * kpad;
  int8_t* dst = planes + vcphip_plane_off(j, k0, kpad);
  dd_residue_words<N>(c, neg, P, [&](int z, unsigned word) {
    *reinterpret_cast<unsigned*>(dst + z * plane_stride) = word;
  });
}

// ---------------------------------------------------------------------------
// reconstruction: CRT -> 256-bit integer -> nearest (hi, lo)
// ---------------------------------------------------------------------------
constexpr int NL = VCPHIP_OZ_NLIMB;
static_assert(NL == vcphip_ddconv::NL, "limb count of the conversion");

template <int N>
__global__ __launch_bounds__(256) void k_dd_recon(int m, int n, int mpad, int npad, VcphipOzParams P,
                                                  const int* __restrict__ ea, const int* __restrict__ fa,
                                                  const int* __restrict__ eb, const int* __restrict__ fb,
                                                  const uint8_t* __restrict__ U, double* __restrict__ C, long ldc) {
  // threads follow the blocked 16 x 16 layout of the residue planes
  // grid: x = 16-row block, y = 16-column block (the plane's block order)
  // the block's planes from a uniform base plus the thread's offset o
  // every lane runs through the CRT's matrix instructions; lanes past m or n
  // (inside the padded planes) read row / column 0's tables and store nothing
  const unsigned o = threadIdx.x;
  const long bbase = ((long)blockIdx.y * (mpad >> 4) + blockIdx.x) * 256;
  const unsigned i0 = blockIdx.x * 16 + 4 * (o >> 6) + (o & 3);
  const unsigned j0 = blockIdx.y * 16 + ((o >> 2) & 15);
  const bool live = i0 < (unsigned)m && j0 < (unsigned)n;
  const unsigned i = live ? i0 : 0u, j = live ? j0 : 0u;
  const long plane = (long)mpad * npad;
  VcphipCrtOps<N> T;
  vcphip_crt_ops_load<N>(P, T);
  unsigned mag[NL];
  const bool neg = vcphip_crt_symmetric<N, NL>(U + bbase, plane, o, T, P, mag);
  const int flag = fa[i] | fb[j];
  if (flag) {   // a zero row / column gives 0, a non-finite one NaN
    const double v = (flag & 2) ? __builtin_nan("") : 0.0;
    if (live) *reinterpret_cast<double2*>(C + 2 * (i + ldc * (long)j)) = make_double2(v, v);
    return;
  }
  const int sc = ea[i] + eb[j] - 2 * P.beta;
  double hi, lo;
  vcphip_dd_from_limbs(mag, sc, hi, lo);
  const double sgn = neg ? -1.0 : 1.0;
  if (live) *reinterpret_cast<double2*>(C + 2 * (i + ldc * (long)j)) = make_double2(sgn * hi, sgn * lo);
}

// ---------------------------------------------------------------------------
// residual R = B - R (dd, elementwise): the r = b - A x step of
// pddblas::linearsolve (vcp/pddblas.hpp:461-466) after the product A x has
// been written to R.  Round-to-nearest double-double subtraction with two
// TwoSums (Knuth) and two renormalisations (Dekker's FastTwoSum); kv::dd's
// own operator bits are not available to pin against.
// ---------------------------------------------------------------------------
VCPHIP_DEVINL void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
VCPHIP_DEVINL void fast_two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  e = b - (s - a);
}

__global__ __launch_bounds__(256) void k_dd_sub_from(int m, int n, const double* __restrict__ B, long ldb,
                                                     double* __restrict__ R, long ldr) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)m * n) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  const double2 b = *reinterpret_cast<const double2*>(B + 2 * (i + ldb * (long)j));
  double2* r = reinterpret_cast<double2*>(R + 2 * (i + ldr * (long)j));
  const double2 p = *r;
  double s, e, t, f;
  two_sum(b.x, -p.x, s, e);
  two_sum(b.y, -p.y, t, f);
  e = e + t;
  fast_two_sum(s, e, s, e);
  e = e + f;
  fast_two_sum(s, e, s, e);
  *r = make_double2(s, e);
}

// hi parts of a dd matrix: D(i, j) = X(i, j).hi, D with leading dimension m
// D(i, j) = hi(X(rows[i], j)), rows = identity when null (a row permutation
// with entries in [0, m), checked by the caller)
__global__ __launch_bounds__(256) void k_dd_hi(int m, int n, const double* __restrict__ X, long ldx,
                                               const int* __restrict__ rows, double* __restrict__ D) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)m * n) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  const int r = rows ? rows[i] : i;
  D[idx] = X[2 * (r + ldx * (long)j)];
}

// the refinement step of pddblas::linearsolve (vcp/pddblas.hpp:450-457):
// x(i, j) += dd(d(i, j)) (TwoSum of the high parts, the low part added,
// FastTwoSum), and per-block maxima of |d| and of the new |x.hi| taken as the
// reference does (a NaN never raises the maximum).  Grid-stride; part[2 b]
// and part[2 b + 1] receive block b's maxima.
__global__ __launch_bounds__(256) void k_dd_add_d(int m, int n, const double* __restrict__ d,
                                                  double* __restrict__ X, long ldx, double* __restrict__ part) {
  __shared__ double sd[256], sx[256];
  const long total = (long)m * n;
  double md = 0.0, mx = 0.0;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    const int i = (int)(idx % m), j = (int)(idx / m);
    double2* xp = reinterpret_cast<double2*>(X + 2 * (i + ldx * (long)j));
    const double2 x = *xp;
    const double dv = d[idx];
    double s, e, h, l;
    two_sum(x.x, dv, s, e);
    e = e + x.y;
    fast_two_sum(s, e, h, l);
    *xp = make_double2(h, l);
    const double ad = __builtin_fabs(dv), ax = __builtin_fabs(h);
    if (ad > md) md = ad;
    if (ax > mx) mx = ax;
  }
  sd[threadIdx.x] = md;
  sx[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      if (sd[threadIdx.x + w] > sd[threadIdx.x]) sd[threadIdx.x] = sd[threadIdx.x + w];
      if (sx[threadIdx.x + w] > sx[threadIdx.x]) sx[threadIdx.x] = sx[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sx[0];
  }
}

// Cholesky refinement (pddblas::Cholesky, vcp/pddblas.hpp:490-563)
// U = dd(Ud) on and above the diagonal, exact zeros below (:219-223 of
// pdblas::Cholesky zero the strictly lower part of dpotrf's output)
__global__ __launch_bounds__(256) void k_dd_from_upper(int n, const double* __restrict__ D, long ldd,
                                                       double* __restrict__ U, long ldu) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  *reinterpret_cast<double2*>(U + 2 * (i + ldu * (long)j)) = make_double2(i <= j ? D[i + ldd * (long)j] : 0.0, 0.0);
}

// W = dd(triu(G, 1) + diag(G) / 2) (:531-545)
__global__ __launch_bounds__(256) void k_dd_w_from_g(int n, const double* __restrict__ G, double* __restrict__ W) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  const double g = G[idx];
  *reinterpret_cast<double2*>(W + 2 * idx) = make_double2(i < j ? g : (i == j ? g * 0.5 : 0.0), 0.0);
}

// U += F in double-double (:548-550), and U's strictly lower part set to
// exact zeros when zero_lower (:553-558)
__global__ __launch_bounds__(256) void k_dd_add_to(int n, const double* __restrict__ F, long ldf, double* __restrict__ U,
                                                   long ldu, int zero_lower) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  double2* up = reinterpret_cast<double2*>(U + 2 * (i + ldu * (long)j));
  if (zero_lower && i > j) {
    *up = make_double2(0.0, 0.0);
    return;
  }
  const double2 f = *reinterpret_cast<const double2*>(F + 2 * (i + ldf * (long)j));
  const double2 u = *up;
  double s, e, t, g;
  two_sum(u.x, f.x, s, e);
  two_sum(u.y, f.y, t, g);
  e = e + t;
  fast_two_sum(s, e, s, e);
  e = e + g;
  fast_two_sum(s, e, s, e);
  *up = make_double2(s, e);
}

// flag[0] = 1 if A(i, j) != A(j, i) in either part for some i, j (the
// reference's is_symmetric); the caller clears the flag first
__global__ __launch_bounds__(256) void k_dd_symmetric(int n, const double* __restrict__ A, long lda,
                                                      int* __restrict__ flag) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  if (i >= j) return;
  const double2 a = *reinterpret_cast<const double2*>(A + 2 * (i + lda * (long)j));
  const double2 b = *reinterpret_cast<const double2*>(A + 2 * (j + lda * (long)i));
  if (!(a.x == b.x && a.y == b.y)) flag[0] = 1;
}

// X = dd(D) for an m x n double matrix D (lo parts zero)
__global__ __launch_bounds__(256) void k_dd_from_d(int m, int n, const double* __restrict__ D, long ldd,
                                                   double* __restrict__ X, long ldx) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)m * n) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  *reinterpret_cast<double2*>(X + 2 * (i + ldx * (long)j)) = make_double2(D[i + ldd * (long)j], 0.0);
}

// ---- the elementwise part of the Ogita-Aishima step (pddblas::oa_step,
// vcp/pddblas.hpp:219-314), in round-to-nearest double-double arithmetic:
// TwoSum / FastTwoSum additions, the exact product error by FMA, and the
// quotient with two correction steps
VCPHIP_DEVINL void dd_add2(double xh, double xl, double yh, double yl, double& h, double& l) {
  double s, e, t, f;
  two_sum(xh, yh, s, e);
  two_sum(xl, yl, t, f);
  e = e + t;
  fast_two_sum(s, e, s, e);
  e = e + f;
  fast_two_sum(s, e, h, l);
}

VCPHIP_DEVINL void two_prod(double a, double b, double& p, double& e) {
  p = a * b;
  e = __builtin_fma(a, b, -p);
}

VCPHIP_DEVINL void dd_mul2(double xh, double xl, double yh, double yl, double& h, double& l) {
  double p, e;
  two_prod(xh, yh, p, e);
  e = e + (xh * yl + xl * yh);
  fast_two_sum(p, e, h, l);
}

VCPHIP_DEVINL void dd_mul_d2(double xh, double xl, double d, double& h, double& l) {
  double p, e;
  two_prod(xh, d, p, e);
  e = e + xl * d;
  fast_two_sum(p, e, h, l);
}

VCPHIP_DEVINL void dd_div2(double xh, double xl, double yh, double yl, double& h, double& l) {
  double ph, pl, rh, rl;
  double q1 = xh / yh;
  dd_mul_d2(yh, yl, q1, ph, pl);
  dd_add2(xh, xl, -ph, -pl, rh, rl);
  double q2 = rh / yh;
  dd_mul_d2(yh, yl, q2, ph, pl);
  dd_add2(rh, rl, -ph, -pl, rh, rl);
  const double q3 = rh / yh;
  fast_two_sum(q1, q2, q1, q2);
  dd_add2(q1, q2, q3, 0.0, h, l);
}

// R = I - G, element (i, j)
VCPHIP_DEVINL void oa_r(int i, int j, double2 g, double& rh, double& rl) {
  if (i == j) dd_add2(-g.x, -g.y, 1.0, 0.0, rh, rl);
  else {
    rh = -g.x;
    rl = -g.y;
  }
}

// lam_i = S_ii / G_ii (:238-243)
__global__ __launch_bounds__(256) void k_oa_lam(int n, const double* __restrict__ S, long lds,
                                                const double* __restrict__ G, long ldg, double* __restrict__ lam) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 s = *reinterpret_cast<const double2*>(S + 2 * (i + lds * (long)i));
  const double2 g = *reinterpret_cast<const double2*>(G + 2 * (i + ldg * (long)i));
  double h, l;
  dd_div2(s.x, s.y, g.x, g.y, h, l);
  lam[2 * i] = h;
  lam[2 * i + 1] = l;
}

// per row i: sum_j |R_ij.hi| and sum_{j != i} |S_ij.hi| (:253-263), added in
// column order as the reference does; lanes take consecutive rows (coalesced)
__global__ __launch_bounds__(256) void k_oa_rows(int n, const double* __restrict__ S, long lds,
                                                 const double* __restrict__ G, long ldg, double* __restrict__ rows) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double sr = 0.0, ss = 0.0;
  for (int j = 0; j < n; ++j) {
    double rh, rl;
    oa_r(i, j, *reinterpret_cast<const double2*>(G + 2 * (i + ldg * (long)j)), rh, rl);
    sr = sr + __builtin_fabs(rh);
    if (j != i) ss = ss + __builtin_fabs(S[2 * (i + lds * (long)j)]);
  }
  rows[2 * i] = sr;
  rows[2 * i + 1] = ss;
}

// delta = 2 (offS + normR * lmax), normR = max_i sum_j |R_ij|, offS = max_i of
// the off-diagonal row sums of |S|, lmax = max |lam| (one block)
__global__ __launch_bounds__(256) void k_oa_delta(int n, const double* __restrict__ rows,
                                                  const double* __restrict__ lam, double* __restrict__ delta) {
  __shared__ double sr[256], ss[256], sl[256];
  double mr = 0.0, ms = 0.0, ml = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    if (rows[2 * i] > mr) mr = rows[2 * i];
    if (rows[2 * i + 1] > ms) ms = rows[2 * i + 1];
    const double al = __builtin_fabs(lam[2 * i]);
    if (al > ml) ml = al;
  }
  sr[threadIdx.x] = mr;
  ss[threadIdx.x] = ms;
  sl[threadIdx.x] = ml;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      if (sr[threadIdx.x + w] > sr[threadIdx.x]) sr[threadIdx.x] = sr[threadIdx.x + w];
      if (ss[threadIdx.x + w] > ss[threadIdx.x]) ss[threadIdx.x] = ss[threadIdx.x + w];
      if (sl[threadIdx.x + w] > sl[threadIdx.x]) sl[threadIdx.x] = sl[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) delta[0] = 2.0 * (ss[0] + sr[0] * sl[0]);
}

// E_ij = (S_ij + lam_j R_ij) / (lam_j - lam_i) where |lam_j - lam_i| > delta
// and i != j, else R_ij / 2 (:264-300)
__global__ __launch_bounds__(256) void k_oa_e(int n, const double* __restrict__ S, long lds,
                                              const double* __restrict__ G, long ldg, const double* __restrict__ lam,
                                              const double* __restrict__ delta, double* __restrict__ E) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  double rh, rl;
  oa_r(i, j, *reinterpret_cast<const double2*>(G + 2 * (i + ldg * (long)j)), rh, rl);
  const double ljh = lam[2 * j], ljl = lam[2 * j + 1];
  double dh, dl;
  dd_add2(ljh, ljl, -lam[2 * i], -lam[2 * i + 1], dh, dl);
  double eh, el;
  if (i != j && __builtin_fabs(dh) > delta[0]) {
    const double2 sv = *reinterpret_cast<const double2*>(S + 2 * (i + lds * (long)j));
    double mh, ml, nh, nl;
    dd_mul2(ljh, ljl, rh, rl, mh, ml);
    dd_add2(sv.x, sv.y, mh, ml, nh, nl);
    dd_div2(nh, nl, dh, dl, eh, el);
  } else {
    dd_mul_d2(rh, rl, 0.5, eh, el);
  }
  *reinterpret_cast<double2*>(E + 2 * idx) = make_double2(eh, el);
}

// out[0], out[1] = the maxima over nb blocks' partials
__global__ __launch_bounds__(256) void k_max_pairs(int nb, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sd[256], sx[256];
  double md = 0.0, mx = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    if (part[2 * b] > md) md = part[2 * b];
    if (part[2 * b + 1] > mx) mx = part[2 * b + 1];
  }
  sd[threadIdx.x] = md;
  sx[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      if (sd[threadIdx.x + w] > sd[threadIdx.x]) sd[threadIdx.x] = sd[threadIdx.x + w];
      if (sx[threadIdx.x + w] > sx[threadIdx.x]) sx[threadIdx.x] = sx[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sd[0];
    out[1] = sx[0];
  }
}

}  // namespace

int vcphip_dd_hi(int m, int n, const double* X, long ldx, double* D, hipStream_t s) {
  if ((long)m * n == 0) return 0;
  return vcphip_dd_hi_rows(m, n, X, ldx, nullptr, D, s);
}

int vcphip_dd_hi_rows(int m, int n, const double* X, long ldx, const int* rows, double* D, hipStream_t s) {
  if ((long)m * n <= 0) return 0;
  hipLaunchKernelGGL(k_dd_hi, dim3((unsigned)(((long)m * n + 255) / 256)), dim3(256), 0, s, m, n, X, ldx, rows, D);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

static inline dim3 grid_nn(int n) { return dim3((unsigned)(((long)n * n + 255) / 256)); }

int vcphip_dd_from_upper(int n, const double* D, long ldd, double* U, long ldu, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_dd_from_upper, grid_nn(n), dim3(256), 0, s, n, D, ldd, U, ldu);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_w_from_g(int n, const double* G, double* W, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_dd_w_from_g, grid_nn(n), dim3(256), 0, s, n, G, W);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_add_to(int n, const double* F, long ldf, double* U, long ldu, int zero_lower, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_dd_add_to, grid_nn(n), dim3(256), 0, s, n, F, ldf, U, ldu, zero_lower);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_symmetric(int n, const double* A, long lda, int* flag, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_dd_symmetric, grid_nn(n), dim3(256), 0, s, n, A, lda, flag);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_from_d(int m, int n, const double* D, long ldd, double* X, long ldx, hipStream_t s) {
  if ((long)m * n == 0) return 0;
  hipLaunchKernelGGL(k_dd_from_d, dim3((unsigned)(((long)m * n + 255) / 256)), dim3(256), 0, s, m, n, D, ldd, X, ldx);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// one Ogita-Aishima step's elementwise part: lam (2n), rows (2n) and delta (1)
// are work arrays; E (n x n dd, leading dimension n) receives the correction
int vcphip_oa_elementwise(int n, const double* S, long lds, const double* G, long ldg, double* lam, double* rows,
                          double* delta, double* E, hipStream_t s) {
  if (n <= 0) return 0;
  const dim3 g1((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_oa_lam, g1, dim3(256), 0, s, n, S, lds, G, ldg, lam);
  hipLaunchKernelGGL(k_oa_rows, g1, dim3(256), 0, s, n, S, lds, G, ldg, rows);
  hipLaunchKernelGGL(k_oa_delta, dim3(1), dim3(256), 0, s, n, rows, lam, delta);
  hipLaunchKernelGGL(k_oa_e, grid_nn(n), dim3(256), 0, s, n, S, lds, G, ldg, lam, delta, E);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// part holds 2 * VCPHIP_DD_ADD_BLOCKS doubles; out[0] = max |d|, out[1] = max |x.hi|
int vcphip_dd_add_d(int m, int n, const double* d, double* X, long ldx, double* part, double* out, hipStream_t s) {
  const long total = (long)m * n;
  long nb = (total + 255) / 256;
  if (nb > VCPHIP_DD_ADD_BLOCKS) nb = VCPHIP_DD_ADD_BLOCKS;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_dd_add_d, dim3((unsigned)nb), dim3(256), 0, s, m, n, d, X, ldx, part);
  hipLaunchKernelGGL(k_max_pairs, dim3(1), dim3(256), 0, s, (int)nb, part, out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_stats(const double* X, long rs, long cs, int rows, int len, int* e, int* flags, int* part,
                    hipStream_t s) {
  if (rows <= 0) return 0;
  if (cs == 1) {   // contiguous lines: one wave each
    hipLaunchKernelGGL(k_dd_stats_lines, dim3((rows + 3) / 4), dim3(256), 0, s, X, rs, rows, len, e, flags);
    return hipGetLastError() == hipSuccess ? 0 : -3;
  }
  const int nchunk = len > 0 ? (len + VCPHIP_DD_CHUNK - 1) / VCPHIP_DD_CHUNK : 0;
  if (nchunk > 0)
    hipLaunchKernelGGL(k_dd_stats_part, dim3((rows + 255) / 256, nchunk), dim3(256), 0, s, X, rs, cs, rows, len,
                       nchunk, part);
  hipLaunchKernelGGL(k_dd_stats_final, dim3((rows + 255) / 256), dim3(256), 0, s, rows, nchunk, part, e, flags);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_split_rows(const double* A, long lda, int m, int k, int mpad, int kpad, const VcphipOzParams& P,
                         const int* e, const int* flags, int8_t* planes, hipStream_t s) {
  const dim3 grid(mpad / SR_ROWS, kpad / SR_K);
  switch (P.nmod) {
#define VCPHIP_DD_SPLIT(NN) \
  case NN: hipLaunchKernelGGL((k_dd_split_rows<NN>), grid, dim3(256), 0, s, A, lda, m, k, mpad, kpad, P, e, flags, planes); break;
    VCPHIP_DD_SPLIT(24) VCPHIP_DD_SPLIT(25) VCPHIP_DD_SPLIT(26) VCPHIP_DD_SPLIT(27) VCPHIP_DD_SPLIT(28)
    VCPHIP_DD_SPLIT(29) VCPHIP_DD_SPLIT(30) VCPHIP_DD_SPLIT(31) VCPHIP_DD_SPLIT(32)
#undef VCPHIP_DD_SPLIT
    default: return -7;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_split_cols(const double* B, long ldb, int k, int n, int npad, int kpad, const VcphipOzParams& P,
                         const int* e, const int* flags, int8_t* planes, hipStream_t s) {
  const long blocks = (long)(npad / 16) * (kpad / 64);
  if (blocks > 0x7FFFFFFFL) return -7;
  switch (P.nmod) {
#define VCPHIP_DD_SPLIT(NN) \
  case NN: hipLaunchKernelGGL((k_dd_split_cols<NN>), dim3((unsigned)blocks), dim3(256), 0, s, B, ldb, k, n, npad, kpad, P, e, flags, planes); break;
    VCPHIP_DD_SPLIT(24) VCPHIP_DD_SPLIT(25) VCPHIP_DD_SPLIT(26) VCPHIP_DD_SPLIT(27) VCPHIP_DD_SPLIT(28)
    VCPHIP_DD_SPLIT(29) VCPHIP_DD_SPLIT(30) VCPHIP_DD_SPLIT(31) VCPHIP_DD_SPLIT(32)
#undef VCPHIP_DD_SPLIT
    default: return -7;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_recon(int m, int n, int mpad, int npad, const VcphipOzParams& P, const int* ea, const int* fa,
                    const int* eb, const int* fb, const uint8_t* U, double* C, long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid((unsigned)(mpad / 16), (unsigned)(npad / 16));
  switch (P.nmod) {
#define VCPHIP_DD_CASE(NN) \
  case NN: hipLaunchKernelGGL((k_dd_recon<NN>), grid, dim3(256), 0, s, m, n, mpad, npad, P, ea, fa, eb, fb, U, C, ldc); break;
    VCPHIP_DD_CASE(24) VCPHIP_DD_CASE(25) VCPHIP_DD_CASE(26) VCPHIP_DD_CASE(27) VCPHIP_DD_CASE(28)
    VCPHIP_DD_CASE(29) VCPHIP_DD_CASE(30) VCPHIP_DD_CASE(31) VCPHIP_DD_CASE(32)
#undef VCPHIP_DD_CASE
    default: return -7;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_dd_sub_from(int m, int n, const double* B, long ldb, double* R, long ldr, hipStream_t s) {
  if ((long)m * n == 0) return 0;
  hipLaunchKernelGGL(k_dd_sub_from, dim3((unsigned)(((long)m * n + 255) / 256)), dim3(256), 0, s, m, n, B, ldb, R, ldr);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
