// Ozaki scheme II interval GEMM on int8 MFMA (gfx950).
//
// The verified product IA*IB (the pidblas::mulmm result, vcp/pidblas.hpp:
// 239-294) is formed from
//   * the midpoint product mid(A) mid(B), computed EXACTLY on scaled
//     integers:  A' = RN(mid A_ik 2^(beta - e_i)),  B' = RN(mid B_kj 2^(beta - f_j))
//     with |A'|, |B'| <= 2^beta.  A'B' is evaluated modulo N pairwise coprime
//     moduli m_t <= 256 with one int8 MFMA GEMM per modulus (residues in
//     [-128, 127], exact int32 accumulation for k <= 2^17), and rebuilt by the
//     Chinese remainder theorem (mixed-radix / Garner form) into the exact
//     integer, which is then rounded once down and once up to double;
//   * a rigorous bound of the scaling error
//       E_ij = 2^(f_j-beta-1) (||mid A_i||_1 + k 2^(e_i-beta-1)) + 2^(e_i-beta-1) ||mid B_j||_1;
//   * the pidblas radius term (|mid A| + rad A) rad B + rad A |mid B|
//     (:277-283) bounded from above by two more int8 MFMA GEMMs of upward-
//     rounded 7-bit digits (exact integer products, scales 2^g / 127).
// Output: [RD(lo_mid - (Rt + E)), RU(hi_mid + (Rt + E))], all in round-up
// mode with RD(x) = -RU(-x).  The CPU oracle restates each step and the GPU
// result is bit-identical to it.
#include "vcphip_device.h"
#include "vcphip_kernels.h"
#include "vcphip_crt.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace {

// ---------------------------------------------------------------------------
// statistics: per row (or column) of a midpoint-radius operand
// ---------------------------------------------------------------------------
// NONFINITE / HASNAN bits of one interval [lo, hi]
VCPHIP_DEVINL int nonfinite_bits(double lo, double hi) {
  const bool nan = __builtin_isnan(lo) || __builtin_isnan(hi);
  const bool inf = __builtin_isinf(lo) || __builtin_isinf(hi);
  return (nan || inf ? VCPHIP_OZ_NONFINITE : 0) | (nan ? VCPHIP_OZ_HASNAN : 0);
}

// flags of a row / column from its maxima and non-finite bits
VCPHIP_DEVINL int side_flags(double mu, double smax, double rmax, int nf) {
  if (nf) return nf | VCPHIP_OZ_MIDZERO | VCPHIP_OZ_SUMZERO | VCPHIP_OZ_RADZERO;
  return (mu == 0.0 ? VCPHIP_OZ_MIDZERO : 0) | (smax == 0.0 ? VCPHIP_OZ_SUMZERO : 0) |
         (rmax == 0.0 ? VCPHIP_OZ_RADZERO : 0);
}

__global__ __launch_bounds__(256) void k_oz_stats_part(const double* __restrict__ X, long rs, long cs,
                                                       int rows, int len, int nchunk,
                                                       double* __restrict__ part) {
  vcphip_mode_round_up();
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= rows) return;
  const int c0 = c * VCPHIP_OZ_CHUNK;
  const int c1 = min(c0 + VCPHIP_OZ_CHUNK, len);
  double mu = 0.0, smax = 0.0, rmax = 0.0, sum = 0.0;
  int nf = 0;
  const double* base = X + 2 * ((long)r * rs);
#pragma unroll 8
  for (int cc = c0; cc < c1; ++cc) {
    const double2 v = *reinterpret_cast<const double2*>(base + 2 * ((long)cc * cs));
    nf |= nonfinite_bits(v.x, v.y);
    double m, rr;
    vcphip_midrad_ru(v.x, v.y, m, rr);
    const double am = __builtin_fabs(m);
    const double as = am + rr;
    mu = am > mu ? am : mu;
    smax = as > smax ? as : smax;
    rmax = rr > rmax ? rr : rmax;
    sum = sum + am;
  }
  const long stride = (long)nchunk * rows;
  part[0 * stride + (long)c * rows + r] = mu;
  part[1 * stride + (long)c * rows + r] = smax;
  part[2 * stride + (long)c * rows + r] = rmax;
  part[3 * stride + (long)c * rows + r] = sum;
  part[4 * stride + (long)c * rows + r] = (double)nf;
}

// columns of B (k x n, column k-contiguous): one wave per column.  Lane l
// accumulates the elements l, l+64, l+128, ... in ascending order (coalesced
// loads); the upward 1-norm is then the sequential upward sum of the 64 lane
// partials in lane order.  Maxima are order-independent.
VCPHIP_DEVINL double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double y = __shfl_xor(x, o, 64);
    x = y > x ? y : x;
  }
  return x;
}

__global__ __launch_bounds__(256) void k_oz_stats_cols(const double* __restrict__ X, long ldx, int cols, int len,
                                                       VcphipOzSide side) {
  vcphip_mode_round_up();
  __shared__ double s_sum[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  double mu = 0.0, smax = 0.0, rmax = 0.0, sum = 0.0;
  int nf = 0;
  if (j < cols) {
    const double* col = X + 2 * ldx * (long)j;
#pragma unroll 4
    for (int kk = lane; kk < len; kk += 64) {
      const double2 v = *reinterpret_cast<const double2*>(col + 2 * (long)kk);
      nf |= nonfinite_bits(v.x, v.y);
      double m, rr;
      vcphip_midrad_ru(v.x, v.y, m, rr);
      const double am = __builtin_fabs(m);
      const double as = am + rr;
      mu = am > mu ? am : mu;
      smax = as > smax ? as : smax;
      rmax = rr > rmax ? rr : rmax;
      sum = sum + am;
    }
  }
  mu = wave_max(mu);
  smax = wave_max(smax);
  rmax = wave_max(rmax);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nf |= __shfl_xor(nf, o, 64);
  s_sum[w][lane] = sum;
  __syncthreads();
  if (lane == 0 && j < cols) {
    double tot = 0.0;
    for (int l = 0; l < 64; ++l) tot = tot + s_sum[w][l];
    side.flags[j] = side_flags(mu, smax, rmax, nf);
    side.e[j] = (!nf && mu > 0.0) ? vcphip_fexp(mu) : 0;
    side.g[j] = (!nf && smax > 0.0) ? vcphip_fexp(smax) : 0;
    side.h[j] = (!nf && rmax > 0.0) ? vcphip_fexp(rmax) : 0;
    side.n1[j] = nf ? 0.0 : tot;
  }
}

__global__ __launch_bounds__(256) void k_oz_stats_final(int rows, int nchunk, const double* __restrict__ part,
                                                        VcphipOzSide side) {
  vcphip_mode_round_up();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  const long stride = (long)nchunk * rows;
  double mu = 0.0, smax = 0.0, rmax = 0.0, tot = 0.0;
  int nf = 0;
  for (int c = 0; c < nchunk; ++c) {
    const double a = part[0 * stride + (long)c * rows + r];
    const double b = part[1 * stride + (long)c * rows + r];
    const double d = part[2 * stride + (long)c * rows + r];
    const double s = part[3 * stride + (long)c * rows + r];
    nf |= (int)part[4 * stride + (long)c * rows + r];
    mu = a > mu ? a : mu;
    smax = b > smax ? b : smax;
    rmax = d > rmax ? d : rmax;
    tot = tot + s;
  }
  side.flags[r] = side_flags(mu, smax, rmax, nf);
  side.e[r] = (!nf && mu > 0.0) ? vcphip_fexp(mu) : 0;
  side.g[r] = (!nf && smax > 0.0) ? vcphip_fexp(smax) : 0;
  side.h[r] = (!nf && rmax > 0.0) ? vcphip_fexp(rmax) : 0;
  side.n1[r] = nf ? 0.0 : tot;
}

// ---------------------------------------------------------------------------
// split: scaled integers -> symmetric residues, upward radius digits
// ---------------------------------------------------------------------------
// nearest integer (ties up) of x without relying on the rounding mode
VCPHIP_DEVINL long long rnint_floor(double x) {
  const double f = __builtin_floor(x);
  const double r = x - f;   // exact
  return (long long)f + (r >= 0.5 ? 1 : 0);
}

// two upward 7-bit digits of x >= 0 against 2^g (x < 2^g), round-up mode:
//   t = RU(127 x 2^-g), h = floor(t), l = ceil(RU(127 (t - h)))
// so that x <= (h + l/127) 2^g / 127 with h, l in [0, 127].
VCPHIP_DEVINL unsigned digits2_ru(double x, int g) {
  const double t = vcphip_mulpow2(x, -g) * 127.0;
  const double hf = __builtin_floor(t);
  const double r = (t - hf) * 127.0;   // t - hf is exact
  return (unsigned)hf | ((unsigned)__builtin_ceil(r) << 8);
}

// the same bound with a non-positive low digit (B's factors):
//   x <= (H + L/127) 2^g / 127,  H = h + [l > 0] <= 127,  L = l - 127 [l > 0] in [-126, 0]
// (h = 127 only when t = 127, and then l = 0).  Against A's (h1, l1 >= 0)
// the product of the two bounds is h1 H + (h1 L + l1 H)/127 + l1 L/127^2,
// whose last term is never positive: the digit GEMMs leave it out.
VCPHIP_DEVINL unsigned digits2_ru_neglo(double x, int g) {
  const unsigned d = digits2_ru(x, g);
  const unsigned h = d & 0xffu, l = d >> 8;
  return l ? ((h + 1u) | (((l - 127u) & 0xffu) << 8)) : d;
}

// operand planes of GEMM output z.  Residue z < nmod: A plane z x B plane z.
// Radius-digit output r = z - nmod: the passes of P's table (rnp[r] of them,
// summed into one int32 output), digit planes counted after the residues.
VCPHIP_DEVINL void oz_pass_planes(const VcphipOzParams& P, int z, int nmod, int& npass, int& pa0, int& pb0, int& pa1,
                                  int& pb1) {
  npass = 1;
  pa0 = pb0 = pa1 = pb1 = z;
  if (z < nmod) return;
  const int r = z - nmod;
  npass = P.rnp[r];
  pa0 = nmod + P.rpa[r][0];
  pb0 = nmod + P.rpb[r][0];
  pa1 = nmod + P.rpa[r][1];
  pb1 = nmod + P.rpb[r][1];
}

// x (|x| <= 2^62) as the eight bytes of x + 2^63 (its two's complement with
// the sign bit flipped), held as two byte quads, low byte first
struct Bytes8 {
  unsigned lo, hi;
};
VCPHIP_DEVINL Bytes8 bytes8(long long x) {
  const unsigned long long u = (unsigned long long)x;
  return Bytes8{(unsigned)u, (unsigned)(u >> 32) ^ 0x80000000u};
}

// the residues of the four elements' x in [-floor(m/2), m - 1 - floor(m/2)]
// (the symmetric int8 range; for m = 256 that is [-128, 127]) as the bytes of
// one word.  y = x (mod m) from two byte dot products of x + 2^63 with the
// weights 256^i mod m, started at the bias ((floor(m/2) - 2^63) mod m) -
// floor(m/2), so y + floor(m/2) lies in [0, 2^19); q = floor((y +
// floor(m/2)) / m) is exact (vcphip_quot24) and c = y - q m is the residue.
VCPHIP_DEVINL unsigned sym_residue_word(const Bytes8 (&b)[4], int t, const VcphipOzParams& P) {
  const unsigned m = (unsigned)P.mod[t] & 0x1FFu, h8 = 8u * (m >> 1);
  unsigned res[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    unsigned y = __builtin_amdgcn_udot4(b[u].lo, P.r_wb[0][t], P.r_bias[t], false);
    y = __builtin_amdgcn_udot4(b[u].hi, P.r_wb[1][t], y, false);
    res[u] = vcphip_mad_i24(vcphip_quot24((y << 3) + h8, P.r_magic[t]), -(int)m, y);
  }
  return vcphip_pack4(res);
}

constexpr int SR_ROWS = 16, SR_K = 64, SR_LD = 68;   // LDS byte tile (one plane block), odd-word row stride

// rows of A (m x k, column-major intervals) -> N+6 planes of mpad x kpad:
// N residue planes of A', then the hi/lo digits of |mid|+rad and of rad, then
// their rounded-up single digits H = hi + [lo > 0].
template <int N>
__global__ __launch_bounds__(256) void k_oz_split_rows(const double* __restrict__ A, long lda, int m, int k,
                                                       int mpad, int kpad, VcphipOzParams P, VcphipOzSide side,
                                                       int8_t* __restrict__ planes) {
  vcphip_mode_round_up();
  extern __shared__ __attribute__((aligned(16))) unsigned char s_tile[];   // [N+6][SR_ROWS][SR_LD]
  constexpr int nplane = N + VCPHIP_OZ_NDIGIT_A;
  const int row0 = blockIdx.x * SR_ROWS, k0 = blockIdx.y * SR_K;
  const int tid = threadIdx.x;
  // each thread: one row, quads of 4 consecutive k (one 32-bit LDS word per plane)
  constexpr int NQ = 256 / SR_ROWS;   // quads per pass
  const int lr = tid % SR_ROWS, kq4 = tid / SR_ROWS;
  const int r = row0 + lr;
  const bool rowok = r < m && !(side.flags[r < m ? r : 0] & VCPHIP_OZ_NONFINITE);
  const int e = rowok ? side.e[r] : 0;
  const int g = rowok ? side.g[r] : 0;
  const int h = rowok ? side.h[r] : 0;
  const bool zr = rowok ? (side.flags[r] & VCPHIP_OZ_MIDZERO) != 0 : true;
  for (int pass = 0; pass < SR_K / (4 * NQ); ++pass) {
    const int lk0 = 4 * (pass * NQ + kq4);
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kk = k0 + lk0 + u;
      v[u] = (rowok && kk < k) ? *reinterpret_cast<const double2*>(A + 2 * (r + lda * kk)) : make_double2(0.0, 0.0);
    }
    Bytes8 ch[4];
    unsigned dsh = 0, dsl = 0, drh = 0, drl = 0, dsc = 0, drc = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double am, ar;
      vcphip_midrad_ru(v[u].x, v[u].y, am, ar);
      const double as = __builtin_fabs(am) + ar;
      const long long ap = zr ? 0 : rnint_floor(vcphip_mulpow2(am, P.beta - e));
      ch[u] = bytes8(ap);
      const unsigned ds = digits2_ru(as, g), dr = digits2_ru(ar, h);
      dsh |= (ds & 0xff) << (8 * u);
      dsl |= (ds >> 8) << (8 * u);
      drh |= (dr & 0xff) << (8 * u);
      drl |= (dr >> 8) << (8 * u);
      dsc |= ((ds & 0xff) + ((ds >> 8) ? 1u : 0u)) << (8 * u);
      drc |= ((dr & 0xff) + ((dr >> 8) ? 1u : 0u)) << (8 * u);
    }
    // modulus by modulus: only one modulus' constants are live at a time
#pragma unroll
    for (int t = 0; t < N; ++t)
      *reinterpret_cast<unsigned*>(&s_tile[(t * SR_ROWS + lr) * SR_LD + lk0]) = sym_residue_word(ch, t, P);
    *reinterpret_cast<unsigned*>(&s_tile[(N * SR_ROWS + lr) * SR_LD + lk0]) = dsh;
    *reinterpret_cast<unsigned*>(&s_tile[((N + 1) * SR_ROWS + lr) * SR_LD + lk0]) = dsl;
    *reinterpret_cast<unsigned*>(&s_tile[((N + 2) * SR_ROWS + lr) * SR_LD + lk0]) = drh;
    *reinterpret_cast<unsigned*>(&s_tile[((N + 3) * SR_ROWS + lr) * SR_LD + lk0]) = drl;
    *reinterpret_cast<unsigned*>(&s_tile[((N + 4) * SR_ROWS + lr) * SR_LD + lk0]) = dsc;
    *reinterpret_cast<unsigned*>(&s_tile[((N + 5) * SR_ROWS + lr) * SR_LD + lk0]) = drc;
  }
  __syncthreads();
  // write out: per plane one 1 KB block (16 rows x 64 k), one word per thread
  static_assert(SR_ROWS * SR_K == 4 * 256, "one word per thread and plane");
  const int orow = tid / 16, oc = (tid % 16) * 4;
  const long plane_stride = (long)mpad * kpad;
  int8_t* dst = planes + vcphip_plane_off(row0 + orow, k0 + oc, kpad);
#pragma unroll 4
  for (int p = 0; p < nplane; ++p)
    if (p < N || ((P.dmask_a >> (p - N)) & 1))   // digit planes the radius scheme reads
      *reinterpret_cast<unsigned*>(dst + p * plane_stride) =
          *reinterpret_cast<const unsigned*>(s_tile + (p * SR_ROWS + orow) * SR_LD + oc);
}

// columns of B (k x n, column-major intervals) -> N+4 planes of npad x kpad:
// N residue planes of B', then the hi/lo digits of rad and of |mid|.
// Thread: a quad of 4 consecutive k of one column; 16 consecutive threads
// cover 64 k of a column (1 KB of contiguous input), a workgroup one 1 KB
// plane block (16 columns x 64 k) per plane, written as 4 contiguous
// 256-byte runs.
template <int N>
__global__ __launch_bounds__(256) void k_oz_split_cols(const double* __restrict__ B, long ldb, int k, int n,
                                                       int npad, int kpad, VcphipOzParams P, VcphipOzSide side,
                                                       int8_t* __restrict__ planes) {
  vcphip_mode_round_up();
  const int kblk = kpad / 64, unit = blockIdx.x, t = threadIdx.x;
  const int j = (unit / kblk) * 16 + (t >> 4), k0 = (unit % kblk) * 64 + 4 * (t & 15);
  const bool colok = j < n && !(side.flags[j < n ? j : 0] & VCPHIP_OZ_NONFINITE);
  const int f = colok ? side.e[j] : 0;
  const int xi = colok ? side.h[j] : 0;
  const bool zc = colok ? (side.flags[j] & VCPHIP_OZ_MIDZERO) != 0 : true;
  const int zeta = zc ? 0 : f;
  double2 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int kk = k0 + u;
    v[u] = (colok && kk < k) ? *reinterpret_cast<const double2*>(B + 2 * (kk + ldb * (long)j)) : make_double2(0.0, 0.0);
  }
  Bytes8 ch[4];
  unsigned drh = 0, drl = 0, dmh = 0, dml = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    double bm, br;
    vcphip_midrad_ru(v[u].x, v[u].y, bm, br);
    const long long bp = zc ? 0 : rnint_floor(vcphip_mulpow2(bm, P.beta - f));
    ch[u] = bytes8(bp);
    const unsigned d1 = digits2_ru_neglo(br, xi);
    const unsigned d2 = digits2_ru_neglo(__builtin_fabs(bm), zeta);
    drh |= (d1 & 0xff) << (8 * u);
    drl |= (d1 >> 8) << (8 * u);
    dmh |= (d2 & 0xff) << (8 * u);
    dml |= (d2 >> 8) << (8 * u);
  }
  const long plane_stride = (long)npad * kpad;
  int8_t* dst = planes + vcphip_plane_off(j, k0, kpad);
#pragma unroll
  for (int z = 0; z < N; ++z) *reinterpret_cast<unsigned*>(dst + z * plane_stride) = sym_residue_word(ch, z, P);
  // the digit planes the radius scheme reads (the low digits are unused by
  // the one-digit schemes)
  if (P.dmask_b & 1) *reinterpret_cast<unsigned*>(dst + N * plane_stride) = drh;
  if (P.dmask_b & 2) *reinterpret_cast<unsigned*>(dst + (N + 1) * plane_stride) = drl;
  if (P.dmask_b & 4) *reinterpret_cast<unsigned*>(dst + (N + 2) * plane_stride) = dmh;
  if (P.dmask_b & 8) *reinterpret_cast<unsigned*>(dst + (N + 3) * plane_stride) = dml;
}

// ---------------------------------------------------------------------------
// int8 MFMA GEMM over the planes: C_z = A_z * B_z^T (both K-contiguous)
// ---------------------------------------------------------------------------
// Plane layout (both operands, written by the split kernels): 1 KB blocks of
// 16 rows x 64 k, row-major inside the block, blocks ordered by row group and
// then by k (vcphip_plane_off), so that one wave's load of 16 rows x 64 k is
// one contiguous kilobyte.
// Output layout is blocked by 16x16 MFMA tiles: block (bi, bj) occupies
// 256 consecutive elements, element (4*(l/16)+v, l%16) of the block at l*4+v,
// which is exactly the v_mfma_i32_16x16x64_i8 accumulator order, so every
// wave stores whole contiguous blocks.
// Workgroup: 4 waves as 2 x 2.

// Variant with direct global->LDS loads (global_load_lds_dwordx4): three LDS
// stages, two in flight while the MFMAs run on the third.  Rows of a stage are
// 64 bytes; the 16-byte chunks of a row are XOR-swizzled by (row >> 2) & 3 so
// the 16-byte fragment reads are bank-conflict free.  Each wave fills whole
// 1 KB segments (16 rows): lane l writes segment byte 16 l, i.e. row l / 4,
// physical chunk l % 4, which holds logical chunk (l % 4) ^ ((l / 16) % 4).
// The residue-plane products reduced mod m into [0, m), four to a packed
// word.  For kpad <= 8192 (|acc| <= kpad 2^14 <= 2^27) in integer
// arithmetic: t = acc + K 2^16 >= 0 (K = ceil(2^11 / m) m, so t < 2^29 and
// t >> 16 < 4352) is congruent to z = t + (t >> 16)(c16 - 2^16) =
// (t >> 16) c16 + (t mod 2^16) < 2^21 with c16 = 2^16 mod m (one
// v_mad_i32_i24), and z mod m = z - q m with q the exact 24-bit quotient:
// six VALU instructions per product.  Larger k: a double-precision floor
// division (by an approximate reciprocal) with two fix-ups.  `fast` is uniform, so each launch takes one
// branch.
struct OzModPlane {
  bool fast;
  int m, c16m;
  unsigned kbias, magic;
  double rc;
  VCPHIP_DEVINL OzModPlane(const VcphipOzParams& P, int z, int kpad)
      : fast(kpad <= 8192), m(P.mod[z]), c16m(P.ep_c16[z]), kbias(P.ep_bias[z]), magic(P.r_magic[z]),
        rc(fast ? 0.0 : 1.0 / (double)P.mod[z]) {}
  VCPHIP_DEVINL unsigned pack(int a0, int a1, int a2, int a3) const {
    const int a[4] = {a0, a1, a2, a3};
    unsigned r[4];
    if (fast) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const unsigned t = (unsigned)a[v] + kbias;
        const unsigned zz = vcphip_mad_i24(t >> 16, c16m, t);
        r[v] = vcphip_mad_i24(vcphip_quot24(zz << 3, magic), -m, zz);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double y = (double)a[v];
        const double qd = __builtin_floor(y * rc);
        int rr = (int)(y - qd * (double)m);
        rr += (rr < 0) ? m : 0;
        rr -= (rr >= m) ? m : 0;
        r[v] = (unsigned)rr;
      }
    }
    return vcphip_pack4(r);
  }
};

template <int FM, int FN>
__global__ __launch_bounds__(256) void k_oz_gemm3(int mpad, int npad, int kpad, int nmod, int nz,
                                                  const int8_t* __restrict__ Aplanes,
                                                  const int8_t* __restrict__ Bplanes, VcphipOzParams P,
                                                  uint8_t* __restrict__ U, int32_t* __restrict__ Z) {
  constexpr int WM = 16 * FM, WN = 16 * FN;
  constexpr int TMB = 2 * WM, TNB = 2 * WN;
  constexpr int NST = 3;                  // LDS stages
  constexpr int SEGA = TMB / 16, SEGB = TNB / 16;   // 1 KB segments per stage
  constexpr int SEGA_W = SEGA / 4, SEGB_W = SEGB / 4;   // per wave
  static_assert(SEGA_W * 4 == SEGA && SEGB_W * 4 == SEGB, "tile/wave mismatch");
  __shared__ __attribute__((aligned(16))) int8_t sA[NST][TMB * 64];
  __shared__ __attribute__((aligned(16))) int8_t sB[NST][TNB * 64];

  const int ntm = mpad / TMB, ntn = npad / TNB;
  const int tiles = ntm * ntn;
  const int z = blockIdx.x / tiles;
  const int local = vcphip_xcd_remap(blockIdx.x - z * tiles, tiles);
  int tm, tn;
  vcphip_tile_coords(local, ntm, ntn, 8, tm, tn);
  const int row0 = tm * TMB, col0 = tn * TNB;
  int npass, pa0, pb0, pa1, pb1;
  oz_pass_planes(P, z, nmod, npass, pa0, pb0, pa1, pb1);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int wu = __builtin_amdgcn_readfirstlane(w);
  const uint32_t sA_base = vcphip_lds_addr(&sA[0][0]), sB_base = vcphip_lds_addr(&sB[0][0]);
  // this lane's source in every 1 KB segment: row lane/4, logical chunk
  const int lrow = lane >> 2, lchunk = (lane & 3) ^ ((lane >> 4) & 3);
  const int src_lane = lrow * 64 + 16 * lchunk;   // inside the 1 KB block of 16 rows x 64 k
  const long kblk = kpad / 64;
  // fragment reads: row R, logical chunk g = lane / 16 -> physical g ^ ((R >> 2) & 3)
  const int frow = lane & 15, g = lane >> 4;
  int fa_off[FM], fb_off[FN];
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int R = wm * WM + 16 * i + frow;
    fa_off[i] = R * 64 + 16 * (g ^ ((R >> 2) & 3));
  }
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int R = wn * WN + 16 * j + frow;
    fb_off[j] = R * 64 + 16 * (g ^ ((R >> 2) & 3));
  }

  v4i acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4i{0, 0, 0, 0};

  const int nk = kpad / 64;
  for (int pass = 0; pass < npass; ++pass) {
    const int8_t* Ab = Aplanes + (long)(pass ? pa1 : pa0) * mpad * kpad + (long)(row0 / 16) * kblk * 1024 + src_lane;
    const int8_t* Bb = Bplanes + (long)(pass ? pb1 : pb0) * npad * kpad + (long)(col0 / 16) * kblk * 1024 + src_lane;
    // The LDS-DMA loads, their wait and the barrier are inline asm, so the
    // compiler neither tracks the DMA (it would wait for all of it before
    // every LDS read) nor fences the barrier; the "memory" clobbers keep the
    // LDS reads between them.
    auto issue = [&](int stage, int kt) {
#pragma unroll
      for (int q = 0; q < SEGA_W; ++q) {
        const int seg = wu * SEGA_W + q;
        vcphip_lds_dma16(Ab + ((long)seg * kblk + kt) * 1024, sA_base + stage * (TMB * 64) + seg * 1024);
      }
#pragma unroll
      for (int q = 0; q < SEGB_W; ++q) {
        const int seg = wu * SEGB_W + q;
        vcphip_lds_dma16(Bb + ((long)seg * kblk + kt) * 1024, sB_base + stage * (TNB * 64) + seg * 1024);
      }
    };
    issue(0, 0);
    if (nk > 1) issue(1, 1);
    for (int kt = 0; kt < nk; ++kt) {
      // stage kt's loads are complete once only stage kt+1's remain in flight
      if (kt + 1 < nk) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(SEGA_W + SEGB_W) : "memory");
      else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      if (kt + 2 < nk) issue((kt + 2) % NST, kt + 2);
      const int8_t* la = sA[kt % NST];
      const int8_t* lb = sB[kt % NST];
      v4i fa[FM];
#pragma unroll
      for (int i = 0; i < FM; ++i) fa[i] = *reinterpret_cast<const v4i*>(la + fa_off[i]);
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const v4i fb = *reinterpret_cast<const v4i*>(lb + fb_off[j]);
#pragma unroll
        for (int i = 0; i < FM; ++i) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb, acc[i][j], 0, 0, 0);
      }
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }

  const long blocks_m = mpad / 16;
  if (z < nmod) {
    const OzModPlane mod(P, z, kpad);
    uint8_t* Uz = U + (long)z * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const long bi = (row0 + wm * WM + 16 * i) / 16, bj = (col0 + wn * WN + 16 * j) / 16;
        const unsigned packed = mod.pack(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
        *reinterpret_cast<unsigned*>(Uz + (bj * blocks_m + bi) * 256 + lane * 4) = packed;
      }
  } else {
    int32_t* Zz = Z + (long)(z - nmod) * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const long bi = (row0 + wm * WM + 16 * i) / 16, bj = (col0 + wn * WN + 16 * j) / 16;
        *reinterpret_cast<v4i*>(Zz + (bj * blocks_m + bi) * 256 + lane * 4) = acc[i][j];
      }
  }
}

// Deeper pipeline: NST LDS stages and fragments double-buffered in registers.
// Iteration kt waits for stage kt+1, refills stage kt (its fragments are
// already in registers) with k-step kt+NST, reads the fragments of kt+1 and
// runs the MFMAs of kt.  The swizzle term (R >> 2) & 3 of a fragment row
// R = 16 i + frow (+ a multiple of 16) is (frow >> 2) & 3 for every i, so one
// base offset per operand serves all fragments.
template <int FM, int FN>
struct OzFrag {
  v4i a[FM];
  v4i b[FN];
};

template <int FM, int FN>
__device__ __forceinline__ void oz_load_frag(const int8_t* la, const int8_t* lb, OzFrag<FM, FN>& f) {
#pragma unroll
  for (int i = 0; i < FM; ++i) f.a[i] = *reinterpret_cast<const v4i*>(la + i * 1024);
#pragma unroll
  for (int j = 0; j < FN; ++j) f.b[j] = *reinterpret_cast<const v4i*>(lb + j * 1024);
}

template <int FM, int FN>
__device__ __forceinline__ void oz_mfma_frag(const OzFrag<FM, FN>& f, v4i (&acc)[FM][FN]) {
#pragma unroll
  for (int j = 0; j < FN; ++j)
#pragma unroll
    for (int i = 0; i < FM; ++i) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
}

template <int FM, int FN, int NST, bool IL = false>
__global__ __launch_bounds__(256) void k_oz_gemm4(int mpad, int npad, int kpad, int nmod, int nz,
                                                  const int8_t* __restrict__ Aplanes,
                                                  const int8_t* __restrict__ Bplanes, VcphipOzParams P,
                                                  uint8_t* __restrict__ U, int32_t* __restrict__ Z) {
  constexpr int WM = 16 * FM, WN = 16 * FN;
  constexpr int TMB = 2 * WM, TNB = 2 * WN;
  constexpr int SEGA = TMB / 16, SEGB = TNB / 16;
  constexpr int SEGA_W = SEGA / 4, SEGB_W = SEGB / 4;
  constexpr int G = SEGA_W + SEGB_W;   // loads per wave per stage
  static_assert(SEGA_W * 4 == SEGA && SEGB_W * 4 == SEGB, "tile/wave mismatch");
  static_assert(NST >= 3 && (NST - 1) * G < 64, "vmcnt range");
  __shared__ __attribute__((aligned(16))) int8_t sA[NST][TMB * 64];
  __shared__ __attribute__((aligned(16))) int8_t sB[NST][TNB * 64];

  const int ntm = mpad / TMB, ntn = npad / TNB;
  const int tiles = ntm * ntn;
  const int z = blockIdx.x / tiles;
  const int local = vcphip_xcd_remap(blockIdx.x - z * tiles, tiles);
  int tm, tn;
  vcphip_tile_coords(local, ntm, ntn, 8, tm, tn);
  const int row0 = tm * TMB, col0 = tn * TNB;
  int npass, pa0, pb0, pa1, pb1;
  oz_pass_planes(P, z, nmod, npass, pa0, pb0, pa1, pb1);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int wu = __builtin_amdgcn_readfirstlane(w);
  const uint32_t sA_base = vcphip_lds_addr(&sA[0][0]), sB_base = vcphip_lds_addr(&sB[0][0]);
  const int lrow = lane >> 2, lchunk = (lane & 3) ^ ((lane >> 4) & 3);
  const int src_lane = lrow * 64 + 16 * lchunk;   // inside the 1 KB block of 16 rows x 64 k
  const long kblk = kpad / 64;
  const int frow = lane & 15, g = lane >> 4;
  const int fsw = 16 * (g ^ ((frow >> 2) & 3));
  const int fa0 = (wm * WM + frow) * 64 + fsw, fb0 = (wn * WN + frow) * 64 + fsw;

  v4i acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4i{0, 0, 0, 0};

  const int nk = kpad / 64;
  OzFrag<FM, FN> f0, f1;
  for (int pass = 0; pass < npass; ++pass) {
    const int8_t* Ab = Aplanes + (long)(pass ? pa1 : pa0) * mpad * kpad + (long)(row0 / 16) * kblk * 1024 + src_lane;
    const int8_t* Bb = Bplanes + (long)(pass ? pb1 : pb0) * npad * kpad + (long)(col0 / 16) * kblk * 1024 + src_lane;
    auto issue = [&](int stage, int kt) {
#pragma unroll
      for (int q = 0; q < SEGA_W; ++q) {
        const int seg = wu * SEGA_W + q;
        vcphip_lds_dma16(Ab + ((long)seg * kblk + kt) * 1024, sA_base + stage * (TMB * 64) + seg * 1024);
      }
#pragma unroll
      for (int q = 0; q < SEGB_W; ++q) {
        const int seg = wu * SEGB_W + q;
        vcphip_lds_dma16(Bb + ((long)seg * kblk + kt) * 1024, sB_base + stage * (TNB * 64) + seg * 1024);
      }
    };
    // Waits until stage kt+1 (or stage 0 when kt = -1) has landed for every
    // wave; the loads issued after it may stay in flight.
    auto wait_next = [&](int kt) {
      if (kt + NST < nk) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((NST - 2) * G) : "memory");
      else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    };
    // one of the G loads of a stage (the first SEGA_W fill A, the rest B)
    auto issue_one = [&](int stage, int kt, int g) {
      if (g < SEGA_W) {
        const int seg = wu * SEGA_W + g;
        vcphip_lds_dma16(Ab + ((long)seg * kblk + kt) * 1024, sA_base + stage * (TMB * 64) + seg * 1024);
      } else {
        const int seg = wu * SEGB_W + (g - SEGA_W);
        vcphip_lds_dma16(Bb + ((long)seg * kblk + kt) * 1024, sB_base + stage * (TNB * 64) + seg * 1024);
      }
    };
    // MFMAs of one stage; with IL the stage's refill loads are spread over
    // them (one per group of FM MFMAs), so that the load queue never drains
    // while the wave issues MFMAs
    auto compute = [&](const OzFrag<FM, FN>& f, int kt) {
      if (!IL) {
        if (kt + NST < nk) issue(kt % NST, kt + NST);
        oz_mfma_frag<FM, FN>(f, acc);
        return;
      }
      const bool refill = kt + NST < nk;
#pragma unroll
      for (int j = 0; j < FN; ++j) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f.a[i], f.b[j], acc[i][j], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        if (refill && (j * G) / FN != ((j + 1) * G) / FN) issue_one(kt % NST, kt + NST, (j * G) / FN);
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    const int npro = nk < NST ? nk : NST;
    for (int st = 0; st < npro; ++st) issue(st, st);
    if (nk >= NST) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((NST - 1) * G) : "memory");
    else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    oz_load_frag<FM, FN>(sA[0] + fa0, sB[0] + fb0, f0);
    for (int kt = 0; kt < nk; kt += 2) {
      if (kt + 1 < nk) {
        wait_next(kt);
        oz_load_frag<FM, FN>(sA[(kt + 1) % NST] + fa0, sB[(kt + 1) % NST] + fb0, f1);
      }
      compute(f0, kt);
      if (kt + 1 < nk) {
        if (kt + 2 < nk) {
          wait_next(kt + 1);
          oz_load_frag<FM, FN>(sA[(kt + 2) % NST] + fa0, sB[(kt + 2) % NST] + fb0, f0);
        }
        compute(f1, kt + 1);
      }
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }

  const long blocks_m = mpad / 16;
  if (z < nmod) {
    const OzModPlane mod(P, z, kpad);
    uint8_t* Uz = U + (long)z * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const long bi = (row0 + wm * WM + 16 * i) / 16, bj = (col0 + wn * WN + 16 * j) / 16;
        const unsigned packed = mod.pack(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
        *reinterpret_cast<unsigned*>(Uz + (bj * blocks_m + bi) * 256 + lane * 4) = packed;
      }
  } else {
    int32_t* Zz = Z + (long)(z - nmod) * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const long bi = (row0 + wm * WM + 16 * i) / 16, bj = (col0 + wn * WN + 16 * j) / 16;
        *reinterpret_cast<v4i*>(Zz + (bj * blocks_m + bi) * 256 + lane * 4) = acc[i][j];
      }
  }
}

// The same pipeline on v_mfma_i32_32x32x32_i8: FM x FN fragments of 32 x 32
// per wave (4 x 4: a 128 x 128 wave tile, 256 accumulator registers).  With
// that many accumulators the 32x32x32 form sustains a far higher issue rate
// than 16x16x64.  Operand lane l holds row (column) l % 32 and the 16 bytes
// of k-quarter 2 h + l / 32 of the stage's k-half h; the accumulator's
// register v of lane l is row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32,
// so registers 4q..4q+3 are one lane-quad of a 16 x 16 output block.
template <int FM, int FN>
struct OzFrag32 {
  v4i a[2][FM];
  v4i b[2][FN];
};

template <int FM, int FN>
__device__ __forceinline__ void oz_load_frag32(const int8_t* la0, const int8_t* la1, const int8_t* lb0,
                                               const int8_t* lb1, OzFrag32<FM, FN>& f) {
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    f.a[0][i] = *reinterpret_cast<const v4i*>(la0 + i * 2048);
    f.a[1][i] = *reinterpret_cast<const v4i*>(la1 + i * 2048);
  }
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    f.b[0][j] = *reinterpret_cast<const v4i*>(lb0 + j * 2048);
    f.b[1][j] = *reinterpret_cast<const v4i*>(lb1 + j * 2048);
  }
}

template <int FM, int FN>
__device__ __forceinline__ void oz_mfma_frag32(const OzFrag32<FM, FN>& f, v16i (&acc)[FM][FN]) {
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int i = 0; i < FM; ++i)
        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.a[h][i], f.b[h][j], acc[i][j], 0, 0, 0);
}

// NW waves per workgroup in a 2 x (NW / 2) grid: with NW = 8 and 4 x 2
// fragments (128 x 64 per wave, 128 accumulator registers) the same 256 x 256
// workgroup tile runs two waves per SIMD, twice the loads in flight per CU.
template <int FM, int FN, int NST, int NW = 4>
__global__ __launch_bounds__(64 * NW) void k_oz_gemm5(int mpad, int npad, int kpad, int nmod, int nz,
                                                      const int8_t* __restrict__ Aplanes,
                                                      const int8_t* __restrict__ Bplanes, VcphipOzParams P,
                                                      uint8_t* __restrict__ U, int32_t* __restrict__ Z) {
  constexpr int WM = 32 * FM, WN = 32 * FN;
  constexpr int TMB = 2 * WM, TNB = (NW / 2) * WN;
  constexpr int SEGA = TMB / 16, SEGB = TNB / 16;
  constexpr int SEGA_W = SEGA / NW, SEGB_W = SEGB / NW;
  constexpr int G = SEGA_W + SEGB_W;
  static_assert(NW == 4 || NW == 8, "wave count");
  static_assert(SEGA_W * NW == SEGA && SEGB_W * NW == SEGB, "tile/wave mismatch");
  static_assert(NST >= 3 && (NST - 1) * G < 64, "vmcnt range");
  __shared__ __attribute__((aligned(16))) int8_t sA[NST][TMB * 64];
  __shared__ __attribute__((aligned(16))) int8_t sB[NST][TNB * 64];

  const int ntm = mpad / TMB, ntn = npad / TNB;
  const int tiles = ntm * ntn;
  const int z = blockIdx.x / tiles;
  const int local = vcphip_xcd_remap(blockIdx.x - z * tiles, tiles);
  int tm, tn;
  vcphip_tile_coords(local, ntm, ntn, 8, tm, tn);
  const int row0 = tm * TMB, col0 = tn * TNB;
  int npass, pa0, pb0, pa1, pb1;
  oz_pass_planes(P, z, nmod, npass, pa0, pb0, pa1, pb1);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int wu = __builtin_amdgcn_readfirstlane(w);
  const uint32_t sA_base = vcphip_lds_addr(&sA[0][0]), sB_base = vcphip_lds_addr(&sB[0][0]);
  const int lrow = lane >> 2, lchunk = (lane & 3) ^ ((lane >> 4) & 3);
  const int src_lane = lrow * 64 + 16 * lchunk;   // inside the 1 KB block of 16 rows x 64 k
  const long kblk = kpad / 64;
  // fragment rows: R = 32 i + lane % 32 (+ a multiple of 32): (R >> 2) & 3 = ((lane % 32) >> 2) & 3
  const int frow = lane & 31, kq = lane >> 5, sw = (frow >> 2) & 3;
  const int fs0 = 16 * ((0 + kq) ^ sw), fs1 = 16 * ((2 + kq) ^ sw);
  const int fa0 = (wm * WM + frow) * 64 + fs0, fa1 = (wm * WM + frow) * 64 + fs1;
  const int fb0 = (wn * WN + frow) * 64 + fs0, fb1 = (wn * WN + frow) * 64 + fs1;

  v16i acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v16i{};

  const int nk = kpad / 64;
  OzFrag32<FM, FN> f0, f1;
  for (int pass = 0; pass < npass; ++pass) {
    const int8_t* Ab = Aplanes + (long)(pass ? pa1 : pa0) * mpad * kpad + (long)(row0 / 16) * kblk * 1024 + src_lane;
    const int8_t* Bb = Bplanes + (long)(pass ? pb1 : pb0) * npad * kpad + (long)(col0 / 16) * kblk * 1024 + src_lane;
    auto issue = [&](int stage, int kt) {
#pragma unroll
      for (int q = 0; q < SEGA_W; ++q) {
        const int seg = wu * SEGA_W + q;
        vcphip_lds_dma16(Ab + ((long)seg * kblk + kt) * 1024, sA_base + stage * (TMB * 64) + seg * 1024);
      }
#pragma unroll
      for (int q = 0; q < SEGB_W; ++q) {
        const int seg = wu * SEGB_W + q;
        vcphip_lds_dma16(Bb + ((long)seg * kblk + kt) * 1024, sB_base + stage * (TNB * 64) + seg * 1024);
      }
    };
    auto wait_next = [&](int kt) {
      if (kt + NST < nk) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((NST - 2) * G) : "memory");
      else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    };
    auto load = [&](int stage, OzFrag32<FM, FN>& f) {
      oz_load_frag32<FM, FN>(sA[stage] + fa0, sA[stage] + fa1, sB[stage] + fb0, sB[stage] + fb1, f);
    };
    const int npro = nk < NST ? nk : NST;
    for (int st = 0; st < npro; ++st) issue(st, st);
    if (nk >= NST) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((NST - 1) * G) : "memory");
    else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    load(0, f0);
    for (int kt = 0; kt < nk; kt += 2) {
      if (kt + 1 < nk) {
        wait_next(kt);
        if (kt + NST < nk) issue(kt % NST, kt + NST);
        load((kt + 1) % NST, f1);
      }
      oz_mfma_frag32<FM, FN>(f0, acc);
      if (kt + 1 < nk) {
        if (kt + 2 < nk) {
          wait_next(kt + 1);
          if (kt + 1 + NST < nk) issue((kt + 1) % NST, kt + 1 + NST);
          load((kt + 2) % NST, f0);
        }
        oz_mfma_frag32<FM, FN>(f1, acc);
      }
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }

  // epilogue: registers 4q..4q+3 of fragment (i, j) are lane l2 of the 16 x 16
  // block (row half q / 2, column half (lane % 32) / 16) of that fragment
  const long blocks_m = mpad / 16;
  const int l2 = 16 * (kq) + (lane & 15), chalf = (lane & 31) >> 4;
  if (z < nmod) {
    const OzModPlane mod(P, z, kpad);
    uint8_t* Uz = U + (long)z * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long bi = (row0 + wm * WM + 32 * i) / 16 + (q >> 1);
          const long bj = (col0 + wn * WN + 32 * j) / 16 + chalf;
          const unsigned packed =
              mod.pack(acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]);
          *reinterpret_cast<unsigned*>(Uz + (bj * blocks_m + bi) * 256 + (l2 + 32 * (q & 1)) * 4) = packed;
        }
  } else {
    int32_t* Zz = Z + (long)(z - nmod) * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long bi = (row0 + wm * WM + 32 * i) / 16 + (q >> 1);
          const long bj = (col0 + wn * WN + 32 * j) / 16 + chalf;
          const v4i val = v4i{acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]};
          *reinterpret_cast<v4i*>(Zz + (bj * blocks_m + bi) * 256 + (l2 + 32 * (q & 1)) * 4) = val;
        }
  }
}

// 256 x 384 workgroup tile, 17% fewer operand bytes per int8 op than the
// 256 x 256 tiles: eight waves in a 2 x 4 grid, each 128 x 96 (4 x 3
// fragments of v_mfma_i32_32x32x32_i8, 192 accumulator registers).  Each
// 64-k stage runs as two 32-k halves whose fragments are read just before
// their MFMAs (the other wave on the SIMD covers that LDS latency), which
// keeps a wave within 256 registers.  NST stages, NST - 1 in flight: the
// barrier that opens stage kt also frees stage kt - 1, which is refilled
// with k-step kt + NST - 1 before the MFMAs of stage kt.  The last column
// tile may pass npad (a multiple of 256): its loads beyond npad re-read the
// last 16-column segment, so every wave issues the same loads per stage, and
// its outputs there are not stored.
template <int NST>
__global__ __launch_bounds__(512) void k_oz_gemm6(int mpad, int npad, int kpad, int nmod, int nz,
                                                  const int8_t* __restrict__ Aplanes,
                                                  const int8_t* __restrict__ Bplanes, VcphipOzParams P,
                                                  uint8_t* __restrict__ U, int32_t* __restrict__ Z) {
  constexpr int FM = 4, FN = 3, NW = 8;
  constexpr int WM = 32 * FM, WN = 32 * FN;
  constexpr int TMB = 2 * WM, TNB = (NW / 2) * WN;   // 256 x 384
  constexpr int SEGA = TMB / 16, SEGB = TNB / 16;
  constexpr int SEGA_W = SEGA / NW, SEGB_W = SEGB / NW;
  constexpr int G = SEGA_W + SEGB_W;                  // loads per wave per stage
  static_assert(SEGA_W * NW == SEGA && SEGB_W * NW == SEGB, "tile/wave mismatch");
  static_assert(NST >= 2 && (NST - 1) * G < 64, "vmcnt range");
  __shared__ __attribute__((aligned(16))) int8_t sA[NST][TMB * 64];
  __shared__ __attribute__((aligned(16))) int8_t sB[NST][TNB * 64];

  const int ntm = mpad / TMB, ntn = (npad + TNB - 1) / TNB;
  const int tiles = ntm * ntn;
  const int z = blockIdx.x / tiles;
  const int local = vcphip_xcd_remap(blockIdx.x - z * tiles, tiles);
  int tm, tn;
  vcphip_tile_coords(local, ntm, ntn, 8, tm, tn);
  const int row0 = tm * TMB, col0 = tn * TNB;
  int npass, pa0, pb0, pa1, pb1;
  oz_pass_planes(P, z, nmod, npass, pa0, pb0, pa1, pb1);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int wu = __builtin_amdgcn_readfirstlane(w);
  const uint32_t sA_base = vcphip_lds_addr(&sA[0][0]), sB_base = vcphip_lds_addr(&sB[0][0]);
  const int lrow = lane >> 2, lchunk = (lane & 3) ^ ((lane >> 4) & 3);
  const int src_lane = lrow * 64 + 16 * lchunk;   // inside the 1 KB block of 16 rows x 64 k
  const long kblk = kpad / 64;
  const int last_seg_b = npad / 16 - 1;
  // this wave's B segments (16 columns each), clamped to the plane
  int bseg[SEGB_W];
#pragma unroll
  for (int q = 0; q < SEGB_W; ++q) {
    const int sg = col0 / 16 + wu * SEGB_W + q;
    bseg[q] = sg < last_seg_b ? sg : last_seg_b;
  }
  // fragment rows: R = 32 i + lane % 32; k half h reads logical chunks 2 h + kq
  const int frow = lane & 31, kq = lane >> 5, sw = (frow >> 2) & 3;
  const int fs0 = 16 * ((0 + kq) ^ sw), fs1 = 16 * ((2 + kq) ^ sw);
  const int fa0 = (wm * WM + frow) * 64 + fs0, fa1 = (wm * WM + frow) * 64 + fs1;
  const int fb0 = (wn * WN + frow) * 64 + fs0, fb1 = (wn * WN + frow) * 64 + fs1;

  v16i acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v16i{};

  const int nk = kpad / 64;
  for (int pass = 0; pass < npass; ++pass) {
    const int8_t* Ab = Aplanes + (long)(pass ? pa1 : pa0) * mpad * kpad + (long)(row0 / 16) * kblk * 1024 + src_lane;
    const int8_t* Bb = Bplanes + (long)(pass ? pb1 : pb0) * npad * kpad + src_lane;
    auto issue = [&](int stage, int kt) {
#pragma unroll
      for (int q = 0; q < SEGA_W; ++q) {
        const int seg = wu * SEGA_W + q;
        vcphip_lds_dma16(Ab + ((long)seg * kblk + kt) * 1024, sA_base + stage * (TMB * 64) + seg * 1024);
      }
#pragma unroll
      for (int q = 0; q < SEGB_W; ++q) {
        const int seg = wu * SEGB_W + q;
        vcphip_lds_dma16(Bb + ((long)bseg[q] * kblk + kt) * 1024, sB_base + stage * (TNB * 64) + seg * 1024);
      }
    };
    auto half = [&](int stage, int h) {
      const int8_t* la = sA[stage] + (h ? fa1 : fa0);
      const int8_t* lb = sB[stage] + (h ? fb1 : fb0);
      v4i a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = *reinterpret_cast<const v4i*>(la + i * 2048);
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = *reinterpret_cast<const v4i*>(lb + j * 2048);
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int i = 0; i < FM; ++i) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], acc[i][j], 0, 0, 0);
    };
    const int npro = nk < NST - 1 ? nk : NST - 1;
    for (int st = 0; st < npro; ++st) issue(st, st);
    for (int kt = 0; kt < nk; ++kt) {
      // stage kt landed for this wave once only the later stages' loads
      // (those issued: up to kt + NST - 2) may still be in flight
      const int later = (kt + NST - 2 < nk ? kt + NST - 2 : nk - 1) - kt;
      if (later >= NST - 2 && NST > 2) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((NST - 2) * G) : "memory");
      else if (later == 1) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(G) : "memory");
      else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      if (kt + NST - 1 < nk) issue((kt + NST - 1) % NST, kt + NST - 1);
      half(kt % NST, 0);
      half(kt % NST, 1);
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }

  // epilogue (as k_oz_gemm5's): registers 4q..4q+3 of fragment (i, j) are
  // lane l2 of the 16 x 16 block (row half q / 2, column half (lane % 32) /
  // 16) of that fragment; fragments past npad are not stored
  const long blocks_m = mpad / 16;
  const int l2 = 16 * (kq) + (lane & 15), chalf = (lane & 31) >> 4;
  if (z < nmod) {
    const OzModPlane mod(P, z, kpad);
    uint8_t* Uz = U + (long)z * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        if (col0 + wn * WN + 32 * j >= npad) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long bi = (row0 + wm * WM + 32 * i) / 16 + (q >> 1);
          const long bj = (col0 + wn * WN + 32 * j) / 16 + chalf;
          const unsigned packed =
              mod.pack(acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]);
          *reinterpret_cast<unsigned*>(Uz + (bj * blocks_m + bi) * 256 + (l2 + 32 * (q & 1)) * 4) = packed;
        }
      }
  } else {
    int32_t* Zz = Z + (long)(z - nmod) * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        if (col0 + wn * WN + 32 * j >= npad) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long bi = (row0 + wm * WM + 32 * i) / 16 + (q >> 1);
          const long bj = (col0 + wn * WN + 32 * j) / 16 + chalf;
          const v4i val = v4i{acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]};
          *reinterpret_cast<v4i*>(Zz + (bj * blocks_m + bi) * 256 + (l2 + 32 * (q & 1)) * 4) = val;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// reconstruction
// ---------------------------------------------------------------------------
// 32-bit limbs of X: M <= 256^N, so |X| < M/2 fits ceil(N / 4) limbs with
// its sign bit to spare
template <int N>
constexpr int oz_limbs() {
  return (N + 3) / 4;
}

// |X| * 2^s rounded toward zero (tz) and away from zero (aw), exactly: both
// come from one 64-bit window of the limbs and the sticky bit below it
template <int NL>
VCPHIP_DEVINL void conv_limbs2(const unsigned (&mag)[NL], int s, double& tz, double& aw) {
  int top = -1;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l)
    if (top < 0 && mag[l] != 0) top = 32 * l + 31 - __builtin_clz(mag[l]);
  if (top < 0) {
    tz = aw = 0.0;
    return;
  }
  int q = top - 52;
  if (q < -1074 - s) q = -1074 - s;
  unsigned long long mant;
  bool rest = false;
  if (q <= 0) {
    mant = (unsigned long long)mag[0] | ((unsigned long long)mag[1] << 32);
    q = 0;
  } else {
    // bits [q, q+64) and the sticky tail below q
    const int ls = q >> 5, bs = q & 31;
    unsigned w0 = 0, w1 = 0, w2 = 0;
    bool tail = false;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      if (l == ls) w0 = mag[l];
      if (l == ls + 1) w1 = mag[l];
      if (l == ls + 2) w2 = mag[l];
      if (l < ls && mag[l] != 0) tail = true;
    }
    if (bs != 0 && (w0 & ((1u << bs) - 1u)) != 0) tail = true;
    const unsigned long long lo = (unsigned long long)w0 | ((unsigned long long)w1 << 32);
    mant = bs ? ((lo >> bs) | ((unsigned long long)w2 << (64 - bs))) : lo;
    if (q >= 32 * NL) {
      mant = 0;
      tail = true;
    }
    rest = tail;
  }
  const int ex = q + s;
  const unsigned long long mant_aw = mant + (rest ? 1ull : 0ull);
  // mant < 2^53 except in the subnormal clamp (q <= 0), where both are exact
  tz = (mant != 0 && ex + (64 - __builtin_clzll(mant)) > 1024) ? 1.7976931348623157e308 : __builtin_ldexp((double)mant, ex);
  aw = (mant_aw != 0 && ex + (64 - __builtin_clzll(mant_aw)) > 1024) ? __builtin_inf() : __builtin_ldexp((double)mant_aw, ex);
}

template <int N>
__global__ __launch_bounds__(256) void k_oz_recon(int m, int n, int k, int mpad, int npad, VcphipOzParams P,
                                                  VcphipOzSide sa, VcphipOzSide sb,
                                                  const uint8_t* __restrict__ U, const int32_t* __restrict__ Z,
                                                  double* __restrict__ C, long ldc) {
  vcphip_mode_round_up();
  // threads follow the blocked 16 x 16 layout of the planes (coalesced
  // residue reads): offset o in a block holds row 4 (o / 64) + o % 4 and
  // column (o / 4) % 16 of that block
  // grid: x = 16-row block, y = 16-column block (the plane's block order)
  // the block's planes are addressed from a uniform base (Ub, Zb) plus the
  // thread's offset o, and the row / column tables by unsigned indices
  // every lane runs through the CRT's matrix instructions; lanes past m or n
  // (inside the padded planes) read row / column 0's tables and store nothing
  const unsigned o = threadIdx.x;
  const long bbase = ((long)blockIdx.y * (mpad >> 4) + blockIdx.x) * 256;
  const unsigned i0 = blockIdx.x * 16 + 4 * (o >> 6) + (o & 3);
  const unsigned j0 = blockIdx.y * 16 + ((o >> 2) & 15);
  const bool live = i0 < (unsigned)m && j0 < (unsigned)n;
  const unsigned i = live ? i0 : 0u, j = live ? j0 : 0u;
  const long plane = (long)mpad * npad;
  const int32_t* __restrict__ Zb = Z + bbase;

  VcphipCrtOps<N> T;
  vcphip_crt_ops_load<N>(P, T);
  constexpr int NL = oz_limbs<N>();
  unsigned mag[NL];
  const bool neg = vcphip_crt_symmetric<N, NL>(U + bbase, plane, o, T, P, mag);
  const int fa = sa.flags[i], fb = sb.flags[j];
  const bool zr = (fa & VCPHIP_OZ_MIDZERO) != 0, zc = (fb & VCPHIP_OZ_MIDZERO) != 0;
  const int ei = sa.e[i], fj = sb.e[j];
  const int sc = ei + fj - 2 * P.beta;
  // RD and RU of the exact midpoint product
  double tz, aw;
  conv_limbs2(mag, sc, tz, aw);
  const double lo_mid = neg ? -aw : tz, hi_mid = neg ? -tz : aw;
  // radius term (|mA|+rA) rB + rA |mB| (pidblas :277-283) from the digit
  // products.  With x1 <= (h1 + l1/127) 2^g1/127 <= H1 2^g1/127 (A, l1 >= 0)
  // and x2 <= (H + L/127) 2^g2/127 <= H 2^g2/127 (B, L <= 0), per term
  //   full    sum x1 x2 <= 2^(g1+g2)/127^2 (Zhh + Zmix/127 + Zll/127^2),
  //           Zhh = sum h1 H, Zmix = sum h1 L + l1 H, Zll = sum l1 L <= 0;
  //   folded  the same without Zll (<= 0, so the bound only rises);
  //   half    (Zhh + Zlh/127) with Zlh = sum l1 H;
  //   ceil    Zcc = sum H1 H.
  // A quotient of a numerator of either sign is bounded above by the
  // numerator times the RU or, when negative, the RD reciprocal.
  double t1 = 0.0, t2 = 0.0;
  const bool ceil_only = P.rscheme == VCPHIP_OZ_RADIUS_CEIL, full = P.rscheme == VCPHIP_OZ_RADIUS_FULL;
  // the directed reciprocals as register values (a per-lane choice between
  // two parameter fields would otherwise become a per-lane load)
  const double c1 = vcphip_uniform(P.c1), c1dn = vcphip_uniform(P.c1dn);
  const double c127 = vcphip_uniform(P.c127), c127dn = vcphip_uniform(P.c127dn);
  auto zload = [&](int z) {   // digit product z of this element
    return (int)__builtin_amdgcn_raw_buffer_load_b32(vcphip_rsrc(Zb + z * plane, 1024), 4 * o, 0, 0);
  };
  auto digit_sum = [&](int zh, int zm, int zl) {
    if (ceil_only) return (double)zload(zh);
    const int mx = zload(zm);
    double u = (double)zload(zh) + (double)mx * (mx >= 0 ? c1 : c1dn);
    if (full) {
      const int ll = zload(zl);
      u = u + (double)ll * (ll >= 0 ? c127 : c127dn);
    }
    return u;
  };
  if (!((fa & VCPHIP_OZ_SUMZERO) || (fb & VCPHIP_OZ_RADZERO)))
    t1 = vcphip_mulpow2(ceil_only ? digit_sum(0, 0, 0) : digit_sum(0, 1, 4), sa.g[i] + sb.h[j]) * P.c127;
  if (!((fa & VCPHIP_OZ_RADZERO) || zc))
    t2 = vcphip_mulpow2(ceil_only ? digit_sum(1, 1, 1) : digit_sum(2, 3, 5), sa.h[i] + fj) * P.c127;
  const double rt = t1 + t2;
  // scaling error of the midpoint product
  double E = 0.0;
  if (!zr && !zc) {
    const double ka = vcphip_mulpow2((double)k, ei - P.beta - 1);
    const double e1 = vcphip_mulpow2(sa.n1[i] + ka, fj - P.beta - 1);
    const double e2 = vcphip_mulpow2(sb.n1[j], ei - P.beta - 1);
    E = e1 + e2;
  }
  const double rad = rt + E;
  double hi = hi_mid + rad;
  double lo = -(rad - lo_mid);
  if ((fa | fb) & VCPHIP_OZ_NONFINITE) {   // a non-finite operand end in the row or column
    const bool nan = ((fa | fb) & VCPHIP_OZ_HASNAN) != 0;
    lo = nan ? __builtin_nan("") : -__builtin_inf();
    hi = nan ? __builtin_nan("") : __builtin_inf();
  }
  if (live) *reinterpret_cast<double2*>(C + 2 * (i + ldc * (long)j)) = make_double2(lo, hi);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int vcphip_oz_stats(const double* X, long rs, long cs, int rows, int len, VcphipOzSide side,
                    hipStream_t s) {
  if (rows <= 0) return 0;
  if (cs == 1) {   // columns of a column-major operand: one wave per column
    hipLaunchKernelGGL(k_oz_stats_cols, dim3((rows + 3) / 4), dim3(256), 0, s, X, rs, rows, len, side);
    return hipGetLastError() == hipSuccess ? 0 : -3;
  }
  const int nchunk = len > 0 ? (len + VCPHIP_OZ_CHUNK - 1) / VCPHIP_OZ_CHUNK : 0;
  if (nchunk > 0) {
    dim3 grid((rows + 255) / 256, nchunk);
    hipLaunchKernelGGL(k_oz_stats_part, grid, dim3(256), 0, s, X, rs, cs, rows, len, nchunk, side.part);
  }
  hipLaunchKernelGGL(k_oz_stats_final, dim3((rows + 255) / 256), dim3(256), 0, s, rows, nchunk, side.part, side);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_oz_split_rows(const double* A, long lda, int m, int k, int mpad, int kpad,
                         const VcphipOzParams& P, VcphipOzSide side, int8_t* planes, hipStream_t s) {
  dim3 grid(mpad / SR_ROWS, kpad / SR_K);
  const size_t lds = (size_t)(P.nmod + VCPHIP_OZ_NDIGIT_A) * SR_ROWS * SR_LD;
  switch (P.nmod) {
#define VCPHIP_SPLIT_CASE(NN) \
  case NN: hipLaunchKernelGGL((k_oz_split_rows<NN>), grid, dim3(256), lds, s, A, lda, m, k, mpad, kpad, P, side, planes); break;
    VCPHIP_SPLIT_CASE(8) VCPHIP_SPLIT_CASE(9) VCPHIP_SPLIT_CASE(10) VCPHIP_SPLIT_CASE(11) VCPHIP_SPLIT_CASE(12)
    VCPHIP_SPLIT_CASE(13) VCPHIP_SPLIT_CASE(14) VCPHIP_SPLIT_CASE(15) VCPHIP_SPLIT_CASE(16) VCPHIP_SPLIT_CASE(17)
    VCPHIP_SPLIT_CASE(18) VCPHIP_SPLIT_CASE(19) VCPHIP_SPLIT_CASE(20)
#undef VCPHIP_SPLIT_CASE
    default: return -7;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_oz_split_cols(const double* B, long ldb, int k, int n, int npad, int kpad,
                         const VcphipOzParams& P, VcphipOzSide side, int8_t* planes, hipStream_t s) {
  const long blocks = (long)(npad / 16) * (kpad / 64);   // one 1 KB plane block each
  if (blocks > 0x7FFFFFFFL) return -7;
  const dim3 grid((unsigned)blocks);
  switch (P.nmod) {
#define VCPHIP_SPLIT_CASE(NN) \
  case NN: hipLaunchKernelGGL((k_oz_split_cols<NN>), grid, dim3(256), 0, s, B, ldb, k, n, npad, kpad, P, side, planes); break;
    VCPHIP_SPLIT_CASE(8) VCPHIP_SPLIT_CASE(9) VCPHIP_SPLIT_CASE(10) VCPHIP_SPLIT_CASE(11) VCPHIP_SPLIT_CASE(12)
    VCPHIP_SPLIT_CASE(13) VCPHIP_SPLIT_CASE(14) VCPHIP_SPLIT_CASE(15) VCPHIP_SPLIT_CASE(16) VCPHIP_SPLIT_CASE(17)
    VCPHIP_SPLIT_CASE(18) VCPHIP_SPLIT_CASE(19) VCPHIP_SPLIT_CASE(20)
#undef VCPHIP_SPLIT_CASE
    default: return -7;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// tile variants (all bit-identical): 6 / 7 = k_oz_gemm3 256 x 256 / 128 x 256;
// 8 / 9 / 10 = k_oz_gemm4 256 x 256 with 4 / 3 stages, 128 x 256 with 4;
// 13 / 14 = k_oz_gemm4 256 x 256 with the refill loads spread over the
// MFMAs (4 / 3 stages); 11 / 12 = k_oz_gemm5 (32 x 32 x 32 MFMA, 4 waves,
// 4 / 3 stages); 15 / 16 = k_oz_gemm5 with 8 waves of 128 x 64 (4 / 3 stages)
int vcphip_oz_gemm(int mpad, int npad, int kpad, int nmod, int nrad, const int8_t* Aplanes,
                   const int8_t* Bplanes, const VcphipOzParams& P, uint8_t* U, int32_t* Z, int variant,
                   hipStream_t s) {
  const int nz = nmod + nrad;
  switch (variant) {
    case 6:
    case 7: {
      const int fm = variant == 6 ? 8 : 4, fn = variant == 6 ? 8 : 8;
      if (kpad % 64) return -7;
      const int tiles = (mpad / (32 * fm)) * (npad / (32 * fn));
      if (variant == 6)
        hipLaunchKernelGGL((k_oz_gemm3<8, 8>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz, Aplanes,
                           Bplanes, P, U, Z);
      else
        hipLaunchKernelGGL((k_oz_gemm3<4, 8>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz, Aplanes,
                           Bplanes, P, U, Z);
      break;
    }
    case 8:
    case 9:
    case 10: {
      if (kpad % 64) return -7;
      const int fm = variant == 10 ? 4 : 8;
      const int tiles = (mpad / (32 * fm)) * (npad / 256);
      if (variant == 8)
        hipLaunchKernelGGL((k_oz_gemm4<8, 8, 4>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      else if (variant == 9)
        hipLaunchKernelGGL((k_oz_gemm4<8, 8, 3>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      else
        hipLaunchKernelGGL((k_oz_gemm4<4, 8, 4>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      break;
    }
    case 13:
    case 14: {
      if (kpad % 64) return -7;
      const int tiles = (mpad / 256) * (npad / 256);
      if (variant == 13)
        hipLaunchKernelGGL((k_oz_gemm4<8, 8, 4, true>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      else
        hipLaunchKernelGGL((k_oz_gemm4<8, 8, 3, true>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      break;
    }
    case 11:
    case 12: {
      if (kpad % 64) return -7;
      const int tiles = (mpad / 256) * (npad / 256);
      if (variant == 11)
        hipLaunchKernelGGL((k_oz_gemm5<4, 4, 4>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      else
        hipLaunchKernelGGL((k_oz_gemm5<4, 4, 3>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      break;
    }
    case 15:
    case 16: {
      if (kpad % 64) return -7;
      const int tiles = (mpad / 256) * (npad / 256);
      if (variant == 15)
        hipLaunchKernelGGL((k_oz_gemm5<4, 2, 4, 8>), dim3(tiles * nz), dim3(512), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      else
        hipLaunchKernelGGL((k_oz_gemm5<4, 2, 3, 8>), dim3(tiles * nz), dim3(512), 0, s, mpad, npad, kpad, nmod, nz,
                           Aplanes, Bplanes, P, U, Z);
      break;
    }
    case 17:
    case 18: {
      if (kpad % 64 || npad % 256) return -7;
      const int tiles = (mpad / 256) * ((npad + 383) / 384);
      if (variant == 17)
        hipLaunchKernelGGL((k_oz_gemm6<3>), dim3(tiles * nz), dim3(512), 0, s, mpad, npad, kpad, nmod, nz, Aplanes,
                           Bplanes, P, U, Z);
      else
        hipLaunchKernelGGL((k_oz_gemm6<2>), dim3(tiles * nz), dim3(512), 0, s, mpad, npad, kpad, nmod, nz, Aplanes,
                           Bplanes, P, U, Z);
      break;
    }
    default: return -7;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_oz_recon(int m, int n, int k, int mpad, int npad, const VcphipOzParams& P, VcphipOzSide sa,
                    VcphipOzSide sb, const uint8_t* U, const int32_t* Z, double* C, long ldc,
                    hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  if (P.rscheme < VCPHIP_OZ_RADIUS_FULL || P.rscheme > VCPHIP_OZ_RADIUS_CEIL) return -7;
  dim3 grid((unsigned)(mpad / 16), (unsigned)(npad / 16));
  switch (P.nmod) {
#define VCPHIP_RECON_CASE(NN) \
  case NN: hipLaunchKernelGGL((k_oz_recon<NN>), grid, dim3(256), 0, s, m, n, k, mpad, npad, P, sa, sb, U, Z, C, ldc); break;
    VCPHIP_RECON_CASE(8) VCPHIP_RECON_CASE(9) VCPHIP_RECON_CASE(10) VCPHIP_RECON_CASE(11)
    VCPHIP_RECON_CASE(12) VCPHIP_RECON_CASE(13) VCPHIP_RECON_CASE(14) VCPHIP_RECON_CASE(15)
    VCPHIP_RECON_CASE(16) VCPHIP_RECON_CASE(17) VCPHIP_RECON_CASE(18) VCPHIP_RECON_CASE(19)
    VCPHIP_RECON_CASE(20)
#undef VCPHIP_RECON_CASE
    default: return -7;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
