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

typedef int v4i __attribute__((ext_vector_type(4)));

namespace {

// ---------------------------------------------------------------------------
// statistics: per row (or column) of a midpoint-radius operand
// ---------------------------------------------------------------------------
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
  const double* base = X + 2 * ((long)r * rs);
  for (int cc = c0; cc < c1; ++cc) {
    const double2 v = *reinterpret_cast<const double2*>(base + 2 * ((long)cc * cs));
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
}

__global__ __launch_bounds__(256) void k_oz_stats_final(int rows, int nchunk, const double* __restrict__ part,
                                                        VcphipOzSide side) {
  vcphip_mode_round_up();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  const long stride = (long)nchunk * rows;
  double mu = 0.0, smax = 0.0, rmax = 0.0, tot = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    const double a = part[0 * stride + (long)c * rows + r];
    const double b = part[1 * stride + (long)c * rows + r];
    const double d = part[2 * stride + (long)c * rows + r];
    const double s = part[3 * stride + (long)c * rows + r];
    mu = a > mu ? a : mu;
    smax = b > smax ? b : smax;
    rmax = d > rmax ? d : rmax;
    tot = tot + s;
  }
  side.zero[r] = mu == 0.0;
  side.e[r] = mu > 0.0 ? vcphip_fexp(mu) : 0;
  side.g[r] = smax > 0.0 ? vcphip_fexp(smax) : 0;
  side.h[r] = rmax > 0.0 ? vcphip_fexp(rmax) : 0;
  side.n1[r] = tot;
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

// upward digit in [0, 127] of x >= 0 against 2^g (x < 2^g); round-up mode
VCPHIP_DEVINL int digit_ru(double x, int g) {
  const double y = vcphip_mulpow2(x, -g) * 127.0;
  return (int)__builtin_ceil(y);
}

// residue of a signed 63-bit integer in the symmetric int8 range of m
VCPHIP_DEVINL int sym_residue(unsigned c0, unsigned c1, unsigned c2, bool neg, int t,
                              const VcphipOzParams& P) {
  const int m = P.mod[t];
  // |x| = c0 + c1 2^21 + c2 2^42,  c_i < 2^21
  const unsigned y = c0 + c1 * (unsigned)P.w21[t] + c2 * (unsigned)P.w42[t];  // < 2^30.1
  const double yd = (double)y;
  const double q = __builtin_floor(yd * P.rcpd[t]);
  int r = (int)(yd - q * (double)m);   // exact; off by at most one modulus
  r += (r < 0) ? m : 0;
  r -= (r >= m) ? m : 0;
  if (neg && r != 0) r = m - r;
  return (r >= (m + 1) / 2) ? r - m : r;
}

VCPHIP_DEVINL void split_parts(long long a, unsigned& c0, unsigned& c1, unsigned& c2, bool& neg) {
  neg = a < 0;
  const unsigned long long u = neg ? (unsigned long long)(-a) : (unsigned long long)a;
  c0 = (unsigned)(u & 0x1FFFFFull);
  c1 = (unsigned)((u >> 21) & 0x1FFFFFull);
  c2 = (unsigned)(u >> 42);
}

constexpr int SR_ROWS = 64, SR_K = 64, SR_LD = 68;   // LDS byte tile, odd-word row stride

// rows of A (m x k, column-major intervals) -> planes [N+2][mpad][kpad]:
// N residue planes of A', then the digits of |mid|+rad and of rad.
__global__ __launch_bounds__(256) void k_oz_split_rows(const double* __restrict__ A, long lda, int m, int k,
                                                       int mpad, int kpad, VcphipOzParams P, VcphipOzSide side,
                                                       int8_t* __restrict__ planes) {
  vcphip_mode_round_up();
  extern __shared__ __attribute__((aligned(16))) unsigned char s_tile[];   // [N+2][64][SR_LD]
  const int nplane = P.nmod + 2;
  const int row0 = blockIdx.x * SR_ROWS, k0 = blockIdx.y * SR_K;
  const int tid = threadIdx.x;
  const int lr = tid % SR_ROWS, kq = tid / SR_ROWS;
  const int r = row0 + lr;
  const bool rowok = r < m;
  const int e = rowok ? side.e[r] : 0;
  const int g = rowok ? side.g[r] : 0;
  const int h = rowok ? side.h[r] : 0;
  const bool zr = rowok ? (side.zero[r] != 0) : true;
  for (int q = 0; q < SR_K / 4; ++q) {
    const int lk = kq + 4 * q;
    const int kk = k0 + lk;
    double lo = 0.0, hi = 0.0;
    if (rowok && kk < k) {
      const double2 v = *reinterpret_cast<const double2*>(A + 2 * (r + lda * kk));
      lo = v.x;
      hi = v.y;
    }
    double am, ar;
    vcphip_midrad_ru(lo, hi, am, ar);
    const double as = __builtin_fabs(am) + ar;
    const long long ap = zr ? 0 : rnint_floor(vcphip_mulpow2(am, P.beta - e));
    unsigned c0, c1, c2;
    bool neg;
    split_parts(ap, c0, c1, c2, neg);
    for (int t = 0; t < P.nmod; ++t)
      s_tile[(t * SR_ROWS + lr) * SR_LD + lk] = (unsigned char)(signed char)sym_residue(c0, c1, c2, neg, t, P);
    s_tile[(P.nmod * SR_ROWS + lr) * SR_LD + lk] = (unsigned char)(as > 0.0 ? digit_ru(as, g) : 0);
    s_tile[((P.nmod + 1) * SR_ROWS + lr) * SR_LD + lk] = (unsigned char)(ar > 0.0 ? digit_ru(ar, h) : 0);
  }
  __syncthreads();
  // write out: per plane 64 rows x 64 bytes, 16 bytes per thread
  const int orow = tid / 4, oc = (tid % 4) * 16;
  const long plane_stride = (long)mpad * kpad;
  for (int p = 0; p < nplane; ++p) {
    const unsigned* src = reinterpret_cast<const unsigned*>(s_tile + (p * SR_ROWS + orow) * SR_LD + oc);
    uint4 v = make_uint4(src[0], src[1], src[2], src[3]);
    *reinterpret_cast<uint4*>(planes + p * plane_stride + (long)(row0 + orow) * kpad + k0 + oc) = v;
  }
}

// columns of B (k x n, column-major intervals) -> planes [N+2][npad][kpad]:
// N residue planes of B', then the digits of rad and of |mid|.  One thread
// handles 16 consecutive k of one column (a 256-byte contiguous read).
__global__ __launch_bounds__(256) void k_oz_split_cols(const double* __restrict__ B, long ldb, int k, int n,
                                                       int npad, int kpad, VcphipOzParams P, VcphipOzSide side,
                                                       int8_t* __restrict__ planes) {
  vcphip_mode_round_up();
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  const int ngrp = kpad / 16;
  const int j = (int)(gid / ngrp), grp = (int)(gid % ngrp);
  if (j >= npad) return;
  const bool colok = j < n;
  const int f = colok ? side.e[j] : 0;
  const int xi = colok ? side.h[j] : 0;
  const bool zc = colok ? (side.zero[j] != 0) : true;
  const int zeta = zc ? 0 : f;
  unsigned c0[16], c1[16], c2[16];
  bool ng[16];
  unsigned dr[4] = {0, 0, 0, 0}, dm[4] = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int kk = grp * 16 + u;
    double lo = 0.0, hi = 0.0;
    if (colok && kk < k) {
      const double2 v = *reinterpret_cast<const double2*>(B + 2 * (kk + ldb * (long)j));
      lo = v.x;
      hi = v.y;
    }
    double bm, br;
    vcphip_midrad_ru(lo, hi, bm, br);
    const long long bp = zc ? 0 : rnint_floor(vcphip_mulpow2(bm, P.beta - f));
    split_parts(bp, c0[u], c1[u], c2[u], ng[u]);
    const unsigned d1 = (unsigned)(br > 0.0 ? digit_ru(br, xi) : 0);
    const unsigned d2 = (unsigned)(bm != 0.0 ? digit_ru(__builtin_fabs(bm), zeta) : 0);
    dr[u >> 2] |= d1 << (8 * (u & 3));
    dm[u >> 2] |= d2 << (8 * (u & 3));
  }
  const long plane_stride = (long)npad * kpad;
  int8_t* dst = planes + (long)j * kpad + grp * 16;
  for (int t = 0; t < P.nmod; ++t) {
    unsigned wv[4] = {0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < 16; ++u)
      wv[u >> 2] |= ((unsigned)(unsigned char)(signed char)sym_residue(c0[u], c1[u], c2[u], ng[u], t, P)) << (8 * (u & 3));
    *reinterpret_cast<uint4*>(dst + t * plane_stride) = make_uint4(wv[0], wv[1], wv[2], wv[3]);
  }
  *reinterpret_cast<uint4*>(dst + P.nmod * plane_stride) = make_uint4(dr[0], dr[1], dr[2], dr[3]);
  *reinterpret_cast<uint4*>(dst + (P.nmod + 1) * plane_stride) = make_uint4(dm[0], dm[1], dm[2], dm[3]);
}

// ---------------------------------------------------------------------------
// int8 MFMA GEMM over the planes: C_z = A_z * B_z^T (both K-contiguous)
// ---------------------------------------------------------------------------
// Output layout is blocked by 16x16 MFMA tiles: block (bi, bj) occupies
// 256 consecutive elements, element (4*(l/16)+v, l%16) of the block at l*4+v,
// which is exactly the v_mfma_i32_16x16x64_i8 accumulator order, so every
// wave stores whole contiguous blocks.
constexpr int GK = 64;        // k bytes per stage
constexpr int GLD = GK + 16;  // LDS row stride (bytes): conflict-free 16-byte fragment reads

template <int FM, int FN>
__global__ __launch_bounds__(256) void k_oz_gemm(int mpad, int npad, int kpad, int nmod, int nz,
                                                 const int8_t* __restrict__ Aplanes,
                                                 const int8_t* __restrict__ Bplanes, VcphipOzParams P,
                                                 uint8_t* __restrict__ U, int32_t* __restrict__ Z) {
  constexpr int WM = 16 * FM, WN = 16 * FN;    // wave tile
  constexpr int TMB = 2 * WM, TNB = 2 * WN;    // workgroup tile (2 x 2 waves)
  constexpr int ACH = TMB * GK / 16 / 256;     // 16-byte chunks of A per thread per stage
  constexpr int BCH = TNB * GK / 16 / 256;
  __shared__ __attribute__((aligned(16))) int8_t sA[2][TMB * GLD];
  __shared__ __attribute__((aligned(16))) int8_t sB[2][TNB * GLD];

  const int ntm = mpad / TMB, ntn = npad / TNB;
  const int tiles = ntm * ntn;
  const int z = blockIdx.x / tiles;
  const int local = vcphip_xcd_remap(blockIdx.x - z * tiles, tiles);
  int tm, tn;
  vcphip_tile_coords(local, ntm, ntn, 8, tm, tn);
  const int row0 = tm * TMB, col0 = tn * TNB;
  const int8_t* Ab = Aplanes + (long)z * mpad * kpad + (long)row0 * kpad;
  const int8_t* Bb = Bplanes + (long)z * npad * kpad + (long)col0 * kpad;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;

  uint4 ra[ACH], rb[BCH];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < ACH; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      ra[q] = *reinterpret_cast<const uint4*>(Ab + (long)row * kpad + k0 + 16 * part);
    }
#pragma unroll
    for (int q = 0; q < BCH; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      rb[q] = *reinterpret_cast<const uint4*>(Bb + (long)row * kpad + k0 + 16 * part);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < ACH; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      *reinterpret_cast<uint4*>(&sA[buf][row * GLD + 16 * part]) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < BCH; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      *reinterpret_cast<uint4*>(&sB[buf][row * GLD + 16 * part]) = rb[q];
    }
  };

  v4i acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4i{0, 0, 0, 0};

  const int nk = kpad / GK;
  gload(0);
  sstore(0);
  __syncthreads();
  const int frow = lane & 15, fk = (lane >> 4) * 16;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * GK);
    v4i fa[FM], fb[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
      fa[i] = *reinterpret_cast<const v4i*>(&sA[buf][(wm * WM + 16 * i + frow) * GLD + fk]);
#pragma unroll
    for (int j = 0; j < FN; ++j)
      fb[j] = *reinterpret_cast<const v4i*>(&sB[buf][(wn * WN + 16 * j + frow) * GLD + fk]);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    if (kt + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }

  const long blocks_m = mpad / 16;
  if (z < nmod) {
    const int md = P.mod[z];
    const double rc = P.rcpd[z];
    uint8_t* Uz = U + (long)z * mpad * npad;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const long bi = (row0 + wm * WM + 16 * i) / 16, bj = (col0 + wn * WN + 16 * j) / 16;
        unsigned packed = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double y = (double)acc[i][j][v];
          const double qd = __builtin_floor(y * rc);
          int rr = (int)(y - qd * (double)md);
          rr += (rr < 0) ? md : 0;
          rr -= (rr >= md) ? md : 0;
          packed |= (unsigned)rr << (8 * v);
        }
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

// ---------------------------------------------------------------------------
// reconstruction
// ---------------------------------------------------------------------------
constexpr int NL = 5;   // 32-bit limbs: M < 2^160 for N <= 20

// |X| * 2^s rounded away from zero (away=true) or toward zero, exactly
VCPHIP_DEVINL double conv_limbs(const unsigned (&mag)[NL], int s, bool away) {
  int top = -1;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l)
    if (top < 0 && mag[l] != 0) top = 32 * l + 31 - __builtin_clz(mag[l]);
  if (top < 0) return 0.0;
  int q = top - 52;
  if (q < -1074 - s) q = -1074 - s;
  unsigned long long mant;
  bool rest = false;
  if (q <= 0) {
    mant = (unsigned long long)mag[0] | ((unsigned long long)mag[1] << 32);
    q = 0;
  } else {
    // gather bits [q, q+64) and the sticky tail below q
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
    if (q >= 32 * NL) { mant = 0; tail = true; }
    rest = tail;
  }
  if (rest && away) mant += 1;
  const int ex = q + s;
  if (mant != 0 && ex + (64 - __builtin_clzll(mant)) > 1024) return away ? __builtin_inf() : 1.7976931348623157e308;
  return __builtin_ldexp((double)mant, ex);
}

template <int N>
__global__ __launch_bounds__(256) void k_oz_recon(int m, int n, int k, int mpad, int npad, VcphipOzParams P,
                                                  VcphipOzSide sa, VcphipOzSide sb,
                                                  const uint8_t* __restrict__ U, const int32_t* __restrict__ Z,
                                                  double* __restrict__ C, long ldc) {
  vcphip_mode_round_up();
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= (long)m * n) return;
  const int i = (int)(gid % m), j = (int)(gid / m);
  const long boff = ((long)(j >> 4) * (mpad / 16) + (i >> 4)) * 256 + ((j & 15) + 16 * ((i & 15) >> 2)) * 4 + (i & 3);
  const long plane = (long)mpad * npad;

  // mixed-radix digits of X = A'B' mod M (Garner):
  //   v_t = (u_t + sum_{s<t} v_s (-(m_0..m_{s-1}) mod m_t)) (m_0..m_{t-1})^{-1} mod m_t
  int v[N];
#pragma unroll
  for (int t = 0; t < N; ++t) {
    const int md = P.mod[t];
    int acc = (int)U[(long)t * plane + boff];
#pragma unroll
    for (int s2 = 0; s2 < t; ++s2) acc += v[s2] * P.negP[s2][t];
    // acc < 2^20: float division estimate, corrected exactly
    int q = (int)__builtin_floorf((float)acc * P.rcpf[t]);
    int r = acc - q * md;
    r += (r < 0) ? md : 0;
    r -= (r >= md) ? md : 0;
    int x = r * P.pinv[t];
    q = (int)__builtin_floorf((float)x * P.rcpf[t]);
    r = x - q * md;
    r += (r < 0) ? md : 0;
    r -= (r >= md) ? md : 0;
    v[t] = r;
  }
  // X = v_0 + m_0 (v_1 + m_1 (v_2 + ...)) in 32-bit limbs
  unsigned mag[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) mag[l] = 0;
  mag[0] = (unsigned)v[N - 1];
#pragma unroll
  for (int t = N - 2; t >= 0; --t) {
    unsigned long long carry = (unsigned)v[t];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const unsigned long long p = (unsigned long long)mag[l] * (unsigned)P.mod[t] + carry;
      mag[l] = (unsigned)p;
      carry = p >> 32;
    }
  }
  // symmetric range: X > M/2 means X - M < 0
  bool gt = false, decided = false;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l) {
    const unsigned hm = P.halfM[l];
    if (!decided && mag[l] != hm) { gt = mag[l] > hm; decided = true; }
  }
  const bool neg = gt;
  if (neg) {   // mag = M - X
    long long borrow = 0;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const long long d = (long long)P.Mlimb[l] - (long long)mag[l] - borrow;
      borrow = d < 0 ? 1 : 0;
      mag[l] = (unsigned)(d + (borrow << 32));
    }
  }
  const bool zr = sa.zero[i] != 0, zc = sb.zero[j] != 0;
  const int ei = sa.e[i], fj = sb.e[j];
  const int sc = ei + fj - 2 * P.beta;
  // RD and RU of the exact midpoint product
  double lo_mid, hi_mid;
  if (!neg) {
    lo_mid = conv_limbs(mag, sc, false);
    hi_mid = conv_limbs(mag, sc, true);
  } else {
    lo_mid = -conv_limbs(mag, sc, true);
    hi_mid = -conv_limbs(mag, sc, false);
  }
  // radius term (|mA|+rA) rB + rA |mB| from the digit products (pidblas :277-283)
  const int z1 = Z[boff], z2 = Z[plane + boff];
  const int zeta = zc ? 0 : fj;
  const double t1 = vcphip_mulpow2((double)z1, sa.g[i] + sb.h[j]) * P.c127;
  const double t2 = vcphip_mulpow2((double)z2, sa.h[i] + zeta) * P.c127;
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
  const double hi = hi_mid + rad;
  const double nlo = rad - lo_mid;
  *reinterpret_cast<double2*>(C + 2 * (i + ldc * (long)j)) = make_double2(-nlo, hi);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int vcphip_oz_stats(const double* X, long rs, long cs, int rows, int len, VcphipOzSide side,
                    hipStream_t s) {
  if (rows <= 0) return 0;
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
  const size_t lds = (size_t)(P.nmod + 2) * SR_ROWS * SR_LD;
  hipLaunchKernelGGL(k_oz_split_rows, grid, dim3(256), lds, s, A, lda, m, k, mpad, kpad, P, side, planes);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_oz_split_cols(const double* B, long ldb, int k, int n, int npad, int kpad,
                         const VcphipOzParams& P, VcphipOzSide side, int8_t* planes, hipStream_t s) {
  const long threads = (long)npad * (kpad / 16);
  hipLaunchKernelGGL(k_oz_split_cols, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, B, ldb, k, n, npad,
                     kpad, P, side, planes);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_oz_gemm(int mpad, int npad, int kpad, int nmod, int nrad, const int8_t* Aplanes,
                   const int8_t* Bplanes, const VcphipOzParams& P, uint8_t* U, int32_t* Z, int variant,
                   hipStream_t s) {
  const int nz = nmod + nrad;
  if (variant == 1) {
    const int tiles = (mpad / 256) * (npad / 256);
    hipLaunchKernelGGL((k_oz_gemm<8, 8>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz, Aplanes,
                       Bplanes, P, U, Z);
  } else {
    const int tiles = (mpad / 128) * (npad / 128);
    hipLaunchKernelGGL((k_oz_gemm<4, 4>), dim3(tiles * nz), dim3(256), 0, s, mpad, npad, kpad, nmod, nz, Aplanes,
                       Bplanes, P, U, Z);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_oz_recon(int m, int n, int k, int mpad, int npad, const VcphipOzParams& P, VcphipOzSide sa,
                    VcphipOzSide sb, const uint8_t* U, const int32_t* Z, double* C, long ldc,
                    hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  const long total = (long)m * n;
  dim3 grid((unsigned)((total + 255) / 256));
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
