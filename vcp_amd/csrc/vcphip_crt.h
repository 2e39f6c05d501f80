// Exact symmetric Chinese-remainder reconstruction for the Ozaki scheme II
// kernels (interval and double-double forms).
//
// The residues u_t = X mod m_t (t < N) of an integer X with |X| < M/2 - the
// splitting parameter beta guarantees |X| <= (M/2) 2^-1e-6 (interval) or
// |X| < M/4 (dd) - determine X as
//     S = sum_t u_t w_t,   w_t = (M/m_t) ((M/m_t)^-1 mod m_t),
//     X = S - q M,         q = round(S / M) = floor(sum_t u_t (w_t / M) + 1/2).
// Both sums are byte-column sums, formed on the matrix cores:
//     col_c = sum_t u_t * byte c of w_t             (c < N, w_t < M < 2^(8 N))
//     fc_b  = sum_t u_t * byte b of F_t,   F_t = floor(2^48 w_t / M)   (b < 6)
// so S = sum_c col_c 2^(8c) and q = floor((sum_b fc_b 2^(8b) + 2^47) / 2^48).
// Each F_t is within 2^-48 of w_t / M, so with u_t < 256 and N <= 32 the
// fixed-point quotient is within 2^-35 of S / M, which stays at least 3.4e-7
// (interval) or 1/4 (dd) away from a half-integer: q is exact.  X is formed
// modulo 2^(32 NL), which suffices because |X| < 2^(32 NL - 1).
//
// The column sums are integer matrix products of int8 operands:
// sum_t u_t v_t = sum_t (u_t - 128)(v_t - 128) + 128 sum_t u_t + K, with
// K = 128 sum_t v_t - 128^2 N a per-column constant.  One
// v_mfma_i32_32x32x32_i8 handles the 64 elements of a wave (one per lane)
// against 16 columns and 16 residues: output row r holds column
// 4 (r / 8) + r % 4 of element group (r / 4) % 2, so a lane's accumulator
// register v is column v of its own element.  Its B operand is the lane's own
// 16 centred residues (k-slots of lane half h carry element group h), and the
// A operand of lane half h is zero in the rows of the other group, which
// keeps the two groups apart.  The table's columns are the six quotient
// bytes (0 .. 5), a column of ones (6), which gives sum_t (u_t - 128), and
// the N weight bytes (7 .. N + 6).  This replaces O(N^2 / 4) byte dot
// products and N double-precision multiply-adds per element by
// ceil((N + 7) / 16) x ceil(N / 16) matrix instructions per 64 elements.
// Every lane of the wave must execute them (the kernels keep all lanes
// through them and mask their stores).
#pragma once

#include "vcphip_device.h"
#include "vcphip_kernels.h"

typedef int vcphip_v4i __attribute__((ext_vector_type(4)));
typedef int vcphip_v16i __attribute__((ext_vector_type(16)));

// The per-lane A operands (weights) of the CRT products.
template <int N>
struct VcphipCrtOps {
  static constexpr int NC = N + 7;           // 6 quotient bytes, the residue sum, N weight bytes
  static constexpr int CH = (NC + 15) / 16;  // 16-column groups
  static constexpr int RH = (N + 15) / 16;   // 16-residue groups
  static_assert(16 * CH <= VCPHIP_OZ_CRTCOL && 4 * RH <= VCPHIP_OZ_MAXMOD / 4, "CRT table");
  vcphip_v4i a[CH][RH];
};

template <int N>
VCPHIP_DEVINL void vcphip_crt_ops_load(const VcphipOzParams& P, VcphipCrtOps<N>& T) {
  const unsigned lane = threadIdx.x & 63u;
  const unsigned h = lane >> 5, r = lane & 31u;
  const bool mine = ((r >> 2) & 1u) == h;
  const unsigned cl = 4 * (r >> 3) + (r & 3u);
#pragma unroll
  for (int ch = 0; ch < T.CH; ++ch)
#pragma unroll
    for (int rh = 0; rh < T.RH; ++rh)
      T.a[ch][rh] = mine ? *reinterpret_cast<const vcphip_v4i*>(&P.crt_a[16 * ch + cl][4 * rh]) : vcphip_v4i{0, 0, 0, 0};
}

// From the residues u[t] (bytes) of this lane's X: returns the sign of X and
// its magnitude in NL little-endian 32-bit limbs.
template <int N, int NL>
VCPHIP_DEVINL bool vcphip_crt_from_residues(const unsigned (&u)[N], const VcphipCrtOps<N>& T, const VcphipOzParams& P,
                                            unsigned (&mag)[NL]) {
  static_assert(N <= 4 * NL, "byte columns beyond the limbs");
  constexpr int CH = VcphipCrtOps<N>::CH, RH = VcphipCrtOps<N>::RH;
  vcphip_v4i b[RH];
#pragma unroll
  for (int rh = 0; rh < RH; ++rh)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      unsigned x = 0;
#pragma unroll
      for (int y = 0; y < 4; ++y)
        if (16 * rh + 4 * e + y < N) x |= u[16 * rh + 4 * e + y] << (8 * y);
      b[rh][e] = (int)(x ^ 0x80808080u);   // u - 128 as int8 (bytes past N meet zero weights)
    }
  // column group 0 first: its columns 0..6 are the quotient bytes and the
  // residue sum, so q is known before the weight columns (7 .. N + 6) of the
  // later groups are formed, and one group's accumulators are live at a time
  vcphip_v16i acc = vcphip_v16i{};
#pragma unroll
  for (int rh = 0; rh < RH; ++rh) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(T.a[0][rh], b[rh], acc, 0, 0, 0);
  const int s128 = 128 * (acc[6] + 128 * N);   // 128 sum_t u_t
  unsigned long long Q = 1ull << 47;
#pragma unroll
  for (int c = 0; c < 6; ++c) Q += (unsigned long long)(unsigned)(acc[c] + s128 + P.crt_k[c]) << (8 * c);
  const long long q = (long long)(Q >> 48);
  long long limb[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) limb[l] = -q * (long long)P.Mlimb[l];
#pragma unroll
  for (int ch = 0; ch < CH; ++ch) {
    if (ch > 0) {
      acc = vcphip_v16i{};
#pragma unroll
      for (int rh = 0; rh < RH; ++rh) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(T.a[ch][rh], b[rh], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int c = 16 * ch + v - 7;   // weight byte c
      if (c < 0 || c >= N) continue;
      limb[c >> 2] += (long long)(unsigned)(acc[v] + s128 + P.crt_k[16 * ch + v]) << (8 * (c & 3));   // < 2^21
    }
  }
  long long carry = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const long long acc_l = limb[l] + carry;
    mag[l] = (unsigned)acc_l;
    carry = acc_l >> 32;
  }
  const bool neg = (int)mag[NL - 1] < 0;
  if (neg) {
    unsigned long long c = 1;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const unsigned long long v = (unsigned long long)(~mag[l]) + c;
      mag[l] = (unsigned)v;
      c = v >> 32;
    }
  }
  return neg;
}

// Residues of element o (< 256) of a 16 x 16 plane block from N planes of
// `plane` bytes, U being the block's base in the first plane (workgroup-
// uniform): one buffer load per plane, from a scalar resource at the lane's
// offset.
template <int N, int NL>
VCPHIP_DEVINL bool vcphip_crt_symmetric(const uint8_t* __restrict__ U, long plane, unsigned o,
                                        const VcphipCrtOps<N>& T, const VcphipOzParams& P, unsigned (&mag)[NL]) {
  unsigned u[N];
#pragma unroll
  for (int t = 0; t < N; ++t) u[t] = __builtin_amdgcn_raw_buffer_load_b8(vcphip_rsrc(U + (long)t * plane, 256), o, 0, 0);
  return vcphip_crt_from_residues<N, NL>(u, T, P, mag);
}
