// Exact symmetric Chinese-remainder reconstruction for the Ozaki scheme II
// kernels (interval and double-double forms).
//
// The residues u_t = X mod m_t (t < N) of an integer X with |X| < M/2 - the
// splitting parameter beta guarantees |X| <= (M/2) 2^-1e-6 (interval) or
// |X| < M/4 (dd) - determine X as
//     S = sum_t u_t w_t,   w_t = (M/m_t) ((M/m_t)^-1 mod m_t),
//     X = S - q M,         q = round(S / M) = floor(sum_t u_t (w_t / M) + 1/2).
// q is evaluated in double precision: with u_t < 256, N <= 32 and w_t / M
// rounded to double, the computed sum is within 2^-34 of S / M, which stays
// at least 3.4e-7 (interval) or 1/4 (dd) away from a half-integer, so the
// floor is exact in every rounding mode.  S is accumulated exactly in byte
// columns with v_dot4_u32_u8 (column c = sum_t u_t * byte c of w_t < 2^21),
// and X is formed modulo 2^(32 NL), which suffices because |X| < 2^(32 NL - 1).
// This replaces a mixed-radix (Garner) reconstruction: O(N * NL) byte dot
// products instead of O(N^2) modular steps and an N-step multiword Horner.
#pragma once

#include "vcphip_device.h"
#include "vcphip_kernels.h"

// From the residues u[t] (bytes) of X: returns the sign of X and its
// magnitude in NL little-endian 32-bit limbs.
template <int N, int NL>
VCPHIP_DEVINL bool vcphip_crt_from_residues(const unsigned (&u)[N], const VcphipOzParams& P, unsigned (&mag)[NL]) {
  constexpr int G = (N + 3) / 4;
  constexpr int NB = N < 4 * NL ? N : 4 * NL;   // byte columns below 2^(32 NL); w_t < M < 2^(8 N)
  unsigned r[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    unsigned x = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (4 * g + e < N) x |= u[4 * g + e] << (8 * e);
    r[g] = x;
  }
  double qf = 0.5;
#pragma unroll
  for (int t = 0; t < N; ++t) qf = __builtin_fma((double)u[t], P.crt_f[t], qf);
  const long long q = (long long)__builtin_floor(qf);
  unsigned col[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    unsigned a = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) a = __builtin_amdgcn_udot4(r[g], P.crt_w8[c][g], a, false);
    col[c] = a;
  }
  long long carry = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    long long acc = carry;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (4 * l + e < NB) acc += (long long)col[4 * l + e] << (8 * e);
    acc -= q * (long long)P.Mlimb[l];
    mag[l] = (unsigned)acc;
    carry = acc >> 32;
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
                                        const VcphipOzParams& P, unsigned (&mag)[NL]) {
  unsigned u[N];
#pragma unroll
  for (int t = 0; t < N; ++t) u[t] = __builtin_amdgcn_raw_buffer_load_b8(vcphip_rsrc(U + (long)t * plane, 256), o, 0, 0);
  return vcphip_crt_from_residues<N, NL>(u, P, mag);
}
