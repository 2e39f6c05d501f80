// Nearest double-double of an integer given as eight 32-bit limbs, scaled by
// 2^s: hi = RN(X 2^s), lo = RN((X - H) 2^s) with H the integer hi stands for
// (ties to even, overflow to infinity, gradual underflow), the conversion of
// the double-double Ozaki product's reconstruction.  Host and device code
// (the CPU test checks the fast path against the exact limb path).
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define VCPHIP_HD __host__ __device__ inline
#else
#define VCPHIP_HD inline
#endif

namespace vcphip_ddconv {

constexpr int NL = 8;
typedef unsigned __int128 u128;

// ---- exact limb path ------------------------------------------------------
VCPHIP_HD int top_bit(const unsigned (&x)[NL]) {
  int top = -1;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l)
    if (top < 0 && x[l] != 0) top = 32 * l + 31 - __builtin_clz(x[l]);
  return top;
}

// bits [q, q+64) of x (q >= 0)
VCPHIP_HD unsigned long long bits64(const unsigned (&x)[NL], int q) {
  const int ls = q >> 5, bs = q & 31;
  unsigned w0 = 0, w1 = 0, w2 = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    if (l == ls) w0 = x[l];
    if (l == ls + 1) w1 = x[l];
    if (l == ls + 2) w2 = x[l];
  }
  const unsigned long long lo = (unsigned long long)w0 | ((unsigned long long)w1 << 32);
  return bs ? ((lo >> bs) | ((unsigned long long)w2 << (64 - bs))) : lo;
}

VCPHIP_HD bool bit_at(const unsigned (&x)[NL], int b) {
  bool v = false;
  if (b < 0) return false;
#pragma unroll
  for (int l = 0; l < NL; ++l)
    if (l == (b >> 5)) v = (x[l] >> (b & 31)) & 1u;
  return v;
}

VCPHIP_HD bool any_below(const unsigned (&x)[NL], int b) {   // any bit < b set
  bool v = false;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const int lo = 32 * l;
    if (b >= lo + 32) v = v || x[l] != 0;
    else if (b > lo) v = v || (x[l] & ((1u << (b - lo)) - 1u)) != 0;
  }
  return v;
}

// nearest double of x 2^s (ties to even); q_out = lowest kept bit, up_out = rounded up
VCPHIP_HD double rn_limbs(const unsigned (&x)[NL], int s, int& q_out, bool& up_out) {
  const int top = top_bit(x);
  q_out = 0;
  up_out = false;
  if (top < 0) return 0.0;
  int q = top - 52;
  if (q < -1074 - s) q = -1074 - s;
  if (q <= 0) {
    q_out = 0;
    return __builtin_ldexp((double)bits64(x, 0), s);
  }
  unsigned long long mant = q < 32 * NL ? bits64(x, q) : 0ull;
  const bool rb = bit_at(x, q - 1);
  const bool st = any_below(x, q - 1);
  const bool up = rb && (st || (mant & 1ull));
  mant += up ? 1ull : 0ull;
  q_out = q;
  up_out = up;
  const int ex = q + s;
  if (mant != 0 && ex + (64 - __builtin_clzll(mant)) > 1024) return __builtin_inf();
  return __builtin_ldexp((double)mant, ex);
}

// hi and lo of X 2^s by the limb path: RN of X, then RN of the remainder
VCPHIP_HD void limbs_hi_lo(const unsigned (&mag)[NL], int sc, double& hi, double& lo) {
  int q;
  bool up;
  hi = rn_limbs(mag, sc, q, up);
  // remainder |X| - round(|X|) at bit q: low bits, or 2^q - low bits if rounded up
  unsigned rem[NL];
  bool rem_neg = false;
  if (q <= 0) {
#pragma unroll
    for (int l = 0; l < NL; ++l) rem[l] = 0;
  } else {
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const int lo = 32 * l;
      unsigned w = mag[l];
      if (q <= lo) w = 0;
      else if (q < lo + 32) w &= (1u << (q - lo)) - 1u;
      rem[l] = w;
    }
    if (up) {   // rem = 2^q - rem
      rem_neg = true;
      long long borrow = 0;
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        const int lo = 32 * l;
        const unsigned long long twoq = (q >= lo && q < lo + 32) ? (1ull << (q - lo)) : 0ull;
        const long long d = (long long)twoq - (long long)rem[l] - borrow;
        borrow = d < 0 ? 1 : 0;
        rem[l] = (unsigned)(d + (borrow << 32));
      }
    }
  }
  int q2;
  bool up2;
  lo = rn_limbs(rem, sc, q2, up2);
  if (rem_neg) lo = -lo;
}

// ---- window path ----------------------------------------------------------
// Both roundings from the 160-bit window W = X >> E, E = 32 (lt - 4), of the
// top limb lt and the four below it (limbs below 0 read as 0), and one bit s0
// for "a set bit below the window".  With b the top bit of limb lt, X's top
// bit is E + 128 + b and hi keeps bits from q = E + 64 + sh, sh = 12 + b, so
// in U = W >> 64 (96 bits, limbs lt .. lt - 2) the mantissa is U >> sh, the
// rounding bit U bit sh - 1, and the sticky bits U's bits below it, limbs
// lt - 3, lt - 4 and s0.  The remainder is V 2^E + R, 0 <= R < 2^E, R != 0
// exactly when s0: V = (U mod 2^sh) 2^64 + limbs lt - 3 : lt - 4 when hi
// rounded down, and 2^(sh + 64) minus that, minus s0, when it rounded up (the
// remainder is then negative).  lo rounds V 2^E at V's top bit pV the same
// way, with s0 as sticky.  The window path applies when neither rounding
// point is clamped by gradual underflow, both lie at bit 1 or above (no
// exact small case) and pV >= 53 (lo's bits inside the window); it returns
// false otherwise, and then the limb path gives the result.  Where it
// applies the two paths agree bit for bit.
VCPHIP_HD bool window_hi_lo(const unsigned (&mag)[NL], int sc, double& hi, double& lo) {
  int lt = -1;
#pragma unroll
  for (int l = 0; l < NL; ++l)
    if (mag[l] != 0) lt = l;
  if (lt < 0) {
    hi = lo = 0.0;
    return true;
  }
  unsigned orp[NL + 1];   // orp[l] = OR of limbs below l
  orp[0] = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) orp[l + 1] = orp[l] | mag[l];
  unsigned w4 = 0, w3 = 0, w2 = 0, w1 = 0, w0 = 0, below = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const bool c = l == lt;
    w4 = c ? mag[l] : w4;
    w3 = c ? (l >= 1 ? mag[l - 1] : 0u) : w3;
    w2 = c ? (l >= 2 ? mag[l - 2] : 0u) : w2;
    w1 = c ? (l >= 3 ? mag[l - 3] : 0u) : w1;
    w0 = c ? (l >= 4 ? mag[l - 4] : 0u) : w0;
    below = c ? (l >= 5 ? orp[l - 4] : 0u) : below;
  }
  const bool s0 = below != 0;
  const int b = 31 - __builtin_clz(w4);
  const int E = 32 * (lt - 4);
  const int sh = 12 + b;
  const int q = E + 64 + sh;   // X's top bit minus 52
  if (q < 1 || q < -1074 - sc) return false;
  const unsigned long long H = ((unsigned long long)w4 << 32) | w3;
  const unsigned long long L = ((unsigned long long)w3 << 32) | w2;   // U's low 64 bits
  const unsigned long long VL = ((unsigned long long)w1 << 32) | w0;
  unsigned long long mant = sh >= 32 ? (H >> (sh - 32)) : ((H << (32 - sh)) | (w2 >> sh));
  const int rp = sh - 1;
  const bool rb = (L >> rp) & 1ull;
  const bool st = (L & ((1ull << rp) - 1ull)) != 0 || VL != 0 || s0;
  const bool up = rb && (st || (mant & 1ull));
  mant += up ? 1ull : 0ull;
  const int ex = q + sc;
  if (ex + (64 - __builtin_clzll(mant)) > 1024) return false;   // overflow: the limb path
  hi = __builtin_ldexp((double)mant, ex);
  // the remainder
  const unsigned long long VH = L & ((1ull << sh) - 1ull);
  u128 V = ((u128)VH << 64) | VL;
  if (up) V = ((u128)1 << (sh + 64)) - V - (s0 ? 1u : 0u);
  const unsigned long long vh = (unsigned long long)(V >> 64), vl = (unsigned long long)V;
  const int pV = vh ? 127 - __builtin_clzll(vh) : (vl ? 63 - __builtin_clzll(vl) : -1);
  if (pV < 53) return false;
  const int d = pV - 52;   // >= 1
  const int q2 = E + d;
  if (q2 < 1 || q2 < -1074 - sc) return false;
  unsigned long long m2 = (unsigned long long)(V >> d);
  const bool rb2 = (unsigned)(V >> (d - 1)) & 1u;
  const bool st2 = (V & (((u128)1 << (d - 1)) - 1u)) != 0 || s0;
  m2 += (rb2 && (st2 || (m2 & 1ull))) ? 1ull : 0ull;
  lo = __builtin_ldexp((double)m2, q2 + sc);
  if (up) lo = -lo;
  return true;
}

}  // namespace vcphip_ddconv

// hi, lo of the double-double nearest X 2^s (X = mag as eight limbs, low first)
VCPHIP_HD void vcphip_dd_from_limbs(const unsigned (&mag)[vcphip_ddconv::NL], int sc, double& hi, double& lo) {
  if (!vcphip_ddconv::window_hi_lo(mag, sc, hi, lo)) vcphip_ddconv::limbs_hi_lo(mag, sc, hi, lo);
}
