// The window path of the double-double conversion (vcphip_ddconv.h) against
// its exact limb path, bit for bit, on random and structured 256-bit
// integers and scales: every case where the window path applies must give
// the limb path's hi and lo (signed zeros included).
#include <cstdio>
#include <cstring>
#include <random>

#include "../../vcp_amd/csrc/vcphip_ddconv.h"

using vcphip_ddconv::NL;

static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

int main(int argc, char** argv) {
  const long cases = argc > 1 ? std::atol(argv[1]) : 2000000;
  std::mt19937_64 rng(12345);
  long applied = 0, bad = 0;
  for (long c = 0; c < cases; ++c) {
    unsigned mag[NL] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int kind = (int)(rng() % 8);
    const int lt = (int)(rng() % NL);
    for (int l = 0; l <= lt; ++l) mag[l] = (unsigned)rng();
    if (mag[lt] == 0) mag[lt] = 1;
    if (kind == 1) {   // a short integer: zero limbs below a few top ones
      const int keep = 1 + (int)(rng() % 3);
      for (int l = 0; l <= lt - keep; ++l) mag[l] = 0;
    } else if (kind == 2) {   // a tie or near-tie after the top 53 bits
      int top = 32 * lt + 31 - __builtin_clz(mag[lt]);
      const int q = top - 52;
      if (q >= 1) {
        for (int bit = 0; bit < q; ++bit) mag[bit >> 5] &= ~(1u << (bit & 31));
        mag[(q - 1) >> 5] |= 1u << ((q - 1) & 31);
        if (rng() & 1) mag[0] |= (unsigned)(rng() & 1);   // just above the tie
      }
    } else if (kind == 3) {   // a mantissa of ones that rounds up to 2^53
      int top = 32 * lt + 31 - __builtin_clz(mag[lt]);
      for (int bit = top - 60; bit <= top; ++bit)
        if (bit >= 0) mag[bit >> 5] |= 1u << (bit & 31);
    } else if (kind == 4) {   // a remainder with leading zeros (deep lo)
      int top = 32 * lt + 31 - __builtin_clz(mag[lt]);
      const int z = (int)(rng() % 120);
      for (int bit = top - 53 - z; bit <= top - 53; ++bit)
        if (bit >= 0) mag[bit >> 5] &= ~(1u << (bit & 31));
    } else if (kind == 5) {   // single bits
      for (int l = 0; l <= lt; ++l) mag[l] = (rng() % 4 == 0) ? (1u << (rng() % 32)) : 0u;
      if (mag[lt] == 0) mag[lt] = 1u << (rng() % 32);
    } else if (kind == 6) {   // all ones
      for (int l = 0; l <= lt; ++l) mag[l] = 0xFFFFFFFFu;
    } else if (kind == 7 && (rng() % 16) == 0) {   // zero
      for (int l = 0; l < NL; ++l) mag[l] = 0;
    }
    int sc;
    switch (rng() % 4) {
      case 0: sc = -220 + (int)(rng() % 40); break;                 // the product's usual scale
      case 1: sc = -1400 + (int)(rng() % 400); break;               // gradual underflow
      case 2: sc = 700 + (int)(rng() % 400); break;                 // overflow
      default: sc = -1400 + (int)(rng() % 2500); break;
    }
    double h0, l0, h1 = 0.0, l1 = 0.0;
    vcphip_ddconv::limbs_hi_lo(mag, sc, h0, l0);
    if (vcphip_ddconv::window_hi_lo(mag, sc, h1, l1)) {
      ++applied;
      if (!same(h0, h1) || !same(l0, l1)) {
        if (++bad <= 5)
          std::printf("mismatch kind %d lt %d sc %d: limbs %a %a window %a %a\n", kind, lt, sc, h0, l0, h1, l1);
      }
    }
    double h2, l2;
    vcphip_dd_from_limbs(mag, sc, h2, l2);
    if (!same(h0, h2) || !same(l0, l2)) {
      if (++bad <= 5) std::printf("mismatch of the combined conversion kind %d\n", kind);
    }
  }
  std::printf("cases %ld window %ld mismatches %ld\n", cases, applied, bad);
  if (bad == 0) std::printf("ddconv ok\n");
  return bad == 0 ? 0 : 1;
}
