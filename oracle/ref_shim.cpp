// C-ABI shim over the reference's own vblas directed-rounding GEMMs, compiled
// from the reference sources where they lie (include path set by the
// Makefile's ref target).  TEST INFRASTRUCTURE ONLY: the result, in
// oracle/_ref/, pins the oracle's RU/RD chains and serves as a CPU baseline.
//
// Reference interfaces: rmatmul(m, n, k, A, B, C, mode) computes the m x k
// product of A (m x n) and B (n x k), column-major, C overwritten
// (vblas/rmatmul.hpp:100-127); udmatmul(m, n, k, A, B, CU, CD) accumulates
// into CU/CD (vblas/udmatmul.hpp:738-760, AVX-512 only).
//
// ref_pidblas_mulmm restates pidblas::mulmm (vcp/pidblas.hpp:239-294) with
// these rmatmul calls for its four products, the oracle's midrad rule and
// the reference's fix-ups in its order; built with -frounding-math.
#include <cfenv>
#include <cmath>
#include <vector>

#include "rmatmul.hpp"
#if defined(__AVX512F__)
#include "udmatmul.hpp"
#endif

extern "C" {

int ref_backend(void) { return (int)vblas_rmatmul_detail::selected_backend(); }

void ref_rmatmul(int m, int n, int k, const double* A, const double* B, double* C, int mode) {
  rmatmul(m, n, k, A, B, C, mode);
}

int ref_udmatmul(int m, int n, int k, double* A, double* B, double* CU, double* CD) {
#if defined(__AVX512F__)
  udmatmul(m, n, k, A, B, CU, CD);
  return 0;
#else
  (void)m; (void)n; (void)k; (void)A; (void)B; (void)CU; (void)CD;
  return -1;
#endif
}


// C = IA * IB, intervals AoS {inf, sup}, column-major with ld = rows:
// lC = RD(mA mB), uC = RU(mA mB), t = RU(RU((|mA| + rA) rB) + RU(rA |mB|)),
// uC = RU(uC + t), lC = RD(lC - t)   (vcp/pidblas.hpp:257-289)
void ref_pidblas_mulmm(int m, int n, int k, const double* IA, const double* IB, double* C) {
  const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
  std::vector<double> mA(mk), rA(mk), sA(mk), mB(kn), rB(kn), aB(kn), lo(mn), hi(mn), p1(mn), p2(mn);
  const int old = std::fegetround();
  std::fesetround(FE_UPWARD);
  auto midrad = [](const double* src, size_t cnt, double* M, double* R) {
    for (size_t t = 0; t < cnt; ++t) {
      const double l = src[2 * t], h = src[2 * t + 1];
      if (l == h) {
        M[t] = l;
        R[t] = 0.0;
        continue;
      }
      const double c = 0.5 * l + 0.5 * h;
      const double r1 = h - c, r2 = c - l;
      M[t] = c;
      R[t] = r1 > r2 ? r1 : r2;
    }
  };
  midrad(IA, mk, mA.data(), rA.data());
  midrad(IB, kn, mB.data(), rB.data());
  for (size_t t = 0; t < mk; ++t) sA[t] = std::fabs(mA[t]) + rA[t];
  for (size_t t = 0; t < kn; ++t) aB[t] = std::fabs(mB[t]);
  std::fesetround(old);
  rmatmul(m, k, n, mA.data(), mB.data(), lo.data(), -1);
  rmatmul(m, k, n, mA.data(), mB.data(), hi.data(), 1);
  rmatmul(m, k, n, sA.data(), rB.data(), p1.data(), 1);
  rmatmul(m, k, n, rA.data(), aB.data(), p2.data(), 1);
  for (size_t t = 0; t < mn; ++t) {
    std::fesetround(FE_UPWARD);
    const double tt = p1[t] + p2[t];
    C[2 * t + 1] = hi[t] + tt;
    std::fesetround(FE_DOWNWARD);
    C[2 * t] = lo[t] - tt;
  }
  std::fesetround(old);
}

}
