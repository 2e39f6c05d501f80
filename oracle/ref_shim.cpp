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
// the reference's fix-ups in its order; built with -frounding-math.  The
// mixed and transposed forms follow the same pattern: mul_im_m (:90-140),
// mul_m_im (:142-196), mulltmm (:296-358) and vmulmm (:198-236), with the
// reference's BLAS calls (pdblas::mulmm / mulltmm, vcp/pdblas.hpp:55-157)
// replaced by rmatmul, whose k-ascending rounded chains are the contract of
// the path (vblas/README.md:167-170).
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


}  // extern "C"

namespace {

// midpoint and radius of `cnt` intervals under the caller's round-up mode:
// [M - R, M + R] contains [l, h] (the oracle's midrad rule)
void midrad_ru(const double* src, size_t cnt, double* M, double* R) {
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
}

// the rows x cols column-major X transposed
std::vector<double> transposed(const double* X, int rows, int cols) {
  std::vector<double> T((size_t)rows * cols);
  for (int j = 0; j < cols; ++j)
    for (int i = 0; i < rows; ++i) T[j + (size_t)cols * i] = X[i + (size_t)rows * j];
  return T;
}

// C[t] = [RD(lo[t] - tt[t]), RU(hi[t] + tt[t])]: addmm_sequential /
// subsmmA_sequential (vcp/pidblas.hpp:61-84)
void fixup(size_t cnt, const double* lo, const double* hi, const double* tt, double* C) {
  const int old = std::fegetround();
  for (size_t t = 0; t < cnt; ++t) {
    std::fesetround(FE_UPWARD);
    C[2 * t + 1] = hi[t] + tt[t];
    std::fesetround(FE_DOWNWARD);
    C[2 * t] = lo[t] - tt[t];
  }
  std::fesetround(old);
}

}  // namespace

extern "C" {

// C = IA * IB, intervals AoS {inf, sup}, column-major with ld = rows:
// lC = RD(mA mB), uC = RU(mA mB), t = RU(RU((|mA| + rA) rB) + RU(rA |mB|)),
// uC = RU(uC + t), lC = RD(lC - t)   (vcp/pidblas.hpp:257-289)
void ref_pidblas_mulmm(int m, int n, int k, const double* IA, const double* IB, double* C) {
  const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
  std::vector<double> mA(mk), rA(mk), sA(mk), mB(kn), rB(kn), aB(kn), lo(mn), hi(mn), p1(mn), p2(mn);
  const int old = std::fegetround();
  std::fesetround(FE_UPWARD);
  auto midrad = midrad_ru;
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

// C = IA * B, B point (vcp/pidblas.hpp:90-140): lC = RD(mA B), uC = RU(mA B),
// t = RU(rA |B|), uC = RU(uC + t), lC = RD(lC - t).  IA m x k, B k x n.
void ref_pidblas_mul_im_m(int m, int n, int k, const double* IA, const double* B, double* C) {
  const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
  std::vector<double> mA(mk), rA(mk), aB(kn), lo(mn), hi(mn), tt(mn);
  const int old = std::fegetround();
  std::fesetround(FE_UPWARD);
  midrad_ru(IA, mk, mA.data(), rA.data());
  std::fesetround(old);
  for (size_t t = 0; t < kn; ++t) aB[t] = std::fabs(B[t]);
  rmatmul(m, k, n, mA.data(), B, lo.data(), -1);
  rmatmul(m, k, n, mA.data(), B, hi.data(), 1);
  rmatmul(m, k, n, rA.data(), aB.data(), tt.data(), 1);
  fixup(mn, lo.data(), hi.data(), tt.data(), C);
}

// C = B * IA, B point (vcp/pidblas.hpp:142-196): lC = RD(B mA), uC = RU(B mA),
// t = RU(|B| rA), then the same fix-ups.  B m x k, IA k x n.
void ref_pidblas_mul_m_im(int m, int n, int k, const double* B, const double* IA, double* C) {
  const size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
  std::vector<double> mA(kn), rA(kn), aB(mk), lo(mn), hi(mn), tt(mn);
  const int old = std::fegetround();
  std::fesetround(FE_UPWARD);
  midrad_ru(IA, kn, mA.data(), rA.data());
  std::fesetround(old);
  for (size_t t = 0; t < mk; ++t) aB[t] = std::fabs(B[t]);
  rmatmul(m, k, n, B, mA.data(), lo.data(), -1);
  rmatmul(m, k, n, B, mA.data(), hi.data(), 1);
  rmatmul(m, k, n, aB.data(), rA.data(), tt.data(), 1);
  fixup(mn, lo.data(), hi.data(), tt.data(), C);
}

// C = IA^T IA (vcp/pidblas.hpp:296-358), IA k x n, C n x n: lC = RD(mA^T mA),
// uC = RU(mA^T mA); temp = RU(|mA|^T rA) and rr = RU(rA^T rA); for i <= j,
// in the reference's loop order, temp(i,j) = RU(temp(i,j) + RU(temp(j,i) +
// rr(i,j))), then the lower triangle mirrors the upper; the fix-ups; and C's
// lower triangle mirrors its upper one.
void ref_pidblas_mulltmm(int k, int n, const double* IA, double* C) {
  const size_t kn = (size_t)k * n, nn = (size_t)n * n;
  std::vector<double> mA(kn), rA(kn), lo(nn), hi(nn), temp(nn), rr(nn);
  const int old = std::fegetround();
  std::fesetround(FE_UPWARD);
  midrad_ru(IA, kn, mA.data(), rA.data());
  std::fesetround(old);
  const std::vector<double> mAt = transposed(mA.data(), k, n);
  std::vector<double> aAt(mAt.size());
  for (size_t t = 0; t < mAt.size(); ++t) aAt[t] = std::fabs(mAt[t]);
  const std::vector<double> rAt = transposed(rA.data(), k, n);
  rmatmul(n, k, n, mAt.data(), mA.data(), lo.data(), -1);
  rmatmul(n, k, n, mAt.data(), mA.data(), hi.data(), 1);
  rmatmul(n, k, n, aAt.data(), rA.data(), temp.data(), 1);
  rmatmul(n, k, n, rAt.data(), rA.data(), rr.data(), 1);
  std::fesetround(FE_UPWARD);
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) temp[i + (size_t)n * j] += temp[j + (size_t)n * i] + rr[i + (size_t)n * j];
  std::fesetround(old);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) temp[j + (size_t)n * i] = temp[i + (size_t)n * j];
  fixup(nn, lo.data(), hi.data(), temp.data(), C);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      C[2 * (j + (size_t)n * i)] = C[2 * (i + (size_t)n * j)];
      C[2 * (j + (size_t)n * i) + 1] = C[2 * (i + (size_t)n * j) + 1];
    }
}

// IC = C^T C for point C (vcp/pidblas.hpp:198-236), C k x n, IC n x n:
// upper ends RU(C^T C), lower ends RD(C^T C), each mirrored from its upper
// triangle as pdblas::mulltmm does (vcp/pdblas.hpp:140-150)
void ref_pidblas_vmulmm(int k, int n, const double* A, double* IC) {
  const size_t nn = (size_t)n * n;
  std::vector<double> up(nn), dn(nn);
  const std::vector<double> At = transposed(A, k, n);
  rmatmul(n, k, n, At.data(), A, up.data(), 1);
  rmatmul(n, k, n, At.data(), A, dn.data(), -1);
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      const size_t u = i + (size_t)n * j, l = j + (size_t)n * i;
      IC[2 * u] = IC[2 * l] = dn[u];
      IC[2 * u + 1] = IC[2 * l + 1] = up[u];
    }
}

}  // extern "C"
