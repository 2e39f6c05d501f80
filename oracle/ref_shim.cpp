// C-ABI shim over the reference's own vblas directed-rounding GEMMs, compiled
// from the reference sources where they lie (include path set by the
// Makefile's ref target).  TEST INFRASTRUCTURE ONLY: the result, in
// oracle/_ref/, pins the oracle's RU/RD chains and serves as a CPU baseline.
//
// Reference interfaces: rmatmul(m, n, k, A, B, C, mode) computes the m x k
// product of A (m x n) and B (n x k), column-major, C overwritten
// (vblas/rmatmul.hpp:100-127); udmatmul(m, n, k, A, B, CU, CD) accumulates
// into CU/CD (vblas/udmatmul.hpp:738-760, AVX-512 only).
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

}
