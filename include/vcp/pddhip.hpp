// pddhip.hpp -- vcp::pddhip, a drop-in kv::dd policy computing the products
// of vcp::pddblas on an MI355X through libvcphip (an Ozaki scheme II on int8
// MFMA in place of the reference's FP64-slice Ozaki scheme, vcp/pddblas.hpp:
// 54-193).
//
//   vcp::matrix< kv::dd, vcp::pddhip > A, B, C;
//   C = A * B;            // GPU
//   C = ltransmul(A);     // A^T A, exactly symmetric
//
// pddblas::mulmm / mulltmm / linearsolve are NOT virtual (vcp/pddblas.hpp:
// 318, :379, :412), so pddhip declares its own, which hide them for
// vcp::matrix<kv::dd, pddhip> (name lookup in matrix's operator* and lss,
// vcp/matrix.hpp:245-267, :607-634); inv() (virtual, :477) is overridden so
// that it reaches pddhip::linearsolve.  linearsolve runs the reference's
// refinement loop on the device (vcphip_dd_solve: LU and corrections by
// rocSOLVER, residuals r = b - A x by the dd Ozaki product); Cholesky, eigsym and eigsymge keep the reference's
// refinements with their dd products on the GPU.  Scalar operands and products below the GPU / host crossover
// (vcphip_policy::crossover, measured by tune_crossover()) keep the
// reference's host semantics.
#pragma once

#ifndef VCP_PDDHIP_HPP
#define VCP_PDDHIP_HPP

#include <cmath>
#include <limits>
#include <random>
#include <vector>

#include <vcp/error.hpp>
#include <vcp/pddblas.hpp>

#include "vcphip_policy_core.hpp"

namespace vcp {

namespace pddhip_detail {
struct Errors {
  [[noreturn]] static void dimension(const std::string& m) { throw vcp::dimension_error(m); }
  [[noreturn]] static void argument(const std::string& m) { throw vcp::invalid_argument(m); }
  [[noreturn]] static void verification(const std::string& m) { throw vcp::verification_error(m); }
  [[noreturn]] static void backend(const char* routine, int code, const std::string& m) {
    throw vcp::backend_error("hip", routine, code, m);
  }
};
static_assert(sizeof(kv::dd) == 2 * sizeof(double), "kv::dd must be {a1, a2}");
inline const double* raw(const std::vector<kv::dd>& v) { return reinterpret_cast<const double*>(v.data()); }
inline double* raw(std::vector<kv::dd>& v) { return reinterpret_cast<double*>(v.data()); }
}  // namespace pddhip_detail

class pddhip : public vcp::pddblas {
  typedef pddhip_detail::Errors E;

 public:
  // Measures where the GPU overtakes the host (pddblas's Ozaki scheme) for
  // square dd products n = 8 ... max_n, installs that crossover and returns
  // it (an m*n*k value).
  static long long tune_crossover(int max_n = 256, int reps = 3) {
    pddhip A, B, C;
    std::mt19937_64 rng(0x5EED0000ULL);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    auto fill = [&](pddhip& X) {
      for (size_t i = 0; i < X.v.size(); ++i) X.v[i] = kv::dd(U(rng));
    };
    auto setup = [&](int n) {
      A.zeros(n, n);
      B.zeros(n, n);
      C.zeros(n, n);
      fill(A);
      fill(B);
    };
    auto host = [&]() { A.vcp::pddblas::mulmm(B, C); };
    auto gpu = [&]() {
      const int n = A.row;
      vcphip_policy::check<E>(vcphip_ddgemm(vcphip_policy::checked_context<E>(), n, n, n, pddhip_detail::raw(A.v), n,
                                            pddhip_detail::raw(B.v), n, pddhip_detail::raw(C.v), n),
                              "vcphip_ddgemm");
    };
    const long long x = vcphip_policy::measure_crossover(setup, host, gpu, max_n, reps);
    vcphip_policy::crossover(vcphip_policy::CROSS_DD) = x;
    return x;
  }

  // C = A*B (pddblas.hpp:318-376)
  void mulmm(const pddhip& B, pddhip& c) const {
    if (vcphip_policy::is_scalar_case(this->type, B.type) ||
        vcphip_policy::on_host(vcphip_policy::CROSS_DD, this->row, B.column, this->column)) {
      vcp::pddblas::mulmm(B, c);
      return;
    }
    vcphip_policy::check<E>(vcphip_policy::check_product(this->type, this->row, this->column, B.type, B.row, B.column),
                            "mulmm");
    const int m = this->row, k = this->column, n = B.column;
    c.zeros(m, n);
    vcphip_policy::check<E>(vcphip_ddgemm(vcphip_policy::checked_context<E>(), m, n, k, pddhip_detail::raw(this->v), m,
                                          pddhip_detail::raw(B.v), k, pddhip_detail::raw(c.v), m),
                            "vcphip_ddgemm");
  }
  // x = A \ b (pddblas.hpp:412-474): LU of the high parts by LAPACK on the
  // host and x += the correction in dd, as the reference; the residual
  // r = b - A x comes from the GPU (the Ozaki-II dd product and a
  // double-double subtraction) in place of split_row/split_col/ozaki_acc.
  void linearsolve(vcp::pddblas& b, vcp::pddblas& x) {
    if (this->row != this->column || b.row != this->row) {
      vcp::throw_error<vcp::dimension_error>("linearsolve: invalid dimensions: A=(", this->row, ", ", this->column,
                                             "), b=(", b.row, ", ", b.column, ")");
    }
    const int an = this->row, nrhs = b.column;
    if (vcphip_policy::on_host(vcphip_policy::CROSS_DD, an, nrhs, an)) {
      vcp::pddblas::linearsolve(b, x);
      return;
    }
    // the whole refinement on the device (vcphip_dd_solve): LU and the
    // corrections by rocSOLVER getrf/getrs for LAPACK's dgetrf/dgetrs, the
    // residuals by the dd Ozaki product, the reference's itmax and stopping rule
    x.zeros(an, nrhs);
    const int rc = vcphip_dd_solve(vcphip_policy::checked_context<E>(), an, nrhs, pddhip_detail::raw(this->v), an,
                                   pddhip_detail::raw(b.v), an, pddhip_detail::raw(x.v), an, 10, nullptr);
    if (rc == VCPHIP_ERR_SINGULAR)
      throw vcp::lapack_error("rocsolver_dgetrf", VCPHIP_ERR_SINGULAR, "linearsolve: zero pivot in the LU of A");
    vcphip_policy::check<E>(rc, "vcphip_dd_solve");
  }

  // A = inv(A): A X = I through pddhip::linearsolve (pddblas.hpp:477-486)
  void inv() override {
    if (this->row != this->column)
      vcp::throw_error<vcp::dimension_error>("inv: matrix must be square: ", this->row, " != ", this->column);
    pddhip b, x;
    b.eye(this->row);
    this->linearsolve(b, x);
    this->v = std::move(x.v);
  }

  // Upper Cholesky factor (pddblas.hpp:488-563): dpotrf of the high parts on
  // the host, then two refinement steps in which E = A - U^T U and F = W U
  // (W = triu(G, 1) + diag(G) / 2, G = Ud^-T E Ud^-1 by dtrsm) are GPU dd
  // products; U += F in kv::dd.
  void Cholesky() override {
    if (!this->is_symmetric()) vcp::throw_error<vcp::domain_error>("Cholesky: matrix must be symmetric");
    const int an = this->row, nn = an * an;
    if (vcphip_policy::on_host(vcphip_policy::CROSS_DD, an, an, an)) {
      vcp::pddblas::Cholesky();
      return;
    }
    // the whole refinement on the device (vcphip_dd_cholesky): dpotrf and
    // dtrsm by rocSOLVER / rocBLAS for LAPACK / BLAS, E = A - U^T U and W U
    // by the dd Ozaki product, the reference's two rounds
    std::vector<kv::dd> U(nn);
    const int rc = vcphip_dd_cholesky(vcphip_policy::checked_context<E>(), an, pddhip_detail::raw(this->v), an,
                                      pddhip_detail::raw(U), an, 2);
    if (rc == VCPHIP_ERR_SINGULAR)
      throw vcp::lapack_error("rocsolver_dpotrf", VCPHIP_ERR_SINGULAR, "Cholesky: matrix not positive definite");
    vcphip_policy::check<E>(rc, "vcphip_dd_cholesky");
    this->v = std::move(U);
  }

  // Symmetric eigenproblem (pddblas.hpp:566-606), every step on the device
  // (vcphip_dd_eigsym): dsyev of the high parts by rocSOLVER, then
  // max(itep, 2) Ogita-Aishima steps (:219-314) with dd products and the
  // elementwise double-double updates on the GPU.
  void eigsym(int itep = 4) override {
    vcp::pddblas V;
    this->eigsym(V, itep);
  }
  void eigsym(vcp::pddblas& V, int itep = 4) {
    if (!this->is_symmetric()) vcp::throw_error<vcp::domain_error>("eigsym: matrix must be symmetric");
    const int an = this->row, nn = an * an;
    if (vcphip_policy::on_host(vcphip_policy::CROSS_DD, an, an, an)) {
      vcp::pddblas::eigsym(V, itep);
      return;
    }
    std::vector<kv::dd> X(nn), lam(an);
    eig_result(vcphip_dd_eigsym(vcphip_policy::checked_context<E>(), an, pddhip_detail::raw(this->v), an,
                                pddhip_detail::raw(lam), pddhip_detail::raw(X), an, itep),
               "vcphip_dd_eigsym");
    V.zeros(an, an);
    V.v = std::move(X);
    this->zeros(an, an);
    for (int i = 0; i < an; ++i) this->v[i + an * i] = lam[i];
  }

  // Generalized problem A x = lambda B x (pddblas.hpp:611-659), every step on
  // the device (vcphip_dd_eigsymge): dsygv of the high parts by rocSOLVER,
  // then max(itep, 2) Ogita-Aishima steps with the B-inner product.
  void eigsymge(vcp::pddblas& B, int itep = 1) {
    vcp::pddblas V;
    this->eigsymge(B, V, itep);
  }
  void eigsymge(vcp::pddblas& B, vcp::pddblas& V, int itep = 1) {
    if (!this->is_symmetric()) vcp::throw_error<vcp::domain_error>("eigsymge: matrix A must be symmetric");
    if (!B.is_symmetric()) vcp::throw_error<vcp::domain_error>("eigsymge: matrix B must be symmetric");
    if (this->row != B.row || this->column != B.column) {
      vcp::throw_error<vcp::dimension_error>("eigsymge: dimension mismatch: A=(", this->row, ", ", this->column,
                                             "), B=(", B.row, ", ", B.column, ")");
    }
    const int an = this->row, nn = an * an;
    if (vcphip_policy::on_host(vcphip_policy::CROSS_DD, an, an, an)) {
      vcp::pddblas::eigsymge(B, V, itep);
      return;
    }
    std::vector<kv::dd> X(nn), lam(an);
    eig_result(vcphip_dd_eigsymge(vcphip_policy::checked_context<E>(), an, pddhip_detail::raw(this->v), an,
                                  pddhip_detail::raw(B.v), an, pddhip_detail::raw(lam), pddhip_detail::raw(X), an, itep),
               "vcphip_dd_eigsymge");
    V.zeros(an, an);
    V.v = std::move(X);
    this->zeros(an, an);
    for (int i = 0; i < an; ++i) this->v[i + an * i] = lam[i];
  }

 private:
  // a failed dsyev / dsygv start is the reference's lapack_error
  // (pddblas.hpp:576-579, :627-630); other codes map as everywhere else
  static void eig_result(int rc, const char* routine) {
    if (rc == VCPHIP_ERR_SINGULAR)
      throw vcp::lapack_error(routine, VCPHIP_ERR_SINGULAR, "eigsym: the double-precision start failed");
    vcphip_policy::check<E>(rc, routine);
  }

 public:
  // C = A^T A (pddblas.hpp:379-409), exactly symmetric
  void mulltmm(pddhip& c) const {
    if (this->type == 'S' || vcphip_policy::on_host(vcphip_policy::CROSS_DD, this->column, this->column, this->row)) {
      vcp::pddblas::mulltmm(c);
      return;
    }
    if (this->type != 'C' && this->type != 'R' && this->type != 'M')
      vcp::throw_error<vcp::state_error>("mulltmm: unsupported matrix type: ", this->type);
    const int k = this->row, n = this->column;
    c.zeros(n, n);
    vcphip_policy::check<E>(vcphip_ddgemm_tn(vcphip_policy::checked_context<E>(), k, n, pddhip_detail::raw(this->v), k,
                                             pddhip_detail::raw(c.v), n),
                            "vcphip_ddgemm_tn");
  }
};

}  // namespace vcp

#endif  // VCP_PDDHIP_HPP
