// pidhip.hpp -- vcp::pidhip, a drop-in interval<double> policy computing the
// verified products of vcp::pidblas on an MI355X through libvcphip.
//
//   #include <kv/interval.hpp>
//   #include <kv/rdouble.hpp>
//   #include <vcp/pidhip.hpp>          // this header (include/ of vcp_amd)
//   vcp::matrix< kv::interval<double>, vcp::pidhip > A, B, C;
//   C = A * B;                          // GPU, enclosure of the exact product
//   x = lss(A, b);                      // verified solve: vcphip_lss
//
// pidhip_t<Base> derives from an interval policy Base (vcp::pidblas,
// vcp/pidblas.hpp:87, for vcp::pidhip; vcp::pidblas_fma for vcp::pidhip_fma,
// pidhip_fma.hpp) and overrides the hooks of pidblas: the public virtual
// mulmm (:239) and mulltmm (:296), and the protected virtual mul_im_m (:90),
// mul_m_im (:142) and vmulmm (:198) that imats::linearsolve / eigsym /
// eigsymge call (vcp/imats.hpp:293-522).  Scalar ('S') operands keep Base's
// host semantics (:241-254), and so do products below the GPU / host
// crossover (vcphip_policy::crossover, measured by tune_crossover()); every
// other product runs on the GPU.  The host floating-point environment is
// never changed (the reference leaves it at nearest via hwround_guard,
// vcp/hwround_guard.hpp:39-48).  Errors map onto vcp/error.hpp.
//
// Requires the VCP and kv headers on the include path and linking with
// -lvcphip.  vcp::pidhip::set_form() selects the interval x interval form
// (VCPHIP_FORM_OZAKI, the default, VCPHIP_FORM_PLAIN or VCPHIP_FORM_INFSUP).
#pragma once

#ifndef VCP_PIDHIP_HPP
#define VCP_PIDHIP_HPP

#include <cmath>
#include <iostream>
#include <random>

#include <vcp/error.hpp>
#include <vcp/pidblas.hpp>

#include "vcphip_policy_core.hpp"

namespace vcp {

namespace pidhip_detail {
struct Errors {
  [[noreturn]] static void dimension(const std::string& m) { throw vcp::dimension_error(m); }
  [[noreturn]] static void argument(const std::string& m) { throw vcp::invalid_argument(m); }
  [[noreturn]] static void verification(const std::string& m) { throw vcp::verification_error(m); }
  [[noreturn]] static void backend(const char* routine, int code, const std::string& m) {
    throw vcp::backend_error("hip", routine, code, m);
  }
};
inline const double* raw(const std::vector<kv::interval<double> >& v) {
  return reinterpret_cast<const double*>(v.data());   // AoS {inf, sup}
}
inline double* raw(std::vector<kv::interval<double> >& v) { return reinterpret_cast<double*>(v.data()); }
static_assert(sizeof(kv::interval<double>) == 2 * sizeof(double), "kv::interval<double> must be {inf, sup}");
}  // namespace pidhip_detail

// FixedForm >= 0 pins the interval x interval form (pidhip_fma); otherwise
// it is vcphip_policy::default_form() (set_form).
template <class Base, int FixedForm = -1>
class pidhip_t : public Base {
  typedef pidhip_detail::Errors E;
  static int form() { return FixedForm >= 0 ? FixedForm : vcphip_policy::default_form(); }
  // crossover of this policy's interval x interval product (the inf-sup host
  // product of pidblas_fma is far slower than pidblas's BLAS one)
  static int cross_kind() {
    return FixedForm == VCPHIP_FORM_INFSUP ? vcphip_policy::CROSS_INFSUP : vcphip_policy::CROSS_INTERVAL;
  }

 protected:
  // C = IA * B (pidblas.hpp:90-140)
  void mul_im_m(const vcp::pdblas& B, vcp::imats<double, vcp::pdblas>& c) const override {
    if (vcphip_policy::is_scalar_case(this->type, B.type) ||
        vcphip_policy::on_host(vcphip_policy::CROSS_INTERVAL, this->row, B.column, this->column)) {
      Base::mul_im_m(B, c);
      return;
    }
    vcphip_policy::check<E>(vcphip_policy::check_product(this->type, this->row, this->column, B.type, B.row, B.column),
                            "mul_im_m");
    const int m = this->row, k = this->column, n = B.column;
    c.zeros(m, n);
    vcphip_policy::check<E>(vcphip_igemm_ip(vcphip_policy::checked_context<E>(), m, n, k, pidhip_detail::raw(this->v), m,
                                            B.v.data(), k, pidhip_detail::raw(c.v), m),
                            "vcphip_igemm_ip");
  }
  // C = B * IA (pidblas.hpp:142-196)
  void mul_m_im(const vcp::pdblas& B, vcp::imats<double, vcp::pdblas>& c) const override {
    if (vcphip_policy::is_scalar_case(B.type, this->type) ||
        vcphip_policy::on_host(vcphip_policy::CROSS_INTERVAL, B.row, this->column, B.column)) {
      Base::mul_m_im(B, c);
      return;
    }
    vcphip_policy::check<E>(vcphip_policy::check_product(B.type, B.row, B.column, this->type, this->row, this->column),
                            "mul_m_im");
    const int m = B.row, k = B.column, n = this->column;
    c.zeros(m, n);
    vcphip_policy::check<E>(vcphip_igemm_pi(vcphip_policy::checked_context<E>(), m, n, k, B.v.data(), m,
                                            pidhip_detail::raw(this->v), k, pidhip_detail::raw(c.v), m),
                            "vcphip_igemm_pi");
  }
  // IC = C^T C with verification (pidblas.hpp:198-236)
  void vmulmm(const vcp::pdblas& C) override {
    if (C.type == 'S' || C.type == 'C' ||
        vcphip_policy::on_host(vcphip_policy::CROSS_INTERVAL, C.column, C.column, C.row)) {
      Base::vmulmm(C);
      return;
    }
    const int k = C.row, n = C.column;
    this->zeros(n, n);
    vcphip_policy::check<E>(vcphip_vgemm_tn(vcphip_policy::checked_context<E>(), k, n, C.v.data(), k,
                                            pidhip_detail::raw(this->v), n),
                            "vcphip_vgemm_tn");
  }

 public:
  static void set_form(int form) { vcphip_policy::default_form() = form; }

  // Measures where the GPU overtakes the host (the reference's BLAS path) for
  // square interval products n = 8 ... max_n in the current form, installs
  // that crossover and returns it (an m*n*k value).
  static long long tune_crossover(int max_n = 512, int reps = 3) {
    pidhip_t A, B, C;
    std::mt19937_64 rng(0x5EED0000ULL);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    auto fill = [&](pidhip_t& X) {
      for (size_t i = 0; i < X.v.size(); ++i) {
        const double m = U(rng), r = std::ldexp(std::fabs(m), -40);
        X.v[i] = kv::interval<double>(m - r, m + r);
      }
    };
    auto setup = [&](int n) {
      A.zeros(n, n);
      B.zeros(n, n);
      C.zeros(n, n);
      fill(A);
      fill(B);
    };
    auto host = [&]() { A.Base::mulmm(B, C); };
    auto gpu = [&]() {
      const int n = A.row;
      vcphip_policy::check<E>(vcphip_igemm(vcphip_policy::checked_context<E>(), form(), n, n, n,
                                           pidhip_detail::raw(A.v), n, pidhip_detail::raw(B.v), n,
                                           pidhip_detail::raw(C.v), n),
                              "vcphip_igemm");
    };
    const long long x = vcphip_policy::measure_crossover(setup, host, gpu, max_n, reps);
    vcphip_policy::crossover(cross_kind()) = x;
    return x;
  }

  // C = IA * IB (pidblas.hpp:239-294)
  void mulmm(const vcp::mats<kv::interval<double> >& B, vcp::mats<kv::interval<double> >& c) const override {
    if (vcphip_policy::is_scalar_case(this->type, B.type) ||
        vcphip_policy::on_host(cross_kind(), this->row, B.column, this->column)) {
      Base::mulmm(B, c);
      return;
    }
    vcphip_policy::check<E>(vcphip_policy::check_product(this->type, this->row, this->column, B.type, B.row, B.column),
                            "mulmm");
    const int m = this->row, k = this->column, n = B.column;
    c.zeros(m, n);
    vcphip_policy::check<E>(vcphip_igemm(vcphip_policy::checked_context<E>(), form(), m, n, k,
                                         pidhip_detail::raw(this->v), m, pidhip_detail::raw(B.v), k,
                                         pidhip_detail::raw(c.v), m),
                            "vcphip_igemm");
  }
  // x = A \ b with verification (imats::linearsolve, vcp/imats.hpp:292-362),
  // every step on the GPU in one call (vcphip_lss with R = NULL): R =
  // inv(mid A) by rocSOLVER getrf/getrs in place of imid + pdblas::inv's
  // LAPACK (:298-301), then pidblas's products and roundings throughout.  Systems of order < 2,
  // those below the crossover and pidhip_fma (whose |RA - I| * T product is
  // the inf-sup one) keep Base::linearsolve, which still reaches the GPU
  // through the hooks above.
  void linearsolve(vcp::imats<double, vcp::pdblas>& b, vcp::imats<double, vcp::pdblas>& x) {
    if (this->row != this->column || b.row != this->row) {
      vcp::throw_error<vcp::dimension_error>("linearsolve: invalid dimensions: A=(", this->row, ", ", this->column,
                                             "), b=(", b.row, ", ", b.column, ")");
    }
    const int n = this->row, nrhs = b.column;
    if (FixedForm >= 0 || n < 2 || nrhs < 1 || vcphip_policy::on_host(vcphip_policy::CROSS_INTERVAL, n, n, n)) {
      Base::linearsolve(b, x);
      return;
    }
    x.zeros(n, nrhs);
    double g = 0.0;
    const int rc = vcphip_lss(vcphip_policy::checked_context<E>(), n, nrhs, pidhip_detail::raw(this->v), n,
                              pidhip_detail::raw(b.v), n, nullptr, n, pidhip_detail::raw(x.v), n, &g);
    if (rc == VCPHIP_ERR_SINGULAR)   // pdblas::inv's dgetrf failure (vcp/pdblas.hpp:191-193)
      throw vcp::lapack_error("rocsolver_dgetrf", VCPHIP_ERR_SINGULAR, "inv: zero pivot in the LU of mid(A)");
    if (rc == VCPHIP_OK || rc == VCPHIP_ERR_VERIFICATION)
      std::cout << "Linear solver Check || I - RA || < 1: " << g << " < 1 ?" << std::endl;   // as :329
    if (rc == VCPHIP_ERR_VERIFICATION)
      vcp::throw_error<vcp::verification_error>("linearsolve verification failed: ||RA - I|| <= ", g);
    vcphip_policy::check<E>(rc, "vcphip_lss");
  }

  // C = IA^T IA (pidblas.hpp:296-358)
  void mulltmm(vcp::mats<kv::interval<double> >& c) const override {
    if (this->type == 'S' || this->type == 'C' ||
        vcphip_policy::on_host(vcphip_policy::CROSS_INTERVAL, this->column, this->column, this->row)) {
      Base::mulltmm(c);
      return;
    }
    if (this->type != 'M' && this->type != 'R')
      vcp::throw_error<vcp::state_error>("mulltmm: unsupported matrix type: ", this->type);
    const int k = this->row, n = this->column;
    c.zeros(n, n);
    vcphip_policy::check<E>(vcphip_igemm_tn(vcphip_policy::checked_context<E>(), k, n, pidhip_detail::raw(this->v), k,
                                            pidhip_detail::raw(c.v), n),
                            "vcphip_igemm_tn");
  }
};

typedef pidhip_t<vcp::pidblas> pidhip;

}  // namespace vcp

#endif  // VCP_PIDHIP_HPP
