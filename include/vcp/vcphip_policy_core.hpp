// vcphip_policy_core.hpp -- the part of the vcp::pidhip / vcp::pddhip
// policies that does not depend on kv or on the rest of VCP: the per-thread
// library context, the operand-type dispatch of the reference's BLAS
// wrappers, and the mapping of C-ABI status codes onto exceptions.
//
// Header-only, C++11.  Links against libvcphip (include/vcphip.h).
#pragma once

#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>

#include "../vcphip.h"

namespace vcphip_policy {

// ---- operand types (the vcp::mats type tags) -------------------------------
// 'S' scalar, 'R' row vector, 'C' column vector, 'M' matrix, as set by
// mats::zeros (vcp/mats.hpp:2801-2817).
inline char type_of_shape(int rows, int cols) {
  if (rows == 1 && cols == 1) return 'S';
  if (cols == 1) return 'C';
  if (rows == 1) return 'R';
  if (rows > 1 && cols > 1) return 'M';
  return 'N';
}

// Product types the reference's BLAS wrapper accepts once the scalar cases
// are handled on the host: R*C (ddot), C*R (outer), M*C (gemv), R*M and M*M
// (gemm), with matching inner dimensions (vcp/pdblas.hpp:75-116 and
// vcp/pddblas.hpp:339-360).  Returns 0 when supported, otherwise a
// VCPHIP_ERR_DIMENSION status.
inline int check_product(char ta, int ra, int ca, char tb, int rb, int cb) {
  const bool inner = (ca == rb);
  if (ta == 'R' && tb == 'C' && inner) return VCPHIP_OK;
  if (ta == 'C' && tb == 'R') return VCPHIP_OK;
  if (ta == 'M' && tb == 'C' && inner) return VCPHIP_OK;
  if (ta == 'R' && tb == 'M' && inner) return VCPHIP_OK;
  if (ta == 'M' && tb == 'M' && inner) return VCPHIP_OK;
  (void)ra;
  (void)cb;
  return VCPHIP_ERR_DIMENSION;
}

inline bool is_scalar_case(char ta, char tb) { return ta == 'S' || tb == 'S'; }

// ---- per-thread context -----------------------------------------------------
// One context per host thread (a context owns one stream); the device is
// VCPHIP_DEVICE from the environment, 0 by default.
struct ContextHolder {
  vcphip_ctx* ctx = nullptr;
  int status = VCPHIP_OK;
  ContextHolder() {
    const char* d = std::getenv("VCPHIP_DEVICE");
    status = vcphip_ctx_create(d ? std::atoi(d) : 0, &ctx);
  }
  ~ContextHolder() {
    if (ctx) vcphip_ctx_destroy(ctx);
  }
  ContextHolder(const ContextHolder&) = delete;
  ContextHolder& operator=(const ContextHolder&) = delete;
};

inline vcphip_ctx* context(int* status = nullptr) {
  static thread_local ContextHolder holder;
  if (status) *status = holder.status;
  return holder.ctx;
}

// process-wide choice of the interval x interval form
inline int& default_form() {
  static int form = VCPHIP_FORM_OZAKI;
  return form;
}

// ---- status -> exception ----------------------------------------------------
// Errors::dimension(msg), ::argument(msg), ::backend(routine, code, msg) and
// ::verification(msg) throw the caller's exception types (the VCP policy
// headers bind them to vcp::dimension_error, vcp::invalid_argument,
// vcp::backend_error("hip", ...) and vcp::verification_error, vcp/error.hpp).
template <class Errors>
inline void check(int status, const char* routine) {
  if (status == VCPHIP_OK) return;
  const std::string msg = std::string(routine) + ": " + vcphip_strerror(status);
  switch (status) {
    case VCPHIP_ERR_DIMENSION: Errors::dimension(msg); break;
    case VCPHIP_ERR_ARGUMENT: Errors::argument(msg); break;
    case VCPHIP_ERR_VERIFICATION: Errors::verification(msg); break;
    default: Errors::backend(routine, status, msg); break;
  }
  throw std::runtime_error(msg);   // Errors must throw; never reached otherwise
}

template <class Errors>
inline vcphip_ctx* checked_context() {
  int st = VCPHIP_OK;
  vcphip_ctx* c = context(&st);
  check<Errors>(c ? VCPHIP_OK : (st != VCPHIP_OK ? st : VCPHIP_ERR_DEVICE), "vcphip_ctx_create");
  return c;
}

}  // namespace vcphip_policy
