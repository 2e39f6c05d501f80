// vcphip_policy_core.hpp -- the part of the vcp::pidhip / vcp::pddhip
// policies that does not depend on kv or on the rest of VCP: the per-thread
// library context, the operand-type dispatch of the reference's BLAS
// wrappers, and the mapping of C-ABI status codes onto exceptions.
//
// Header-only, C++11.  Links against libvcphip (include/vcphip.h).
#pragma once

#include <chrono>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

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

// ---- GPU / host crossover ----------------------------------------------------
// Products with m*n*k below the crossover of their kind run on the host
// through the base policy (the reference's host path): there the staging
// over PCIe and the launch latency exceed the work.  The environment
// variables VCPHIP_CROSSOVER (pidblas products), VCPHIP_CROSSOVER_FMA
// (pidblas_fma's inf-sup product) and VCPHIP_CROSSOVER_DD (dd products) set
// it as an m*n*k value; 0 sends everything to the GPU.  The policies'
// tune_crossover() measures it with measure_crossover() below.
enum CrossoverKind { CROSS_INTERVAL = 0, CROSS_DD = 1, CROSS_INFSUP = 2 };

inline long long env_crossover(const char* name, long long dflt) {
  const char* e = std::getenv(name);
  return e ? std::atoll(e) : dflt;
}

inline long long& crossover(int kind) {
  static long long v[3] = {env_crossover("VCPHIP_CROSSOVER", 64LL * 64 * 64),
                           env_crossover("VCPHIP_CROSSOVER_DD", 32LL * 32 * 32),
                           env_crossover("VCPHIP_CROSSOVER_FMA", 16LL * 16 * 16)};
  return v[(kind >= 0 && kind <= 2) ? kind : 0];
}

inline bool on_host(int kind, long long m, long long n, long long k) { return m * n * k < crossover(kind); }

// best wall time of `reps` runs of f()
template <class F>
inline double best_time(F& f, int reps) {
  typedef std::chrono::steady_clock clk;
  double best = 0.0;
  for (int r = 0; r < reps; ++r) {
    const clk::time_point t0 = clk::now();
    f();
    const double t = std::chrono::duration<double>(clk::now() - t0).count();
    if (r == 0 || t < best) best = t;
  }
  return best;
}

// For n = 8, 16, ..., max_n: setup(n) prepares an n x n x n product, then
// host() and gpu() each run it `reps` times (best time kept).  Returns the
// smallest n^3 from which the GPU is faster at that and every larger n tried,
// or max_n^3 + 1 when it is not faster at max_n.
template <class Setup, class Host, class Gpu>
long long measure_crossover(Setup setup, Host host, Gpu gpu, int max_n = 512, int reps = 3) {
  std::vector<int> ns;
  std::vector<bool> gpu_wins;
  for (int n = 8; n <= max_n; n *= 2) {
    setup(n);
    const double th = best_time(host, reps), tg = best_time(gpu, reps);
    ns.push_back(n);
    gpu_wins.push_back(tg < th);
  }
  size_t first = ns.size();
  while (first > 0 && gpu_wins[first - 1]) --first;
  if (first == ns.size()) return (long long)max_n * max_n * max_n + 1;
  const long long n0 = ns[first];
  return n0 * n0 * n0;
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
