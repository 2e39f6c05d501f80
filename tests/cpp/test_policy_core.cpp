// C++ test of the policy core (include/vcp/vcphip_policy_core.hpp) through the
// C ABI: operand-type dispatch of the reference's BLAS wrappers, status ->
// exception mapping, the per-thread context, and one interval product per form
// checked bit for bit against the CPU oracle.
//   ./test_policy_core          logic only (no GPU)
//   ./test_policy_core gpu      also the device part
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vcp/vcphip_policy_core.hpp"
#include "../../oracle/oracle.h"

namespace {
int failures = 0;
#define EXPECT(cond)                                                  \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

struct dimension_error : std::runtime_error { using std::runtime_error::runtime_error; };
struct argument_error : std::runtime_error { using std::runtime_error::runtime_error; };
struct verification_error : std::runtime_error { using std::runtime_error::runtime_error; };
struct backend_error : std::runtime_error {
  int code;
  backend_error(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};
struct Errors {
  [[noreturn]] static void dimension(const std::string& m) { throw dimension_error(m); }
  [[noreturn]] static void argument(const std::string& m) { throw argument_error(m); }
  [[noreturn]] static void verification(const std::string& m) { throw verification_error(m); }
  [[noreturn]] static void backend(const char*, int code, const std::string& m) { throw backend_error(m, code); }
};

void test_logic() {
  using namespace vcphip_policy;
  EXPECT(type_of_shape(1, 1) == 'S' && type_of_shape(5, 1) == 'C' && type_of_shape(1, 5) == 'R' &&
         type_of_shape(3, 4) == 'M');
  EXPECT(check_product('M', 3, 4, 'M', 4, 5) == VCPHIP_OK);
  EXPECT(check_product('M', 3, 4, 'M', 5, 5) == VCPHIP_ERR_DIMENSION);
  EXPECT(check_product('R', 1, 4, 'C', 4, 1) == VCPHIP_OK);
  EXPECT(check_product('C', 4, 1, 'R', 1, 6) == VCPHIP_OK);
  EXPECT(check_product('M', 3, 4, 'C', 4, 1) == VCPHIP_OK);
  EXPECT(check_product('R', 1, 4, 'M', 4, 7) == VCPHIP_OK);
  EXPECT(check_product('C', 4, 1, 'M', 1, 7) == VCPHIP_ERR_DIMENSION);   // unsupported in the reference
  EXPECT(check_product('M', 3, 4, 'R', 1, 4) == VCPHIP_ERR_DIMENSION);
  EXPECT(is_scalar_case('S', 'M') && is_scalar_case('M', 'S') && !is_scalar_case('M', 'M'));
  bool thrown = false;
  try { check<Errors>(VCPHIP_ERR_DIMENSION, "x"); } catch (const dimension_error&) { thrown = true; }
  EXPECT(thrown);
  thrown = false;
  try { check<Errors>(VCPHIP_ERR_OOM, "x"); } catch (const backend_error& e) { thrown = e.code == VCPHIP_ERR_OOM; }
  EXPECT(thrown);
  thrown = false;
  try { check<Errors>(VCPHIP_ERR_VERIFICATION, "x"); } catch (const verification_error&) { thrown = true; }
  EXPECT(thrown);
  check<Errors>(VCPHIP_OK, "x");
}

void test_gpu() {
  using namespace vcphip_policy;
  vcphip_ctx* ctx = checked_context<Errors>();
  EXPECT(ctx != nullptr);
  EXPECT(context() == ctx);   // one context per thread
  const int m = 37, k = 91, n = 29;
  std::mt19937_64 rng(0x5EED);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<double> A(2 * m * k), B(2 * k * n), C(2 * m * n), R(2 * m * n);
  auto fill = [&](std::vector<double>& X) {
    for (size_t i = 0; i < X.size(); i += 2) {
      const double mid = std::ldexp(u(rng), (int)(rng() % 41) - 20), rad = std::ldexp(std::fabs(mid), -45);
      X[i] = mid - rad;
      X[i + 1] = mid + rad;
    }
  };
  fill(A);
  fill(B);
  check<Errors>(vcphip_igemm(ctx, VCPHIP_FORM_PLAIN, m, n, k, A.data(), m, B.data(), k, C.data(), m), "plain");
  orc_pidblas_mulmm(m, n, k, A.data(), m, B.data(), k, R.data(), m);
  EXPECT(std::memcmp(C.data(), R.data(), sizeof(double) * C.size()) == 0 || C == R);
  check<Errors>(vcphip_igemm(ctx, VCPHIP_FORM_OZAKI, m, n, k, A.data(), m, B.data(), k, C.data(), m), "ozaki");
  orc_ozaki_igemm(m, n, k, A.data(), m, B.data(), k, R.data(), m, 16);
  EXPECT(C == R);
  bool thrown = false;
  try {
    check<Errors>(vcphip_igemm(ctx, VCPHIP_FORM_PLAIN, m, n, k, A.data(), m - 1, B.data(), k, C.data(), m), "bad lda");
  } catch (const dimension_error&) {
    thrown = true;
  }
  EXPECT(thrown);
}
}  // namespace

// busy-wait for `ns` nanoseconds (a stand-in for work of known cost)
void spin_ns(double ns) {
  const auto t0 = std::chrono::steady_clock::now();
  while (std::chrono::duration<double, std::nano>(std::chrono::steady_clock::now() - t0).count() < ns) {
  }
}

void test_crossover() {
  using namespace vcphip_policy;
  int cur = 0;
  auto setup = [&cur](int n) { cur = n; };
  // host work n^3 ns, GPU a constant 150 us: the GPU first wins at n = 64
  auto host = [&cur]() { spin_ns((double)cur * cur * cur); };
  auto gpu_const = []() { spin_ns(150e3); };
  EXPECT(measure_crossover(setup, host, gpu_const, 128, 3) == 64LL * 64 * 64);
  // a GPU that is always twice as slow never wins
  auto gpu_slow = [&cur]() { spin_ns(2.0 * cur * cur * cur); };
  EXPECT(measure_crossover(setup, host, gpu_slow, 64, 2) == 64LL * 64 * 64 + 1);
  const long long saved = crossover(CROSS_INTERVAL);
  crossover(CROSS_INTERVAL) = 1000;
  EXPECT(on_host(CROSS_INTERVAL, 9, 10, 10) && !on_host(CROSS_INTERVAL, 10, 10, 10));
  crossover(CROSS_INTERVAL) = 0;
  EXPECT(!on_host(CROSS_INTERVAL, 1, 1, 1));
  crossover(CROSS_INTERVAL) = saved;
  EXPECT(crossover(CROSS_DD) >= 0);
}

int main(int argc, char** argv) {
  test_logic();
  test_crossover();
  if (argc > 1 && std::string(argv[1]) == "gpu") test_gpu();
  if (failures) {
    std::fprintf(stderr, "%d failure(s)\n", failures);
    return 1;
  }
  std::printf("policy core ok%s\n", argc > 1 ? " (with gpu)" : "");
  return 0;
}
