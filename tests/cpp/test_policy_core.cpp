// C++ test of the policy core (include/vcp/vcphip_policy_core.hpp) through the
// C ABI: operand-type dispatch of the reference's BLAS wrappers, status ->
// exception mapping, the per-thread context, one interval product per form
// and a batched product checked bit for bit against the CPU oracle, and the
// device-side solvers (verified solve with R formed by rocSOLVER, dd solve)
// in a process that loads ROCm's own rocSOLVER / rocBLAS.
//   ./test_policy_core          logic only (no GPU)
//   ./test_policy_core gpu      also the device part
#include <algorithm>
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
  orc_ozaki_igemm(m, n, k, A.data(), m, B.data(), k, R.data(), m, 16, k >= 4096 ? ORC_OZ_RADIUS_FOLDED : ORC_OZ_RADIUS_FULL);
  EXPECT(C == R);
  bool thrown = false;
  try {
    check<Errors>(vcphip_igemm(ctx, VCPHIP_FORM_PLAIN, m, n, k, A.data(), m - 1, B.data(), k, C.data(), m), "bad lda");
  } catch (const dimension_error&) {
    thrown = true;
  }
  EXPECT(thrown);

  // batched interval products: each C_p is the single plain product
  {
    const int m2 = 70, k2 = 13, n2 = 40;
    std::vector<double> A2(2 * m2 * k2), B2(2 * k2 * n2), C2(2 * m2 * n2), C1(2 * m * n), R2(2 * m2 * n2);
    fill(A2);
    fill(B2);
    const int ms[2] = {m, m2}, ns[2] = {n, n2}, ks[2] = {k, k2}, las[2] = {m, m2}, lbs[2] = {k, k2};
    const double* As[2] = {A.data(), A2.data()};
    const double* Bs[2] = {B.data(), B2.data()};
    double* Cs[2] = {C1.data(), C2.data()};
    check<Errors>(vcphip_igemm_batched(ctx, 2, ms, ns, ks, As, las, Bs, lbs, Cs, las), "batched");
    orc_pidblas_mulmm(m, n, k, A.data(), m, B.data(), k, R.data(), m);
    orc_pidblas_mulmm(m2, n2, k2, A2.data(), m2, B2.data(), k2, R2.data(), m2);
    EXPECT(C1 == R && C2 == R2);
  }

  // the verified solve with R formed on the device (rocSOLVER in this
  // process's own ROCm libraries), against the oracle given that R:
  // tridiag(-1, 2, -1) of order 15 with b = 1..15, whose solution is dyadic
  {
    const int N = 15;
    std::vector<double> T(2 * N * N, 0.0), b(2 * N), X(2 * N), Xo(2 * N), Rd(N * N);
    for (int i = 0; i < N; ++i) {
      for (int j = 0; j < N; ++j) {
        const double v = (i == j) ? 2.0 : ((i - j == 1 || j - i == 1) ? -1.0 : 0.0);
        T[2 * (i + N * j)] = T[2 * (i + N * j) + 1] = v;
      }
      b[2 * i] = b[2 * i + 1] = i + 1.0;
    }
    double g = 1.0, go = 1.0;
    check<Errors>(vcphip_lss(ctx, N, 1, T.data(), N, b.data(), N, nullptr, N, X.data(), N, &g), "lss");
    check<Errors>(vcphip_inv_mid(ctx, N, T.data(), N, Rd.data(), N), "inv_mid");
    EXPECT(orc_lss(N, 1, T.data(), N, b.data(), N, Rd.data(), N, Xo.data(), N, &go) == 0);
    EXPECT(g == go && X == Xo && g < 1e-12);
    for (int i = 0; i < N; ++i) {   // x_i = sum_j min(i, j) (16 - max(i, j)) / 16 * j, 1-based
      double x = 0.0;
      for (int j = 1; j <= N; ++j) x += (double)(std::min(i + 1, j) * (N + 1 - std::max(i + 1, j))) / (N + 1) * j;
      EXPECT(X[2 * i] <= x && x <= X[2 * i + 1]);
    }
  }

  // the dd solve: the residual of its solution, formed by the oracle
  {
    const int N = 9;
    std::vector<double> Ad(2 * N * N), bd(2 * N), xd(2 * N), rd(2 * N);
    for (int i = 0; i < N * N; ++i) {
      Ad[2 * i] = u(rng) + ((i % (N + 1)) == 0 ? 4.0 : 0.0);
      Ad[2 * i + 1] = 0.0;
    }
    for (int i = 0; i < N; ++i) {
      bd[2 * i] = u(rng);
      bd[2 * i + 1] = 0.0;
    }
    int rounds = 0;
    check<Errors>(vcphip_dd_solve(ctx, N, 1, Ad.data(), N, bd.data(), N, xd.data(), N, 10, &rounds), "dd_solve");
    EXPECT(rounds >= 2 && rounds <= 6);
    orc_dd_residual(N, 1, N, Ad.data(), N, xd.data(), N, bd.data(), N, rd.data(), N, 31);
    for (int i = 0; i < N; ++i) EXPECT(std::fabs(rd[2 * i]) <= 1e-28);
  }
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
