// libvcphip host side: the C ABI declared in include/vcphip.h.
//
// A context owns one HIP stream (default priority) and a cached device
// workspace; host-pointer entry points stage operands through it, device
// entry points enqueue on the stream and return.  Parameters of the Ozaki
// form (moduli, scale exponent, CRT constants) are derived here on the host.
#include <fenv.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/vcphip.h"
#include "vcphip_kernels.h"

#define VCPHIP_TRY(x)          \
  do {                         \
    const int _rc = (x);       \
    if (_rc) return _rc;       \
  } while (0)

struct vcphip_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  long long oz_nmod = 0, oz_variant = 0, timing = 0, dd_nmod = 0, oz_radius = 0, batch_tiles = 0, plane_rows = 0;
  // workspace slots (device memory, grown on demand)
  enum { WS_PA, WS_PB, WS_U, WS_Z, WS_SA, WS_SB, WS_HA, WS_HB, WS_HC, WS_HC2, WS_BATCH, WS_DD, WS_LSS, WS_HR, WS_LSSB,
         WS_INV, WS_DDS, WS_IGB, WS_TT, WS_MR, WS_Q, WS_N };
  void* ws[WS_N] = {};
  size_t ws_size[WS_N] = {};
  // stage timing: events recorded by every call since the last synchronize,
  // folded into per-stage totals there
  std::vector<hipEvent_t> ev;
  std::vector<std::string> ev_name;
  std::vector<std::string> st_name;
  std::vector<double> st_ms;
  std::vector<long long> st_cnt;
  // batched residual: descriptors of the last problem set and its graph
  struct Resid {
    std::vector<VcphipResidProblem> probs;
    std::vector<VcphipResidTile> tiles;
    std::vector<int> row_problem;
    double* normub = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int n64 = 0, n32 = 0;       // tiles of each shape (64 x 64 first in the tile array)
    int grid = 0, grid32 = 0;   // persistent grids of the two tile launches
    // device descriptors of the cached set (for re-timing by the tuner)
    const VcphipResidTile* dt = nullptr;
    const VcphipResidProblem* dp = nullptr;
    const int* drp = nullptr;
    double* drs = nullptr;
    int nrows = 0;
  } resid;
  int num_cu = 0;
  // autotuner table: choices per product kind and size class (ceil(log2) of
  // m, n, k); 0 in a field means "not tuned"
  struct Tune {
    int kind, lm, ln, lk, variant, nmod, radius;
  };
  std::vector<Tune> tune;
  int resid_per_cu = 3;   // persistent workgroups per CU of the batched residual (its occupancy)
  // tile shape of the batched launches per size class of a problem's rows
  // (ceil(log2)): VCPHIP_BATCH_TILES_*, 0 = not tuned (64 x 64); [0] the
  // residual, [1] the interval products
  unsigned char tshape[2][32] = {};
  // second stream (default priority) of the batched solve: the right-hand
  // side steps run on it beside the residual's persistent launch
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // rocBLAS handle of the approximate inverses (rocSOLVER getrf/getri),
  // created on first use
  rocblas_handle blas = nullptr;
  // three more streams (default priority), each with its own rocBLAS handle,
  // over which a batch of approximate inverses spreads its getrf/getrs
  // groups: rocSOLVER's panel kernels fill a few CUs each, so groups of
  // different orders run side by side; created on first use
  static constexpr int INV_LANES = 3;
  hipStream_t inv_s[INV_LANES] = {};
  rocblas_handle inv_h[INV_LANES] = {};
  hipEvent_t inv_fork = nullptr, inv_join[INV_LANES] = {};
  int inv_lanes = -1;   // extra lanes available; -1 before the first try
  // two pinned host buffers (hipHostMalloc) that stage the large copies of
  // the host-pointer entry points, and the events of their last DMA
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  // batched interval products: descriptors of the last problem set and its graph
  struct Igb {
    std::vector<VcphipIgemmProblem> probs;
    std::vector<VcphipResidTile> tiles;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int n64 = 0, n32 = 0, grid = 0, grid32 = 0;
    const VcphipIgemmProblem* dp = nullptr;   // device descriptors of the cached set
  } igb;
};

namespace {

int ws_get(vcphip_ctx* c, int slot, size_t bytes, void** out) {
  if (bytes == 0) bytes = 256;
  if (c->ws_size[slot] < bytes) {
    if (c->ws[slot]) {
      // every stream of the context may still read or write the slot: the
      // context stream, the second stream of the batched solve and the
      // inverse lanes (they join the context stream before each entry point
      // returns, but not before an early error return)
      hipStreamSynchronize(c->stream);
      if (c->aux) hipStreamSynchronize(c->aux);
      for (int l = 0; l < vcphip_ctx::INV_LANES; ++l)
        if (c->inv_s[l]) hipStreamSynchronize(c->inv_s[l]);
      hipFree(c->ws[slot]);
      c->ws[slot] = nullptr;
      c->ws_size[slot] = 0;
    }
    const size_t want = (bytes + 0xFFFFF) & ~(size_t)0xFFFFF;   // 1 MiB granularity
    if (hipMalloc(&c->ws[slot], want) != hipSuccess) {
      hipGetLastError();
      return VCPHIP_ERR_OOM;
    }
    c->ws_size[slot] = want;
  }
  *out = c->ws[slot];
  return VCPHIP_OK;
}

struct StageTimer {
  vcphip_ctx* c;
  bool on;
  // allowed = false: no events at all (the no-allocation entry point)
  explicit StageTimer(vcphip_ctx* ctx, bool allowed = true) : c(ctx), on(allowed && ctx->timing != 0) {
    if (!on) return;
    mark("");   // call boundary: the next stage is timed from here
  }
  void mark(const char* name) {
    if (!on) return;
    hipEvent_t e;
    hipEventCreate(&e);
    hipEventRecord(e, c->stream);
    c->ev.push_back(e);
    c->ev_name.push_back(name);
  }
};

int check_launch(int rc) { return rc == 0 ? VCPHIP_OK : (rc == -7 ? VCPHIP_ERR_UNSUPPORTED : VCPHIP_ERR_DEVICE); }

// autotuner table
enum { TUNE_IGEMM = 1, TUNE_DDGEMM = 2 };
// untuned int8 GEMM tile: 256 x 256, eight waves of 128 x 64, 4 stages.  At
// n = 8192 it ties the four-wave variant 8 (10.7 ms per 22 plane products);
// at n = 4096, with one 256 x 256 tile per CU and plane, its second wave per
// SIMD hides what the four-wave tile exposes (interval 1.38 against 1.67 ms,
// dd 2.03 against 2.10 ms)
constexpr int DEFAULT_VARIANT = 15;

int size_class(int x) {
  int l = 0;
  while (l < 31 && (1 << l) < x) ++l;
  return l;
}

const vcphip_ctx::Tune* tune_find(const vcphip_ctx* c, int kind, int m, int n, int k) {
  const int lm = size_class(m), ln = size_class(n), lk = size_class(k);
  for (const auto& t : c->tune)
    if (t.kind == kind && t.lm == lm && t.ln == ln && t.lk == lk) return &t;
  return nullptr;
}

void tune_put(vcphip_ctx* c, int kind, int lm, int ln, int lk, int variant, int nmod, int radius = 0) {
  for (auto& t : c->tune)
    if (t.kind == kind && t.lm == lm && t.ln == ln && t.lk == lk) {
      if (variant) t.variant = variant;
      if (nmod) t.nmod = nmod;
      if (radius) t.radius = radius;
      return;
    }
  c->tune.push_back(vcphip_ctx::Tune{kind, lm, ln, lk, variant, nmod, radius});
}

// radius-digit scheme of an interval Ozaki product: the parameter, else the
// tuned choice, else by k: the folded scheme's extra term (A's low digits
// against a unit of B's digit) averages out over thousands of products, so
// from k = 4096 it gives the full scheme's radii (max ratio to the plain form
// 0.88 either way at n = k = 8192 on exponents spread over 41 binades) for
// two plane products less; below that the full scheme
int oz_radius_scheme(const vcphip_ctx* c, const vcphip_ctx::Tune* tuned, int k) {
  if (c->oz_radius) return (int)c->oz_radius;
  if (tuned && tuned->radius) return tuned->radius;
  return k >= VCPHIP_OZ_FOLDED_MINK ? VCPHIP_OZ_RADIUS_FOLDED : VCPHIP_OZ_RADIUS_FULL;
}
bool oz_radius_valid(long long v) { return v >= VCPHIP_OZ_RADIUS_FULL && v <= VCPHIP_OZ_RADIUS_CEIL; }

// int8 plane products of the radius term under each scheme
int oz_radius_products(int radius) {
  switch (radius) {
    case VCPHIP_OZ_RADIUS_FULL: return 8;
    case VCPHIP_OZ_RADIUS_FOLDED: return 6;
    case VCPHIP_OZ_RADIUS_HALF: return 4;
    default: return 2;
  }
}

// the radius-digit GEMM outputs of a scheme (digit planes: A 0 |m|+r hi,
// 1 lo, 2 r hi, 3 lo, 4 |m|+r rounded-up, 5 r rounded-up; B 0 r H, 1 L,
// 2 |m| H, 3 L), in the order the reconstruction reads them
void oz_set_radius(VcphipOzParams* P, int radius) {
  P->rscheme = radius;
  auto out = [&](int r, int a0, int b0, int a1 = -1, int b1 = -1) {
    P->rnp[r] = a1 < 0 ? 1 : 2;
    P->rpa[r][0] = a0;
    P->rpb[r][0] = b0;
    P->rpa[r][1] = a1 < 0 ? a0 : a1;
    P->rpb[r][1] = b1 < 0 ? b0 : b1;
  };
  switch (radius) {
    case VCPHIP_OZ_RADIUS_FULL:
    case VCPHIP_OZ_RADIUS_FOLDED:
      out(0, 0, 0);            // term 1: hi x H
      out(1, 0, 1, 1, 0);      //         hi x L + lo x H
      out(2, 2, 2);            // term 2
      out(3, 2, 3, 3, 2);
      P->rnz = 4;
      if (radius == VCPHIP_OZ_RADIUS_FULL) {
        out(4, 1, 1);          // lo x L of each term
        out(5, 3, 3);
        P->rnz = 6;
      }
      break;
    case VCPHIP_OZ_RADIUS_HALF:
      out(0, 0, 0);            // term 1: hi x H, lo x H
      out(1, 1, 0);
      out(2, 2, 2);            // term 2
      out(3, 3, 2);
      P->rnz = 4;
      break;
    default:                   // ceil: rounded-up x H of each term
      out(0, 4, 0);
      out(1, 5, 2);
      P->rnz = 2;
      break;
  }
  P->dmask_a = P->dmask_b = 0;
  for (int r = 0; r < P->rnz; ++r)
    for (int q = 0; q < P->rnp[r]; ++q) {
      P->dmask_a |= 1 << P->rpa[r][q];
      P->dmask_b |= 1 << P->rpb[r][q];
    }
}

int gcd_i(int a, int b) {
  while (b) {
    int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

int modinv(int a, int m) {   // a^{-1} mod m, gcd(a, m) = 1
  int t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
  while (nr) {
    const int q = r / nr;
    int tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  return ((t % m) + m) % m;
}

// floor(2^48 w / M) for w < M, both of VCPHIP_OZ_NLIMB limbs (binary long division)
unsigned long long frac48(const uint32_t (&w)[VCPHIP_OZ_NLIMB], const uint32_t (&M)[VCPHIP_OZ_NLIMB]) {
  uint32_t r[VCPHIP_OZ_NLIMB + 1];
  for (int l = 0; l < VCPHIP_OZ_NLIMB; ++l) r[l] = w[l];
  r[VCPHIP_OZ_NLIMB] = 0;
  unsigned long long F = 0;
  for (int b = 0; b < 48; ++b) {   // invariant r < M
    uint32_t cin = 0;
    for (int l = 0; l <= VCPHIP_OZ_NLIMB; ++l) {
      const uint32_t cout = r[l] >> 31;
      r[l] = (r[l] << 1) | cin;
      cin = cout;
    }
    bool ge = r[VCPHIP_OZ_NLIMB] != 0;
    if (!ge) {
      ge = true;   // equal counts as >=
      for (int l = VCPHIP_OZ_NLIMB - 1; l >= 0; --l)
        if (r[l] != M[l]) {
          ge = r[l] > M[l];
          break;
        }
    }
    F <<= 1;
    if (ge) {
      long long borrow = 0;
      for (int l = 0; l <= VCPHIP_OZ_NLIMB; ++l) {
        const long long d = (long long)r[l] - (l < VCPHIP_OZ_NLIMB ? (long long)M[l] : 0) - borrow;
        borrow = d < 0;
        r[l] = (uint32_t)(d + (borrow << 32));
      }
      F |= 1;
    }
  }
  return F;
}

// greedy pairwise-coprime moduli from 256 downward
int oz_moduli(int nmod, int* out) {
  int cnt = 0;
  for (int m = 256; m >= 2 && cnt < nmod; --m) {
    bool ok = true;
    for (int t = 0; t < cnt; ++t)
      if (gcd_i(m, out[t]) != 1) { ok = false; break; }
    if (ok) out[cnt++] = m;
  }
  return cnt;
}

// CRT constants shared by the interval and dd forms; beta from the no-wrap
// condition k (2^beta + slack)^2 < M/2, i.e. k 2^(2 beta + extra) < M
int oz_fill_params(int k, int nmod, int extra, int beta_cap, VcphipOzParams* P) {
  memset(P, 0, sizeof(*P));
  P->dmask_a = P->dmask_b = ~0;   // every digit plane until a radius scheme is set
  if (nmod < 8 || nmod > VCPHIP_OZ_MAXMOD) return VCPHIP_ERR_ARGUMENT;
  if (oz_moduli(nmod, P->mod) != nmod) return VCPHIP_ERR_ARGUMENT;
  P->nmod = nmod;
  {
    double lm = 0.0;
    for (int t = 0; t < nmod; ++t) lm += log2((double)P->mod[t]);
    int lk = 0;
    while ((1ll << lk) < (long long)(k > 1 ? k : 1)) ++lk;
    const int beta = (int)floor((lm - (double)extra - (double)lk - 1e-6) / 2.0);
    P->beta = beta > beta_cap ? beta_cap : beta;
  }
  if (P->beta < 8) return VCPHIP_ERR_UNSUPPORTED;
  for (int t = 0; t < nmod; ++t) {
    const int m = P->mod[t];
    if (m <= 32) return VCPHIP_ERR_UNSUPPORTED;   // the splits' 24-bit quotient needs r_magic < 2^24
    unsigned wb[16];   // 256^i mod m, i = 0..15 (each < 256)
    unsigned long long pw = 1 % (unsigned)m;
    for (int i = 0; i < 16; ++i) {
      wb[i] = (unsigned)pw;
      pw = pw * 256ull % (unsigned)m;
    }
    for (int q = 0; q < 4; ++q)
      P->dd_wb[t][q] = wb[4 * q] | (wb[4 * q + 1] << 8) | (wb[4 * q + 2] << 16) | (wb[4 * q + 3] << 24);
    P->dd_kc[t] = -128 * 128 * 16;
    for (int i = 0; i < 16; ++i) P->dd_kc[t] += 128 * (int32_t)wb[i];
    P->r_magic[t] = (uint32_t)(((1u << 29) + (unsigned)m - 1u) / (unsigned)m);
    P->ep_bias[t] = (uint32_t)((2048 + m - 1) / m * m) << 16;
    P->ep_c16[t] = (int32_t)(65536 % m) - 65536;
    for (int q = 0; q < 2; ++q) P->r_wb[q][t] = P->dd_wb[t][q];
    const unsigned two63 = (unsigned)(wb[7] * 128ull % (unsigned)m);   // 2^63 = 256^7 * 128
    P->r_bias[t] = (uint32_t)((int)(((unsigned)(m >> 1) + (unsigned)m - two63) % (unsigned)m) - (m >> 1));
  }
  uint32_t M[VCPHIP_OZ_NLIMB] = {1, 0, 0, 0, 0, 0, 0, 0};
  for (int t = 0; t < nmod; ++t) {
    unsigned long long carry = 0;
    for (int l = 0; l < VCPHIP_OZ_NLIMB; ++l) {
      const unsigned long long p = (unsigned long long)M[l] * (unsigned)P->mod[t] + carry;
      M[l] = (uint32_t)p;
      carry = p >> 32;
    }
    if (carry) return VCPHIP_ERR_UNSUPPORTED;
  }
  for (int l = 0; l < VCPHIP_OZ_NLIMB; ++l) P->Mlimb[l] = M[l];
  // CRT weights w_t = (M/m_t) y_t, y_t = (M/m_t)^-1 mod m_t, their bytes and
  // F_t = floor(2^48 w_t / M) as the centred int8 columns of the MFMA CRT
  for (int t = 0; t < nmod; ++t) {
    const unsigned m = (unsigned)P->mod[t];
    uint32_t Mt[VCPHIP_OZ_NLIMB];
    unsigned long long rem = 0;
    for (int l = VCPHIP_OZ_NLIMB - 1; l >= 0; --l) {   // M / m_t (exact)
      const unsigned long long cur = (rem << 32) | M[l];
      Mt[l] = (uint32_t)(cur / m);
      rem = cur % m;
    }
    unsigned long long mt_mod = 0;   // (M / m_t) mod m_t
    for (int l = VCPHIP_OZ_NLIMB - 1; l >= 0; --l) mt_mod = ((mt_mod << 32) | Mt[l]) % m;
    const unsigned y = (unsigned)modinv((int)mt_mod, (int)m);
    uint32_t wt[VCPHIP_OZ_NLIMB];
    unsigned long long carry = 0;
    for (int l = 0; l < VCPHIP_OZ_NLIMB; ++l) {
      const unsigned long long p = (unsigned long long)Mt[l] * y + carry;
      wt[l] = (uint32_t)p;
      carry = p >> 32;
    }
    if (carry) return VCPHIP_ERR_UNSUPPORTED;
    const unsigned long long F = frac48(wt, M);
    auto put = [&](int c, unsigned byte, bool centred) {
      const unsigned v = centred ? ((byte - 128u) & 0xFFu) : byte;
      P->crt_a[c][t / 4] |= v << (8 * (t % 4));
      if (centred) P->crt_k[c] += 128 * (int32_t)byte;
    };
    for (int b = 0; b < 6; ++b) put(b, (unsigned)(F >> (8 * b)) & 0xFFu, true);
    put(6, 1u, false);
    for (int c = 0; c < nmod; ++c) put(7 + c, (wt[c / 4] >> (8 * (c % 4))) & 0xFFu, true);   // w_t < M < 2^(8 N)
  }
  for (int c = 0; c < nmod + 7; ++c)
    if (c != 6) P->crt_k[c] -= 128 * 128 * nmod;
  {
    const int old = fegetround();
    volatile double one = 1.0, d = 16129.0, d1 = 127.0;
    fesetround(FE_DOWNWARD);
    P->c1dn = one / d1;
    P->c127dn = one / d;
    fesetround(FE_UPWARD);
    P->c127 = one / d;
    P->c1 = one / d1;
    fesetround(old);
  }
  return VCPHIP_OK;
}

// interval form: |A'| <= 2^beta, k 2^(2 beta + 1) < M, beta <= 62 (int64)
int oz_make_params(int k, int nmod, VcphipOzParams* P) {
  if (nmod > VCPHIP_OZ_MAXMOD_IV) return VCPHIP_ERR_ARGUMENT;
  return oz_fill_params(k, nmod, 1, 62, P);
}

// dd form: |A'| <= 2^beta + 1, k 2^(2 beta + 2) < M, beta <= 115 (int128)
int dd_make_params(int k, int nmod, VcphipOzParams* P) { return oz_fill_params(k, nmod, 2, 115, P); }

int default_dd_nmod(int k) {
  // smallest count giving beta >= 110 (a product accurate to ~2^-108 of the
  // row / column scales, the dd precision)
  for (int nmod = 24; nmod <= VCPHIP_OZ_MAXMOD; ++nmod) {
    VcphipOzParams P;
    if (dd_make_params(k, nmod, &P) == VCPHIP_OK && P.beta >= 110) return nmod;
  }
  return VCPHIP_OZ_MAXMOD;
}

int default_nmod(int k) {
  // 16 moduli: beta = 55 at k = 8192, a scaling error far below the FP64
  // directed-rounding width of the same product
  (void)k;
  return 16;
}

}  // namespace

// ===========================================================================
// context
// ===========================================================================
extern "C" int vcphip_version(void) { return 1; }

namespace {
void* scratch_init(int device);   // the no-allocation entry point's scratch, below
}  // namespace

extern "C" const char* vcphip_strerror(int code) {
  switch (code) {
    case VCPHIP_OK: return "ok";
    case VCPHIP_ERR_DIMENSION: return "dimension mismatch";
    case VCPHIP_ERR_ARGUMENT: return "invalid argument";
    case VCPHIP_ERR_DEVICE: return "HIP device error";
    case VCPHIP_ERR_OOM: return "device memory exhausted";
    case VCPHIP_ERR_ROUNDING: return "directed rounding not in effect";
    case VCPHIP_ERR_VERIFICATION: return "verification failed";
    case VCPHIP_ERR_UNSUPPORTED: return "unsupported size or form";
    case VCPHIP_ERR_RANGE: return "exponent range not supported";
    case VCPHIP_ERR_IO: return "matrix file I/O or format error";
    case VCPHIP_ERR_SINGULAR: return "factorization failed: zero pivot or not positive definite";
    case VCPHIP_ERR_SCRATCH: return "product needs more scratch than the no-allocation entry point holds";
    default: return "unknown error";
  }
}

extern "C" int vcphip_ctx_create(int device, vcphip_ctx** out) {
  if (!out) return VCPHIP_ERR_ARGUMENT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    hipGetLastError();
    return VCPHIP_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return VCPHIP_ERR_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return VCPHIP_ERR_DEVICE;
  vcphip_ctx* c = new vcphip_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return VCPHIP_ERR_DEVICE;
  }
  c->stream = c->own;
  hipDeviceProp_t prop;
  c->num_cu = (hipGetDeviceProperties(&prop, device) == hipSuccess) ? prop.multiProcessorCount : 256;
  // the no-allocation entry point's scratch address on this device (once per
  // device; a failure only disables that entry point)
  scratch_init(device);
  *out = c;
  return VCPHIP_OK;
}

extern "C" int vcphip_ctx_destroy(vcphip_ctx* c) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  for (int s = 0; s < vcphip_ctx::WS_N; ++s)
    if (c->ws[s]) hipFree(c->ws[s]);
  for (auto e : c->ev) hipEventDestroy(e);
  if (c->resid.exec) hipGraphExecDestroy(c->resid.exec);
  if (c->resid.graph) hipGraphDestroy(c->resid.graph);
  if (c->aux) {
    hipStreamSynchronize(c->aux);
    hipStreamDestroy(c->aux);
    hipEventDestroy(c->ev_fork);
    hipEventDestroy(c->ev_join);
  }
  if (c->blas) rocblas_destroy_handle(c->blas);
  for (int l = 0; l < vcphip_ctx::INV_LANES; ++l) {
    if (c->inv_s[l]) hipStreamSynchronize(c->inv_s[l]);
    if (c->inv_h[l]) rocblas_destroy_handle(c->inv_h[l]);
    if (c->inv_s[l]) hipStreamDestroy(c->inv_s[l]);
    if (c->inv_join[l]) hipEventDestroy(c->inv_join[l]);
  }
  if (c->inv_fork) hipEventDestroy(c->inv_fork);
  for (int h = 0; h < 2; ++h) {
    if (c->pin_ev[h]) hipEventDestroy(c->pin_ev[h]);
    if (c->pin[h]) hipHostFree(c->pin[h]);
  }
  if (c->igb.exec) hipGraphExecDestroy(c->igb.exec);
  if (c->igb.graph) hipGraphDestroy(c->igb.graph);
  hipStreamDestroy(c->own);
  delete c;
  return VCPHIP_OK;
}

extern "C" int vcphip_ctx_set_stream(vcphip_ctx* c, void* stream) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  c->stream = stream ? (hipStream_t)stream : c->own;
  return VCPHIP_OK;
}

extern "C" int vcphip_ctx_synchronize(vcphip_ctx* c) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return VCPHIP_ERR_DEVICE;
  for (size_t i = 1; i < c->ev.size(); ++i) {
    if (c->ev_name[i].empty()) continue;   // a call boundary ends no stage
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev[i - 1], c->ev[i]);
    size_t s = 0;
    while (s < c->st_name.size() && c->st_name[s] != c->ev_name[i]) ++s;
    if (s == c->st_name.size()) {
      c->st_name.push_back(c->ev_name[i]);
      c->st_ms.push_back(0.0);
      c->st_cnt.push_back(0);
    }
    c->st_ms[s] += ms;
    c->st_cnt[s] += 1;
  }
  for (auto e : c->ev) hipEventDestroy(e);
  c->ev.clear();
  c->ev_name.clear();
  return VCPHIP_OK;
}

extern "C" int vcphip_ctx_set_param(vcphip_ctx* c, int key, long long v) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  switch (key) {
    case VCPHIP_PARAM_OZ_NMOD:
      if (v != 0 && (v < 8 || v > VCPHIP_OZ_MAXMOD_IV)) return VCPHIP_ERR_ARGUMENT;
      c->oz_nmod = v;
      return VCPHIP_OK;
    case VCPHIP_PARAM_OZ_VARIANT:
      if (v != 0 && (v < 6 || v > 18)) return VCPHIP_ERR_ARGUMENT;
      c->oz_variant = v;
      return VCPHIP_OK;
    case VCPHIP_PARAM_STAGE_TIMING:   // (re)enabling clears the totals
      c->timing = v ? 1 : 0;
      c->st_name.clear();
      c->st_ms.clear();
      c->st_cnt.clear();
      return VCPHIP_OK;
    case VCPHIP_PARAM_DD_NMOD:
      if (v != 0 && (v < 24 || v > VCPHIP_OZ_MAXMOD)) return VCPHIP_ERR_ARGUMENT;
      c->dd_nmod = v;
      return VCPHIP_OK;
    case VCPHIP_PARAM_OZ_RADIUS:
      if (v != 0 && !oz_radius_valid(v)) return VCPHIP_ERR_ARGUMENT;
      c->oz_radius = v;
      return VCPHIP_OK;
    case VCPHIP_PARAM_BATCH_TILES:
      if (v != 0 && (v < VCPHIP_BATCH_TILES_64 || v > VCPHIP_BATCH_TILES_32)) return VCPHIP_ERR_ARGUMENT;
      if (v != c->batch_tiles) {   // cached tile lists are rebuilt on the next call
        c->resid.probs.clear();
        c->igb.probs.clear();
      }
      c->batch_tiles = v;
      return VCPHIP_OK;
    case VCPHIP_PARAM_PLANE_ROWS:
      if (v != 0 && v != 32 && v != 64) return VCPHIP_ERR_ARGUMENT;
      c->plane_rows = v;
      return VCPHIP_OK;
    default: return VCPHIP_ERR_ARGUMENT;
  }
}

extern "C" int vcphip_ctx_get_param(vcphip_ctx* c, int key, long long* v) {
  if (!c || !v) return VCPHIP_ERR_ARGUMENT;
  switch (key) {
    case VCPHIP_PARAM_OZ_NMOD: *v = c->oz_nmod; return VCPHIP_OK;
    case VCPHIP_PARAM_OZ_VARIANT: *v = c->oz_variant; return VCPHIP_OK;
    case VCPHIP_PARAM_STAGE_TIMING: *v = c->timing; return VCPHIP_OK;
    case VCPHIP_PARAM_DD_NMOD: *v = c->dd_nmod; return VCPHIP_OK;
    case VCPHIP_PARAM_OZ_RADIUS: *v = c->oz_radius; return VCPHIP_OK;
    case VCPHIP_PARAM_BATCH_TILES: *v = c->batch_tiles; return VCPHIP_OK;
    case VCPHIP_PARAM_PLANE_ROWS: *v = c->plane_rows; return VCPHIP_OK;
    default: return VCPHIP_ERR_ARGUMENT;
  }
}

extern "C" int vcphip_ctx_stage_times(vcphip_ctx* c, double* ms, int max, int* count, char* names, int names_len) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  const int n = (int)c->st_name.size();
  if (count) *count = n;
  std::string all;
  for (int i = 0; i < n; ++i) {
    if (ms && i < max) ms[i] = c->st_cnt[i] ? c->st_ms[i] / (double)c->st_cnt[i] : 0.0;
    if (i) all += ";";
    all += c->st_name[i];
  }
  if (names && names_len > 0) {
    strncpy(names, all.c_str(), (size_t)names_len - 1);
    names[names_len - 1] = 0;
  }
  return VCPHIP_OK;
}

extern "C" int vcphip_ozaki_config(int k, int nmod, int* beta, int* moduli, int max_moduli) {
  if (nmod <= 0) nmod = default_nmod(k);
  VcphipOzParams P;
  const int rc = oz_make_params(k, nmod, &P);
  if (rc) return rc;
  if (beta) *beta = P.beta;
  for (int t = 0; t < nmod && t < max_moduli && moduli; ++t) moduli[t] = P.mod[t];
  return nmod;
}

extern "C" int vcphip_ozaki_plan(vcphip_ctx* c, int m, int n, int k, int* nmod, int* radius, int* variant,
                                 int* products) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (m < 0 || n < 0 || k < 0) return VCPHIP_ERR_DIMENSION;
  const vcphip_ctx::Tune* tuned = tune_find(c, TUNE_IGEMM, m, n, k);
  const int nm = c->oz_nmod ? (int)c->oz_nmod : (tuned && tuned->nmod ? tuned->nmod : default_nmod(k));
  const int rad = oz_radius_scheme(c, tuned, k);
  const int var = c->oz_variant ? (int)c->oz_variant : (tuned && tuned->variant ? tuned->variant : DEFAULT_VARIANT);
  if (nmod) *nmod = nm;
  if (radius) *radius = rad;
  if (variant) *variant = var;
  // residue products plus hi x hi, two mixed and (full) lo x lo per radius term
  if (products) *products = nm + oz_radius_products(rad);
  return VCPHIP_OK;
}

// ===========================================================================
// device entry points
// ===========================================================================
static int dims_ok(int m, int n, int k) { return m >= 0 && n >= 0 && k >= 0; }

// the moduli count, radius scheme, tile variant and buffer sizes of one
// interval Ozaki product in this context
struct OzPlan {
  VcphipOzParams P;
  int nmod = 0, nz = 0, variant = 0, mpad = 0, npad = 0, kpad = 0, nchunk = 0;
  size_t b_pa = 0, b_pb = 0, b_pu = 0, b_pz = 0, b_sa = 0, b_sb = 0;   // bytes of each workspace buffer
};

static int oz_plan(vcphip_ctx* c, int m, int n, int k, OzPlan& p) {
  if (k > VCPHIP_OZ_MAXK) return VCPHIP_ERR_UNSUPPORTED;
  const vcphip_ctx::Tune* tuned = tune_find(c, TUNE_IGEMM, m, n, k);
  p.nmod = c->oz_nmod ? (int)c->oz_nmod : (tuned && tuned->nmod ? tuned->nmod : default_nmod(k));
  const int rc = oz_make_params(k, p.nmod, &p.P);
  if (rc) return rc;
  oz_set_radius(&p.P, oz_radius_scheme(c, tuned, k));
  p.nz = p.P.rnz;
  p.variant = c->oz_variant ? (int)c->oz_variant : (tuned && tuned->variant ? tuned->variant : DEFAULT_VARIANT);
  p.mpad = (m + VCPHIP_OZ_PADMN - 1) / VCPHIP_OZ_PADMN * VCPHIP_OZ_PADMN;
  p.npad = (n + VCPHIP_OZ_PADMN - 1) / VCPHIP_OZ_PADMN * VCPHIP_OZ_PADMN;
  p.kpad = k > 0 ? (k + VCPHIP_OZ_PADK - 1) / VCPHIP_OZ_PADK * VCPHIP_OZ_PADK : VCPHIP_OZ_PADK;
  p.nchunk = k > 0 ? (k + VCPHIP_OZ_CHUNK - 1) / VCPHIP_OZ_CHUNK : 0;
  p.b_pa = ((size_t)p.nmod + VCPHIP_OZ_NDIGIT_A) * p.mpad * p.kpad;
  p.b_pb = ((size_t)p.nmod + VCPHIP_OZ_NDIGIT_B) * p.npad * p.kpad;
  p.b_pu = (size_t)p.nmod * p.mpad * p.npad;
  p.b_pz = (size_t)p.nz * sizeof(int32_t) * (size_t)p.mpad * p.npad;
  auto side_bytes = [&](int rows) {
    return (size_t)rows * (4 * sizeof(int) + sizeof(double)) +
           sizeof(double) * 5 * (size_t)(p.nchunk > 0 ? p.nchunk : 1) * rows + 1024;
  };
  p.b_sa = side_bytes(m);
  p.b_sb = side_bytes(n);
  return VCPHIP_OK;
}

// statistics, split, int8 GEMM and reconstruction of one product into the
// given buffers (sizes from oz_plan; every kernel writes the whole padded
// extent it or a later kernel reads, so the buffers need no clearing)
static int oz_run(vcphip_ctx* c, const OzPlan& p, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                  double* C, int ldc, void* pa, void* pb, void* pu, void* pz, void* sa, void* sb, bool timed) {
  auto carve = [&](void* base, int rows) {
    VcphipOzSide s;
    char* q = (char*)base;
    s.n1 = (double*)q; q += sizeof(double) * rows;
    s.part = (double*)q; q += sizeof(double) * 5 * (size_t)(p.nchunk > 0 ? p.nchunk : 1) * rows;
    s.e = (int*)q; q += sizeof(int) * rows;
    s.g = (int*)q; q += sizeof(int) * rows;
    s.h = (int*)q; q += sizeof(int) * rows;
    s.flags = (int*)q;
    return s;
  };
  const VcphipOzSide SA = carve(sa, m), SB = carve(sb, n);
  const int mpad = p.mpad, npad = p.npad, kpad = p.kpad;
  int rc;
  StageTimer T(c, timed);
  if ((rc = check_launch(vcphip_oz_stats(A, 1, lda, m, k, SA, c->stream)))) return rc;
  if ((rc = check_launch(vcphip_oz_stats(B, ldb, 1, n, k, SB, c->stream)))) return rc;
  T.mark("stats");
  if ((rc = check_launch(vcphip_oz_split_rows(A, lda, m, k, mpad, kpad, p.P, SA, (int8_t*)pa, c->stream)))) return rc;
  if ((rc = check_launch(vcphip_oz_split_cols(B, ldb, k, n, npad, kpad, p.P, SB, (int8_t*)pb, c->stream)))) return rc;
  T.mark("split");
  if ((rc = check_launch(vcphip_oz_gemm(mpad, npad, kpad, p.nmod, p.nz, (const int8_t*)pa, (const int8_t*)pb, p.P,
                                        (uint8_t*)pu, (int32_t*)pz, p.variant, c->stream))))
    return rc;
  T.mark("gemm");
  if ((rc = check_launch(vcphip_oz_recon(m, n, k, mpad, npad, p.P, SA, SB, (const uint8_t*)pu,
                                         (const int32_t*)pz, C, ldc, c->stream))))
    return rc;
  T.mark("recon");
  return VCPHIP_OK;
}

static int oz_igemm_dev(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                        double* C, int ldc) {
  OzPlan p;
  int rc = oz_plan(c, m, n, k, p);
  if (rc) return rc;
  void *pa, *pb, *pu, *pz, *sa, *sb;
  if ((rc = ws_get(c, vcphip_ctx::WS_PA, p.b_pa, &pa))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_PB, p.b_pb, &pb))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_U, p.b_pu, &pu))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_Z, p.b_pz, &pz))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_SA, p.b_sa, &sa))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_SB, p.b_sb, &sb))) return rc;
  return oz_run(c, p, m, n, k, A, lda, B, ldb, C, ldc, pa, pb, pu, pz, sa, sb, true);
}

// ---------------------------------------------------------------------------
// the Ozaki product without allocation: its workspace is one file-scope
// device array of VCPHIP_NOALLOC_SCRATCH_BYTES, held by every GPU the library
// runs on until the process ends.  Host code never uses the array's name as
// an address (it is a host-side stand-in): vcphip_ctx_create takes each
// device's own address with hipGetSymbolAddress, once per device, and the
// entry point passes that address on.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[VCPHIP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
bool g_scratch_tried[SCRATCH_MAX_DEVICES] = {};

// the scratch array's address on `device`, taken with that device current;
// nullptr if the runtime cannot give it
void* scratch_init(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_tried[device]) {
    g_scratch_tried[device] = true;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess && ptr) {
      g_scratch_ptr[device] = ptr;
    } else {
      hipGetLastError();
    }
  }
  return g_scratch_ptr[device];
}

void* scratch_of(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

extern "C" int vcphip_igemm_ozaki_noalloc_dev(vcphip_ctx* c, int m, int n, int k, const double* IA, int lda,
                                              const double* IB, int ldb, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  if (!IA || !IB || !C) return VCPHIP_ERR_ARGUMENT;
  OzPlan p;
  int rc = oz_plan(c, m, n, k, p);
  if (rc) return rc;
  const size_t o_pb = al256(p.b_pa), o_pu = o_pb + al256(p.b_pb), o_pz = o_pu + al256(p.b_pu);
  const size_t o_sa = o_pz + al256(p.b_pz), o_sb = o_sa + al256(p.b_sa), total = o_sb + al256(p.b_sb);
  if (total > VCPHIP_NOALLOC_SCRATCH_BYTES) return VCPHIP_ERR_SCRATCH;
  unsigned char* base = (unsigned char*)scratch_of(c->device);
  if (!base) return VCPHIP_ERR_DEVICE;
  hipSetDevice(c->device);
  return oz_run(c, p, m, n, k, IA, lda, IB, ldb, C, ldc, base, base + o_pb, base + o_pu, base + o_pz, base + o_sa,
                base + o_sb, false);
}

namespace {

// Tile rows of a plane-fed product.  The 32-row tiles give a CU more
// resident waves (3 against 2 per SIMD for interval x interval, 4-6 against
// 3-5 for the other forms) and a finer last round of tiles, at 1.5x the LDS
// reads per FMA; measured on MI355X they are the faster choice up to about 4
// (interval x interval) or 3 (the other forms) 64-row tiles per CU: n = 2048
// interval x interval 1.43 against 1.55 ms, n = 1024 0.21 against 0.25 ms,
// while from n = 3072 the 64-row tiles are 3-8% faster.
int plane_rows(const vcphip_ctx* c, int kind, size_t mpad, size_t npad) {
  if (c->plane_rows) return (int)c->plane_rows;
  const size_t cus = c->num_cu > 0 ? (size_t)c->num_cu : 256;
  const size_t per_cu = kind == VCPHIP_DR_II ? 4 : 3;
  return (mpad / 64) * (npad / 64) <= per_cu * cus ? 32 : 64;
}

// A single product of kind VCPHIP_DR_*: the operand planes (midpoint and
// radius of an interval operand, the point operand itself), padded and laid
// out for whole k-slices, then the product from them.  k = 0 (an all-zero or,
// accumulating, unchanged product) takes the tile kernel.
int planes_product(vcphip_ctx* c, int kind, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                   double* C, int ldc, double* C2 = nullptr, int accumulate = 0) {
  StageTimer T(c);
  if (k == 0) {
    const int rc =
        check_launch(vcphip_launch_drgemm(kind, m, n, k, A, lda, B, ldb, C, ldc, C2, accumulate, c->stream));
    T.mark("drgemm");
    return rc;
  }
  const bool ai = kind == VCPHIP_DR_II || kind == VCPHIP_DR_IP, bi = kind == VCPHIP_DR_II || kind == VCPHIP_DR_PI;
  const size_t mpad = ((size_t)m + 63) / 64 * 64, npad = ((size_t)n + 63) / 64 * 64, kpad = ((size_t)k + 15) / 16 * 16;
  const size_t na = (ai ? 2 : 1) * mpad * kpad, nb = (bi ? 2 : 1) * kpad * npad;
  void* base;
  int rc = ws_get(c, vcphip_ctx::WS_MR, sizeof(double) * (na + nb), &base);
  if (rc) return rc;
  double* A0 = (double*)base;
  double* A1 = A0 + mpad * kpad;   // rad(A) when A is an interval
  double* B0 = A0 + na;
  double* B1 = B0 + kpad * npad;   // rad(B) when B is an interval
  rc = check_launch(vcphip_launch_planes(kind, m, n, k, A, lda, B, ldb, A0, ai ? A1 : nullptr, B0,
                                         bi ? B1 : nullptr, c->stream));
  T.mark("planes");
  if (rc) return rc;
  rc = check_launch(vcphip_launch_drgemm_planes(kind, m, n, k, A0, ai ? A1 : nullptr, B0, bi ? B1 : nullptr, C, ldc,
                                                C2, accumulate, plane_rows(c, kind, mpad, npad), c->stream));
  T.mark("drgemm");
  return rc;
}

// A^T A (mulltmm / vmulmm) from the planes of A; k = 0 takes the tile kernel
int syrk_planes(vcphip_ctx* c, int interval, int k, int n, const double* A, int lda, double* C, int ldc) {
  StageTimer T(c);
  if (k == 0) {
    const int rc = check_launch(vcphip_launch_drsyrk(interval, k, n, A, lda, C, ldc, c->stream));
    T.mark("drsyrk");
    return rc;
  }
  const size_t npad = ((size_t)n + 63) / 64 * 64, kpad = ((size_t)k + 15) / 16 * 16;
  void* base;
  int rc = ws_get(c, vcphip_ctx::WS_MR, sizeof(double) * (interval ? 2 : 1) * kpad * npad, &base);
  if (rc) return rc;
  double* P0 = (double*)base;
  rc = check_launch(vcphip_launch_drsyrk_planes(interval, k, n, A, lda, P0, interval ? P0 + kpad * npad : nullptr, C,
                                                ldc, c->stream));
  T.mark("drsyrk");
  return rc;
}

}  // namespace

extern "C" int vcphip_igemm_dev(vcphip_ctx* c, int form, int m, int n, int k, const double* IA, int lda,
                                const double* IB, int ldb, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  if (form == VCPHIP_FORM_OZAKI) return oz_igemm_dev(c, m, n, k, IA, lda, IB, ldb, C, ldc);
  if (form == VCPHIP_FORM_INFSUP) {
    StageTimer T(c);
    const int rc = check_launch(vcphip_launch_isgemm(m, n, k, IA, lda, IB, ldb, C, ldc, c->stream));
    T.mark("isgemm");
    return rc;
  }
  if (form != VCPHIP_FORM_PLAIN) return VCPHIP_ERR_ARGUMENT;
  return planes_product(c, VCPHIP_DR_II, m, n, k, IA, lda, IB, ldb, C, ldc);
}

extern "C" int vcphip_igemm_ip_dev(vcphip_ctx* c, int m, int n, int k, const double* IA, int lda, const double* B,
                                   int ldb, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  return planes_product(c, VCPHIP_DR_IP, m, n, k, IA, lda, B, ldb, C, ldc);
}

extern "C" int vcphip_igemm_pi_dev(vcphip_ctx* c, int m, int n, int k, const double* B, int ldb, const double* IA,
                                   int lda, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || ldb < (m > 1 ? m : 1) || lda < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  return planes_product(c, VCPHIP_DR_PI, m, n, k, B, ldb, IA, lda, C, ldc);
}

extern "C" int vcphip_igemm_tn_dev(vcphip_ctx* c, int k, int n, const double* IA, int lda, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (k < 0 || n < 0 || lda < (k > 1 ? k : 1) || ldc < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  return syrk_planes(c, 1, k, n, IA, lda, C, ldc);
}

extern "C" int vcphip_vgemm_tn_dev(vcphip_ctx* c, int k, int n, const double* A, int lda, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (k < 0 || n < 0 || lda < (k > 1 ? k : 1) || ldc < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  return syrk_planes(c, 0, k, n, A, lda, C, ldc);
}

extern "C" int vcphip_gemm_rurd_dev(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* B,
                                    int ldb, double* Cup, double* Cdn, int ldc, int accumulate) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  return planes_product(c, VCPHIP_DR_PP, m, n, k, A, lda, B, ldb, Cup, ldc, Cdn, accumulate ? 1 : 0);
}

// dd product C = op(A) * B through the Ozaki-II pipeline.  tn: C = A^T A,
// where both operands are the columns of A (k x n).
// ta: C = A^T B with A k x m (A's columns are the row operand, split like
// B's columns)
static int dd_gemm_dev(vcphip_ctx* c, bool tn, int m, int n, int k, const double* A, int lda, const double* B,
                       int ldb, double* C, int ldc, bool ta = false) {
  if (k > VCPHIP_OZ_MAXK) return VCPHIP_ERR_UNSUPPORTED;
  const int nmod = c->dd_nmod ? (int)c->dd_nmod : default_dd_nmod(k);
  VcphipOzParams P;
  int rc = dd_make_params(k, nmod, &P);
  if (rc) return rc;
  const int mpad = (m + VCPHIP_OZ_PADMN - 1) / VCPHIP_OZ_PADMN * VCPHIP_OZ_PADMN;
  const int npad = (n + VCPHIP_OZ_PADMN - 1) / VCPHIP_OZ_PADMN * VCPHIP_OZ_PADMN;
  const int kpad = k > 0 ? (k + VCPHIP_OZ_PADK - 1) / VCPHIP_OZ_PADK * VCPHIP_OZ_PADK : VCPHIP_OZ_PADK;
  const int nchunk = k > 0 ? (k + VCPHIP_DD_CHUNK - 1) / VCPHIP_DD_CHUNK : 1;
  void *pa = nullptr, *pb, *pu, *st;
  if (!tn && (rc = ws_get(c, vcphip_ctx::WS_PA, (size_t)nmod * mpad * kpad, &pa))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_PB, (size_t)nmod * npad * kpad, &pb))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_U, (size_t)nmod * mpad * npad, &pu))) return rc;
  const size_t rows_all = (size_t)m + n;
  if ((rc = ws_get(c, vcphip_ctx::WS_DD, sizeof(int) * (2 * rows_all + (size_t)nchunk * rows_all) + 1024, &st))) return rc;
  int* ea = (int*)st;
  int* fa = ea + m;
  int* eb = fa + m;
  int* fb = eb + n;
  int* part = fb + n;
  StageTimer T(c);
  if (!tn) {
    if ((rc = check_launch(ta ? vcphip_dd_stats(A, lda, 1, m, k, ea, fa, part, c->stream)
                              : vcphip_dd_stats(A, 1, lda, m, k, ea, fa, part, c->stream))))
      return rc;
  }
  if ((rc = check_launch(vcphip_dd_stats(B, ldb, 1, n, k, eb, fb, part, c->stream)))) return rc;
  T.mark("stats");
  if (!tn && (rc = check_launch(ta ? vcphip_dd_split_cols(A, lda, k, m, mpad, kpad, P, ea, fa, (int8_t*)pa, c->stream)
                                   : vcphip_dd_split_rows(A, lda, m, k, mpad, kpad, P, ea, fa, (int8_t*)pa,
                                                          c->stream))))
    return rc;
  if ((rc = check_launch(vcphip_dd_split_cols(B, ldb, k, n, npad, kpad, P, eb, fb, (int8_t*)pb, c->stream)))) return rc;
  T.mark("split");
  const vcphip_ctx::Tune* tuned = tune_find(c, TUNE_DDGEMM, m, n, k);
  const int variant = c->oz_variant ? (int)c->oz_variant : (tuned && tuned->variant ? tuned->variant : DEFAULT_VARIANT);
  const int8_t* Aplanes = tn ? (const int8_t*)pb : (const int8_t*)pa;
  if ((rc = check_launch(vcphip_oz_gemm(mpad, npad, kpad, nmod, 0, Aplanes, (const int8_t*)pb, P, (uint8_t*)pu,
                                        nullptr, variant, c->stream))))
    return rc;
  T.mark("gemm");
  if ((rc = check_launch(vcphip_dd_recon(m, n, mpad, npad, P, tn ? eb : ea, tn ? fb : fa, eb, fb, (const uint8_t*)pu,
                                         C, ldc, c->stream))))
    return rc;
  T.mark("recon");
  return VCPHIP_OK;
}

extern "C" int vcphip_ddgemm_dev(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* B,
                                 int ldb, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  return dd_gemm_dev(c, false, m, n, k, A, lda, B, ldb, C, ldc);
}

// C = A^T B for dd matrices (the X^T Y products of the Ogita-Aishima step)
extern "C" int vcphip_ddgemm_ta_dev(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* B,
                                    int ldb, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (k > 1 ? k : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  return dd_gemm_dev(c, false, m, n, k, A, lda, B, ldb, C, ldc, true);
}

// R = B - A*X (dd): the residual of pddblas::linearsolve's refinement
// (vcp/pddblas.hpp:459-466) with the Ozaki-II dd product
extern "C" int vcphip_dd_residual_dev(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* X,
                                      int ldx, const double* B, int ldb, double* R, int ldr) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldx < (k > 1 ? k : 1) || ldb < (m > 1 ? m : 1) ||
      ldr < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  const int rc = dd_gemm_dev(c, false, m, n, k, A, lda, X, ldx, R, ldr);  // R = A X
  if (rc) return rc;
  return check_launch(vcphip_dd_sub_from(m, n, B, ldb, R, ldr, c->stream));
}

// ===========================================================================
// batched interval products C_p = IA_p * IB_p (pidblas::mulmm per problem, the
// plain form): one persistent launch over a size-sorted tile list, captured
// into a hipGraph and replayed while the problem set is unchanged
// ===========================================================================
// ---- tile lists of the batched persistent launches ---------------------------
namespace {

enum { TS_RESID = 0, TS_IGEMM = 1 };

// shape mode of a problem with `rows` output rows: the parameter, else the
// tuned table, else 64 x 64 tiles
int batch_mode(const vcphip_ctx* c, int kind, int rows) {
  if (c->batch_tiles) return (int)c->batch_tiles;
  const int m = c->tshape[kind][size_class(rows)];
  return m ? m : VCPHIP_BATCH_TILES_64;
}

struct TileSet {
  std::vector<VcphipResidTile> t64, t32;
  double tile_work = 0.0, useful = 0.0;   // multiply-adds of the tiles (padded) and of the problems
};

// the tiles of one m x n output (k deep) under a shape mode, in groups of
// eight row blocks with the columns inside a group, so that a run of
// consecutive tiles covers a compact block of the output
void add_tiles(TileSet& s, int pi, int m, int n, int k, int mode) {
  if (m <= 0 || n <= 0) return;
  std::vector<int> r0s, hs;
  for (int r0 = 0; r0 < m;) {
    const int h = mode == VCPHIP_BATCH_TILES_32 ? 32 : mode == VCPHIP_BATCH_TILES_EDGE && m - r0 <= 32 ? 32 : 64;
    r0s.push_back(r0);
    hs.push_back(h);
    r0 += h;
  }
  const int tn = (n + VCPHIP_RESID_TILE - 1) / VCPHIP_RESID_TILE;
  const size_t GH = 8;
  for (size_t g = 0; g < r0s.size(); g += GH) {
    const size_t ge = std::min(r0s.size(), g + GH);
    for (int b = 0; b < tn; ++b)
      for (size_t a = g; a < ge; ++a) {
        (hs[a] == 64 ? s.t64 : s.t32).push_back(VcphipResidTile{pi, r0s[a], b * VCPHIP_RESID_TILE, 0});
        s.tile_work += (double)hs[a] * VCPHIP_RESID_TILE * k;
      }
  }
  s.useful += (double)m * n * k;
}

// the two work-queue counters of the batched launches on the context stream
// (a fixed 256-byte slot: it is never regrown, so captured graphs keep it)
unsigned* batch_queue(vcphip_ctx* c) {
  void* q = nullptr;
  return ws_get(c, vcphip_ctx::WS_Q, 256, &q) == VCPHIP_OK ? (unsigned*)q : nullptr;
}

// grid of one shape's persistent launch
int batch_grid(const vcphip_ctx* c, size_t ntiles, int per_cu) {
  return (int)std::max<size_t>(1, std::min(ntiles, (size_t)per_cu * (c->num_cu > 0 ? c->num_cu : 256)));
}

// the launch's tile array: 64 x 64 tiles, then 32 x 64 ones.  (Reordering
// each part so that every XCD's workgroups take contiguous runs of tiles was
// measured on 256 problems of n in [65, 1021]: no change in time.)
std::vector<VcphipResidTile> tile_array(const TileSet& s) {
  std::vector<VcphipResidTile> all(s.t64);
  all.insert(all.end(), s.t32.begin(), s.t32.end());
  return all;
}

TileSet igemm_tiles(const vcphip_ctx* c, const std::vector<VcphipIgemmProblem>& probs) {
  // problems with the most work first (longest-processing-time order)
  std::vector<int> order(probs.size());
  for (size_t i = 0; i < probs.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (double)probs[a].m * probs[a].n * probs[a].k > (double)probs[b].m * probs[b].n * probs[b].k;
  });
  TileSet s;
  for (int pi : order)
    add_tiles(s, pi, probs[pi].m, probs[pi].n, probs[pi].k, batch_mode(c, TS_IGEMM, probs[pi].m));
  return s;
}

TileSet resid_tiles(const vcphip_ctx* c, const std::vector<VcphipResidProblem>& probs) {
  std::vector<int> order(probs.size());
  for (size_t i = 0; i < probs.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return probs[a].n > probs[b].n; });
  TileSet s;
  for (int pi : order) add_tiles(s, pi, probs[pi].n, probs[pi].n, probs[pi].n, batch_mode(c, TS_RESID, probs[pi].n));
  return s;
}

// Shape choice per size class of a batched set, by measurement: the classes
// in order of their work, each tried with the three shapes while the others
// keep their current choice, the set timed whole (its tile launches only), so
// that a shape is judged by what it does to the set's time.  time_set(ms)
// builds and times the set under the current table.
template <class F>
int tune_shapes(vcphip_ctx* c, int kind, const std::vector<int>& rows, const std::vector<double>& work, F time_set,
                double* best_ms) {
  double cw[32] = {};
  for (size_t i = 0; i < rows.size(); ++i) cw[size_class(rows[i])] += work[i];
  std::vector<int> classes;
  for (int l = 0; l < 32; ++l)
    if (cw[l] > 0.0) classes.push_back(l);
  std::stable_sort(classes.begin(), classes.end(), [&](int a, int b) { return cw[a] > cw[b]; });
  const long long forced = c->batch_tiles;
  c->batch_tiles = 0;
  for (int l : classes)
    if (!c->tshape[kind][l]) c->tshape[kind][l] = VCPHIP_BATCH_TILES_64;
  double best = 0.0;
  int rc = time_set(&best);
  for (size_t ci = 0; ci < classes.size() && rc == VCPHIP_OK; ++ci) {
    const int l = classes[ci];
    const int keep = c->tshape[kind][l];
    int choice = keep;
    for (int mode = VCPHIP_BATCH_TILES_64; mode <= VCPHIP_BATCH_TILES_32 && rc == VCPHIP_OK; ++mode) {
      if (mode == keep) continue;
      c->tshape[kind][l] = (unsigned char)mode;
      double ms = 0.0;
      rc = time_set(&ms);
      // a shape replaces the current one only when it is faster by more
      // than 1%: equal shapes should not trade places on noise
      if (rc == VCPHIP_OK && ms < 0.99 * best) {
        best = ms;
        choice = mode;
      }
    }
    c->tshape[kind][l] = (unsigned char)choice;
  }
  c->batch_tiles = forced;
  if (best_ms) *best_ms = best;
  return rc;
}

// one timing of `reps` back-to-back runs of enqueue() on the context stream,
// per run
template <class F>
int time_runs_once(vcphip_ctx* c, F enqueue, int reps, double* ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return VCPHIP_ERR_DEVICE;
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return VCPHIP_ERR_DEVICE;
  }
  int rc = enqueue() ? VCPHIP_ERR_DEVICE : VCPHIP_OK;   // warm-up
  if (rc == VCPHIP_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = VCPHIP_ERR_DEVICE;
  for (int r = 0; r < reps && rc == VCPHIP_OK; ++r)
    if (enqueue()) rc = VCPHIP_ERR_DEVICE;
  if (rc == VCPHIP_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
    rc = VCPHIP_ERR_DEVICE;
  float t = 0.f;
  if (rc == VCPHIP_OK) hipEventElapsedTime(&t, e0, e1);
  if (rc) hipStreamSynchronize(c->stream);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms = (double)t / reps;
  return rc;
}

// time of `reps` back-to-back runs of enqueue() on the context stream, per
// run: the least of three such measurements (run-to-run noise is a few per
// cent, the differences the tuners look for not much larger)
template <class F>
int time_runs(vcphip_ctx* c, F enqueue, int reps, double* ms) {
  double best = 0.0;
  for (int round = 0; round < 3; ++round) {
    double t = 0.0;
    const int rc = time_runs_once(c, enqueue, reps, &t);
    if (rc) return rc;
    if (round == 0 || t < best) best = t;
  }
  *ms = best;
  return VCPHIP_OK;
}

}  // namespace

// a captured graph is kept only when it holds exactly the `expect` kernel
// nodes that the capture enqueued: a capture that ends successfully but
// recorded nothing (as a tool intercepting launches can cause) would replay
// as a no-op
static bool graph_has_kernels(hipGraph_t g, size_t expect) {
  size_t nn = 0;
  if (hipGraphGetNodes(g, nullptr, &nn) != hipSuccess || nn != expect) return false;
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn && hipGraphGetNodes(g, nodes.data(), &nn) != hipSuccess) return false;
  for (hipGraphNode_t nd : nodes) {
    hipGraphNodeType t;
    if (hipGraphNodeGetType(nd, &t) != hipSuccess || t != hipGraphNodeTypeKernel) return false;
  }
  return true;
}

static bool same_igemm(const VcphipIgemmProblem& a, const VcphipIgemmProblem& b) {
  return a.A == b.A && a.B == b.B && a.C == b.C && a.lda == b.lda && a.ldb == b.ldb && a.ldc == b.ldc &&
         a.m == b.m && a.n == b.n && a.k == b.k;
}

extern "C" int vcphip_igemm_batched_dev(vcphip_ctx* c, int count, const int* m, const int* n, const int* k,
                                        const double* const* IA, const int* lda, const double* const* IB,
                                        const int* ldb, double* const* C, const int* ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count < 0 || (count > 0 && (!m || !n || !k || !IA || !lda || !IB || !ldb || !C || !ldc)))
    return VCPHIP_ERR_ARGUMENT;
  if (count == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  std::vector<VcphipIgemmProblem> probs((size_t)count);
  for (int i = 0; i < count; ++i) {
    if (!dims_ok(m[i], n[i], k[i]) || lda[i] < (m[i] > 1 ? m[i] : 1) || ldb[i] < (k[i] > 1 ? k[i] : 1) ||
        ldc[i] < (m[i] > 1 ? m[i] : 1))
      return VCPHIP_ERR_DIMENSION;
    if (m[i] > 0 && n[i] > 0 && (!C[i] || (k[i] > 0 && (!IA[i] || !IB[i])))) return VCPHIP_ERR_ARGUMENT;
    probs[i] = VcphipIgemmProblem{IA[i], IB[i], C[i], lda[i], ldb[i], ldc[i], m[i], n[i], k[i], 0};
  }
  vcphip_ctx::Igb& g = c->igb;
  bool same = g.exec && g.stream == c->stream && g.probs.size() == probs.size();
  for (size_t i = 0; same && i < probs.size(); ++i) same = same_igemm(g.probs[i], probs[i]);
  if (!same) {
    // tile lists per shape, the most work first (longest-processing-time order)
    TileSet ts = igemm_tiles(c, probs);
    g.n64 = (int)ts.t64.size();
    g.n32 = (int)ts.t32.size();
    g.grid = batch_grid(c, ts.t64.size(), 2);
    g.grid32 = batch_grid(c, ts.t32.size(), 2);
    g.tiles = tile_array(ts);
    g.probs = probs;
    const size_t bp = sizeof(VcphipIgemmProblem) * count, bt = sizeof(VcphipResidTile) * g.tiles.size();
    const size_t o_t = (bp + 255) & ~(size_t)255;
    void* base;
    const int wrc = ws_get(c, vcphip_ctx::WS_IGB, o_t + bt + 256, &base);
    if (wrc) {
      g.probs.clear();
      return wrc;
    }
    char* b8 = (char*)base;
    if (hipMemcpyAsync(b8, g.probs.data(), bp, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        (bt && hipMemcpyAsync(b8 + o_t, g.tiles.data(), bt, hipMemcpyHostToDevice, c->stream) != hipSuccess)) {
      g.probs.clear();
      return VCPHIP_ERR_DEVICE;
    }
    if (g.exec) { hipGraphExecDestroy(g.exec); g.exec = nullptr; }
    if (g.graph) { hipGraphDestroy(g.graph); g.graph = nullptr; }
    g.stream = c->stream;
    const VcphipResidTile* dt = (const VcphipResidTile*)(b8 + o_t);
    const VcphipIgemmProblem* dp = (const VcphipIgemmProblem*)b8;
    g.dp = dp;
    unsigned* queue = batch_queue(c);
    if (!queue) {
      g.probs.clear();
      return VCPHIP_ERR_OOM;
    }
    auto launch = [&]() {
      return vcphip_launch_igemm_tiles(g.n64, g.n32, dt, dp, g.grid, g.grid32, queue, c->stream);
    };
    bool captured = false;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      const int lrc = launch();
      hipGraph_t gr = nullptr;
      if (hipStreamEndCapture(c->stream, &gr) == hipSuccess && lrc == 0 && gr &&
          graph_has_kernels(gr, (size_t)vcphip_igemm_tile_kernels(g.n64, g.n32))) {
        if (hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0) == hipSuccess) {
          g.graph = gr;
          captured = true;
        } else {
          hipGraphDestroy(gr);
        }
      } else if (gr) {
        hipGraphDestroy(gr);
      }
    }
    if (!captured) {
      hipGetLastError();
      g.exec = nullptr;
      StageTimer T(c);
      const int lrc = launch();
      T.mark("igemm_batched");
      g.probs.clear();
      return check_launch(lrc);
    }
  }
  StageTimer T(c);
  if (hipGraphLaunch(g.exec, c->stream) != hipSuccess) return VCPHIP_ERR_DEVICE;
  T.mark("igemm_batched");
  return VCPHIP_OK;
}

static bool same_problem(const VcphipResidProblem& a, const VcphipResidProblem& b) {
  return a.R == b.R && a.A == b.A && a.C == b.C && a.ldr == b.ldr && a.lda == b.lda && a.ldc == b.ldc && a.n == b.n;
}

extern "C" int vcphip_residual_batched_dev(vcphip_ctx* c, int count, const int* n, const double* const* R,
                                           const int* ldr, const double* const* IA, const int* lda, double* const* C,
                                           const int* ldc, double* normub_dev) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count < 0 || (count > 0 && (!n || !R || !ldr || !IA || !lda || !C || !ldc || !normub_dev)))
    return VCPHIP_ERR_ARGUMENT;
  if (count == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  std::vector<VcphipResidProblem> probs((size_t)count);
  int nrows = 0;
  for (int i = 0; i < count; ++i) {
    const int ni = n[i];
    if (ni < 0 || ldr[i] < (ni > 1 ? ni : 1) || lda[i] < (ni > 1 ? ni : 1) || ldc[i] < (ni > 1 ? ni : 1))
      return VCPHIP_ERR_DIMENSION;
    probs[i] = VcphipResidProblem{R[i], IA[i], C[i], ldr[i], lda[i], ldc[i], ni, nrows};
    nrows += ni;
  }
  vcphip_ctx::Resid& rs = c->resid;
  bool same = rs.exec && rs.normub == normub_dev && rs.stream == c->stream && rs.probs.size() == probs.size();
  for (size_t i = 0; same && i < probs.size(); ++i) same = same_problem(rs.probs[i], probs[i]);
  if (!same) {
    // tile list, largest problems first (longest-processing-time order)
    TileSet ts = resid_tiles(c, probs);
    rs.n64 = (int)ts.t64.size();
    rs.n32 = (int)ts.t32.size();
    rs.grid = batch_grid(c, ts.t64.size(), c->resid_per_cu);
    rs.grid32 = batch_grid(c, ts.t32.size(), c->resid_per_cu);
    rs.tiles = tile_array(ts);
    rs.row_problem.assign((size_t)nrows, 0);
    for (int i = 0; i < count; ++i)
      for (int r = 0; r < probs[i].n; ++r) rs.row_problem[(size_t)probs[i].row_base + r] = i;
    rs.probs = probs;
    const size_t bp = sizeof(VcphipResidProblem) * count, bt = sizeof(VcphipResidTile) * rs.tiles.size();
    const size_t br = sizeof(int) * (size_t)nrows, bs = sizeof(double) * (size_t)nrows;
    const size_t o_t = (bp + 255) & ~(size_t)255, o_r = o_t + ((bt + 255) & ~(size_t)255);
    const size_t o_s = o_r + ((br + 255) & ~(size_t)255), total = o_s + bs + 256;
    void* base;
    int rc = ws_get(c, vcphip_ctx::WS_BATCH, total, &base);
    if (rc) return rc;
    char* b8 = (char*)base;
    if (hipMemcpyAsync(b8, rs.probs.data(), bp, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        (bt && hipMemcpyAsync(b8 + o_t, rs.tiles.data(), bt, hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
        (br && hipMemcpyAsync(b8 + o_r, rs.row_problem.data(), br, hipMemcpyHostToDevice, c->stream) != hipSuccess))
      return VCPHIP_ERR_DEVICE;
    if (rs.exec) { hipGraphExecDestroy(rs.exec); rs.exec = nullptr; }
    if (rs.graph) { hipGraphDestroy(rs.graph); rs.graph = nullptr; }
    rs.normub = normub_dev;
    rs.stream = c->stream;
    // capture the persistent launches and the norm kernels once; replays reuse them
    const VcphipResidTile* dt = (const VcphipResidTile*)(b8 + o_t);
    const VcphipResidProblem* dp = (const VcphipResidProblem*)b8;
    const int* drp = (const int*)(b8 + o_r);
    double* drs = (double*)(b8 + o_s);
    rs.dt = dt;
    rs.dp = dp;
    rs.drp = drp;
    rs.drs = drs;
    rs.nrows = nrows;
    unsigned* queue = batch_queue(c);
    if (!queue) {
      rs.probs.clear();
      return VCPHIP_ERR_OOM;
    }
    bool captured = false;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      const int lrc = vcphip_launch_resid(rs.n64, rs.n32, dt, dp, nrows, drp, drs, count, normub_dev, rs.grid,
                                          rs.grid32, queue, c->stream);
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(c->stream, &g) == hipSuccess && lrc == 0 && g &&
          graph_has_kernels(g, (size_t)vcphip_resid_kernels(rs.n64, rs.n32, nrows, count))) {
        if (hipGraphInstantiate(&rs.exec, g, nullptr, nullptr, 0) == hipSuccess) {
          rs.graph = g;
          captured = true;
        } else {
          hipGraphDestroy(g);
        }
      } else if (g) {
        hipGraphDestroy(g);
      }
    }
    if (!captured) {
      hipGetLastError();
      rs.exec = nullptr;
      // a stream that cannot be captured (e.g. the legacy null stream) runs
      // the same kernels as direct launches
      StageTimer T(c);
      const int lrc = vcphip_launch_resid(rs.n64, rs.n32, dt, dp, nrows, drp, drs, count, normub_dev, rs.grid,
                                          rs.grid32, queue, c->stream);
      T.mark("resid");
      rs.probs.clear();   // nothing cached to replay
      return check_launch(lrc);
    }
  }
  StageTimer T(c);
  // the bounds fail closed: +inf until the replay's norm kernel overwrites
  // them, so a replay that runs nothing cannot pass as verified
  if (vcphip_fill_f64(normub_dev, count, __builtin_inf(), c->stream)) return VCPHIP_ERR_DEVICE;
  if (hipGraphLaunch(rs.exec, c->stream) != hipSuccess) return VCPHIP_ERR_DEVICE;
  T.mark("resid");
  return VCPHIP_OK;
}

// ===========================================================================
// approximate inverse of the midpoint matrix: imid + pdblas::inv
// (vcp/imats.hpp:298-301, vcp/pdblas.hpp:180-204)
// ===========================================================================
namespace {

// orders are bordered up to a multiple of this, so that problems of nearby
// order share one batched rocSOLVER call
constexpr int INV_BORDER = 64;

inline int inv_order(int n) { return (n + INV_BORDER - 1) / INV_BORDER * INV_BORDER; }

// the context's rocBLAS handle (created on first use), bound to its stream
int blas_on_stream(vcphip_ctx* c) {
  if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success) {
    c->blas = nullptr;
    return VCPHIP_ERR_DEVICE;
  }
  return rocblas_set_stream(c->blas, c->stream) == rocblas_status_success ? VCPHIP_OK : VCPHIP_ERR_DEVICE;
}

// the extra inverse lanes, created once; a context that cannot create them
// runs every group on its own stream (the results are the same)
int inv_lanes(vcphip_ctx* c) {
  if (c->inv_lanes >= 0) return c->inv_lanes;
  c->inv_lanes = 0;
  if (hipEventCreateWithFlags(&c->inv_fork, hipEventDisableTiming) != hipSuccess) {
    c->inv_fork = nullptr;
    return 0;
  }
  for (int l = 0; l < vcphip_ctx::INV_LANES; ++l) {
    if (hipStreamCreateWithFlags(&c->inv_s[l], hipStreamNonBlocking) != hipSuccess) {
      c->inv_s[l] = nullptr;
      break;
    }
    if (hipEventCreateWithFlags(&c->inv_join[l], hipEventDisableTiming) != hipSuccess) {
      c->inv_join[l] = nullptr;
      break;
    }
    if (rocblas_create_handle(&c->inv_h[l]) != rocblas_status_success) {
      c->inv_h[l] = nullptr;
      break;
    }
    if (rocblas_set_stream(c->inv_h[l], c->inv_s[l]) != rocblas_status_success) break;
    c->inv_lanes = l + 1;
  }
  return c->inv_lanes;
}

struct InvJob {
  int n;
  const double* IA;
  long lda;
  double* P;   // 2 npad^2 doubles: the LU factors (leading dimension npad), then the inverse
  int npad;    // inv_order(n)
  double* R;   // if not null: receives the leading n x n block of the inverse
  long ldr;
};

inline size_t inv_work_doubles(int n) { return 2 * (size_t)inv_order(n) * inv_order(n); }
inline double* inv_result(const InvJob& j) { return j.P + (size_t)j.npad * j.npad; }

struct InvBatch {
  std::vector<int> ord;          // jobs sorted by npad
  std::vector<double*> hptr;     // LU then inverse pointers in that order; the source of an async copy
  rocblas_int* info = nullptr;   // device: getrf's info per sorted position
};

// inv(mid(IA_j)) for every job, enqueued on the context stream: the midpoint
// matrix (kv's mid()) bordered by the identity to order npad (k_inv_pad),
// LU with partial pivoting (rocSOLVER dgetrf), then the inverse as the
// solution of L U X = I (dgetrs; LAPACK's dgetrf + dgetri in pdblas::inv
// compute the same inverse, and getrs is the faster route on the device),
// one batched call per padded order.  inv(M) is the leading n x n block of
// inv_result(job).  The caller synchronises the stream before `b` goes out
// of scope.
int inv_mid_enqueue(vcphip_ctx* c, const std::vector<InvJob>& jobs, InvBatch& b) {
  const int cnt = (int)jobs.size();
  hipStream_t s = c->stream;
  VCPHIP_TRY(blas_on_stream(c));
  b.ord.resize(cnt);
  for (int i = 0; i < cnt; ++i) b.ord[i] = i;
  std::stable_sort(b.ord.begin(), b.ord.end(), [&](int x, int y) { return jobs[x].npad < jobs[y].npad; });
  size_t npiv = 0;
  for (const auto& j : jobs) npiv += (size_t)j.npad;
  auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
  const size_t o_piv = al(sizeof(double*) * 2 * cnt), o_info = o_piv + al(sizeof(rocblas_int) * npiv);
  void* base;
  const int wrc = ws_get(c, vcphip_ctx::WS_INV, o_info + al(sizeof(rocblas_int) * cnt), &base);
  if (wrc) return wrc;
  char* b8 = (char*)base;
  double** dptr = (double**)b8;   // LU pointers, then inverse pointers
  rocblas_int* piv = (rocblas_int*)(b8 + o_piv);
  b.info = (rocblas_int*)(b8 + o_info);
  b.hptr.resize(2 * (size_t)cnt);
  for (int t = 0; t < cnt; ++t) {
    b.hptr[t] = jobs[b.ord[t]].P;
    b.hptr[cnt + t] = inv_result(jobs[b.ord[t]]);
  }
  if (hipMemcpyAsync(dptr, b.hptr.data(), sizeof(double*) * 2 * cnt, hipMemcpyHostToDevice, s) != hipSuccess)
    return VCPHIP_ERR_DEVICE;
  for (const auto& j : jobs)
    if (vcphip_inv_launch_pad(j.n, j.npad, j.IA, j.lda, j.P, s) ||
        vcphip_inv_launch_pad(0, j.npad, nullptr, 1, inv_result(j), s))   // the identity
      return VCPHIP_ERR_DEVICE;
  // groups of equal padded order: sorted position range and pivot offset
  struct Group {
    int t, e;
    size_t poff;
    double cost;
    int lane;
  };
  std::vector<Group> groups;
  size_t poff = 0;
  for (int t = 0; t < cnt;) {
    const int np = jobs[b.ord[t]].npad;
    int e = t + 1;
    while (e < cnt && jobs[b.ord[e]].npad == np) ++e;
    groups.push_back(Group{t, e, poff, (double)(e - t) * np * np * np, 0});
    poff += (size_t)np * (e - t);
    t = e;
  }
  // the groups over the lanes, the most work first to the least loaded lane;
  // lane 0 is the context stream
  const int lanes = groups.size() > 1 ? 1 + std::min((int)groups.size() - 1, inv_lanes(c)) : 1;
  if (lanes > 1) {
    std::vector<int> by_cost(groups.size());
    for (size_t g = 0; g < groups.size(); ++g) by_cost[g] = (int)g;
    std::stable_sort(by_cost.begin(), by_cost.end(), [&](int x, int y) { return groups[x].cost > groups[y].cost; });
    double load[1 + vcphip_ctx::INV_LANES] = {};
    for (int g : by_cost) {
      int best = 0;
      for (int l = 1; l < lanes; ++l)
        if (load[l] < load[best]) best = l;
      groups[g].lane = best;
      load[best] += groups[g].cost;
    }
    // the extra lanes start after everything enqueued so far on the context stream
    if (hipEventRecord(c->inv_fork, s) != hipSuccess) return VCPHIP_ERR_DEVICE;
    for (int l = 1; l < lanes; ++l)
      if (hipStreamWaitEvent(c->inv_s[l - 1], c->inv_fork, 0) != hipSuccess) return VCPHIP_ERR_DEVICE;
  }
  int rc = VCPHIP_OK;
  for (const Group& g : groups) {
    const int np = jobs[b.ord[g.t]].npad, bc = g.e - g.t;
    rocblas_handle h = g.lane ? c->inv_h[g.lane - 1] : c->blas;
    // the batched routines also for a group of one: their rounding differs
    // from the plain ones', and a problem's R must not depend on its group
    rocblas_status st = rocsolver_dgetrf_batched(h, np, np, dptr + g.t, np, piv + g.poff, np, b.info + g.t, bc);
    if (st == rocblas_status_success)
      st = rocsolver_dgetrs_batched(h, rocblas_operation_none, np, np, dptr + g.t, np, piv + g.poff, np,
                                    dptr + cnt + g.t, np, bc);
    if (st != rocblas_status_success) {
      rc = VCPHIP_ERR_DEVICE;
      break;
    }
  }
  // the context stream waits for every lane, also after a failure, so that
  // synchronising it covers all work on the workspace
  for (int l = 1; l < lanes; ++l)
    if (hipEventRecord(c->inv_join[l - 1], c->inv_s[l - 1]) != hipSuccess ||
        hipStreamWaitEvent(s, c->inv_join[l - 1], 0) != hipSuccess) {
      hipStreamSynchronize(c->inv_s[l - 1]);
      rc = VCPHIP_ERR_DEVICE;
    }
  if (rc) return rc;
  for (const auto& j : jobs)
    if (j.R && vcphip_inv_launch_copy(j.n, inv_result(j), j.npad, j.R, j.ldr, s)) return VCPHIP_ERR_DEVICE;
  return VCPHIP_OK;
}

// after the stream has been synchronised: singular[j] = 1 when job j met an
// exactly zero pivot (LAPACK's info > 0, the reference's lapack_error)
int inv_mid_singular(vcphip_ctx* c, const InvBatch& b, std::vector<char>& singular) {
  const int cnt = (int)b.ord.size();
  std::vector<rocblas_int> info((size_t)cnt);
  if (hipMemcpy(info.data(), b.info, sizeof(rocblas_int) * cnt, hipMemcpyDeviceToHost) != hipSuccess)
    return VCPHIP_ERR_DEVICE;
  singular.assign(cnt, 0);
  for (int t = 0; t < cnt; ++t) singular[b.ord[t]] = info[t] != 0 ? 1 : 0;
  return VCPHIP_OK;
}

}  // namespace

extern "C" int vcphip_inv_mid_dev(vcphip_ctx* c, int n, const double* IA, int lda, double* R, int ldr) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || lda < (n > 1 ? n : 1) || ldr < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  if (!IA || !R) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  const int np = inv_order(n);
  void* pw;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_LSS, sizeof(double) * inv_work_doubles(n), &pw));
  StageTimer T(c);
  InvBatch b;
  const int rc = inv_mid_enqueue(c, {InvJob{n, IA, lda, (double*)pw, np, R, ldr}}, b);
  T.mark("inv_mid");
  if (hipStreamSynchronize(c->stream) != hipSuccess) return VCPHIP_ERR_DEVICE;
  if (rc) return rc;
  std::vector<char> singular;
  VCPHIP_TRY(inv_mid_singular(c, b, singular));
  return singular[0] ? VCPHIP_ERR_SINGULAR : VCPHIP_OK;
}

// ===========================================================================
// dd linear solve: pddblas::linearsolve (vcp/pddblas.hpp:412-474) on the device
// ===========================================================================
extern "C" int vcphip_dd_solve_dev(vcphip_ctx* c, int n, int nrhs, const double* A, int lda, const double* b, int ldb,
                                   double* x, int ldx, int itmax, int* iters) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || nrhs < 0 || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1) || ldx < (n > 1 ? n : 1))
    return VCPHIP_ERR_DIMENSION;
  if (iters) *iters = 0;
  if (n == 0 || nrhs == 0) return VCPHIP_OK;
  if (!A || !b || !x || itmax < 0) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  // workspace: LU of hi(A) (n x n), the correction d (n x nrhs), the residual
  // r (dd, n x nrhs), the block maxima and the two norms, pivots and info
  auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
  const size_t nn = (size_t)n * n, nr = (size_t)n * nrhs;
  const size_t o_d = al(8 * nn), o_r = o_d + al(8 * nr), o_part = o_r + al(16 * nr);
  const size_t o_norm = o_part + al(16 * (size_t)VCPHIP_DD_ADD_BLOCKS), o_piv = o_norm + 256;
  const size_t o_info = o_piv + al(sizeof(rocblas_int) * n), total = o_info + 256;
  void* base;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_DDS, total, &base));
  char* b8 = (char*)base;
  double *LU = (double*)b8, *d = (double*)(b8 + o_d), *r = (double*)(b8 + o_r);
  double *part = (double*)(b8 + o_part), *norms = (double*)(b8 + o_norm);
  rocblas_int *piv = (rocblas_int*)(b8 + o_piv), *info = (rocblas_int*)(b8 + o_info);
  auto fail = [&](int code) {
    hipStreamSynchronize(s);
    return code;
  };
  VCPHIP_TRY(blas_on_stream(c));
  StageTimer T(c);
  // LU of double(A), reused in every round (:423-432)
  if (vcphip_dd_hi(n, n, A, lda, LU, s) || rocsolver_dgetrf(c->blas, n, n, LU, n, piv, info) != rocblas_status_success)
    return fail(VCPHIP_ERR_DEVICE);
  rocblas_int hinfo = 0;
  if (hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return VCPHIP_ERR_DEVICE;
  if (hinfo != 0) return VCPHIP_ERR_SINGULAR;   // the reference's lapack_error("dgetrf")
  // the row interchanges as one permutation, applied while taking hi(r):
  // LAPACK's getrs would apply them with a row-swap pass per call, which
  // with n right-hand sides (the inverse) took longer than the two
  // triangular solves
  {
    std::vector<rocblas_int> hp(n);
    if (hipMemcpyAsync(hp.data(), piv, sizeof(rocblas_int) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
      const int j = hp[i] - 1;
      if (j < i || j >= n) return VCPHIP_ERR_DEVICE;   // getrf pivots lie in [i + 1, n] (1-based)
      std::swap(perm[i], perm[j]);
    }
    if (hipMemcpyAsync(piv, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
  }
  static_assert(sizeof(rocblas_int) == sizeof(int), "the permutation reuses the pivot buffer");
  const int* rowperm = (const int*)piv;
  const double one = 1.0;
  T.mark("dd_solve_lu");
  // x = 0, r = b (:438-439)
  if (hipMemset2DAsync(x, 16 * (size_t)ldx, 0, 16 * (size_t)n, nrhs, s) != hipSuccess) return fail(VCPHIP_ERR_DEVICE);
  const double* rc = b;
  long rld = ldb;
  double prev = HUGE_VAL;
  int it = 0;
  while (it < itmax) {
    // d = LU \ hi(r); x += dd(d); the norms (:446-458)
    if (vcphip_dd_hi_rows(n, nrhs, rc, rld, rowperm, d, s) ||
        rocblas_dtrsm(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit,
                      n, nrhs, &one, LU, n, d, n) != rocblas_status_success ||
        rocblas_dtrsm(c->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                      rocblas_diagonal_non_unit, n, nrhs, &one, LU, n, d, n) != rocblas_status_success ||
        vcphip_dd_add_d(n, nrhs, d, x, ldx, part, norms, s))
      return fail(VCPHIP_ERR_DEVICE);
    double h[2];
    if (hipMemcpyAsync(h, norms, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
    ++it;
    const double normd = h[0], normx = h[1];
    // converged to dd accuracy or stagnated (:460)
    if (normd == 0.0 || normd <= 0.25e-32 * normx || normd >= 0.5 * prev) break;
    prev = normd;
    // r = b - A x with the dd Ozaki product (:463-470)
    const int rrc = vcphip_dd_residual_dev(c, n, nrhs, n, A, lda, x, ldx, b, ldb, r, n);
    if (rrc) return fail(rrc);
    rc = r;
    rld = n;
  }
  T.mark("dd_solve_refine");
  if (iters) *iters = it;
  return VCPHIP_OK;
}

// ===========================================================================
// dd Cholesky: pddblas::Cholesky (vcp/pddblas.hpp:488-563) on the device
// ===========================================================================
extern "C" int vcphip_dd_cholesky_dev(vcphip_ctx* c, int n, const double* A, int lda, double* U, int ldu, int itep) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || lda < (n > 1 ? n : 1) || ldu < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  if (!A || !U || itep < 0) return VCPHIP_ERR_ARGUMENT;
  if (n > VCPHIP_OZ_MAXK) return VCPHIP_ERR_UNSUPPORTED;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  // workspace: Ud and G (n x n doubles), P and W (n x n dd), flag / info
  auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
  const size_t nn = (size_t)n * n;
  const size_t o_g = al(8 * nn), o_p = o_g + al(8 * nn), o_w = o_p + al(16 * nn), o_i = o_w + al(16 * nn);
  void* base;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_DDS, o_i + 256, &base));
  char* b8 = (char*)base;
  double *Ud = (double*)b8, *G = (double*)(b8 + o_g), *P = (double*)(b8 + o_p), *W = (double*)(b8 + o_w);
  rocblas_int* info = (rocblas_int*)(b8 + o_i);
  int* flag = (int*)(b8 + o_i + 64);
  auto fail = [&](int code) {
    hipStreamSynchronize(s);
    return code;
  };
  VCPHIP_TRY(blas_on_stream(c));
  StageTimer T(c);
  // the reference's is_symmetric (:491-493), then dpotrf of double(A) (:496-498)
  if (hipMemsetAsync(flag, 0, sizeof(int), s) != hipSuccess || vcphip_dd_symmetric(n, A, lda, flag, s) ||
      vcphip_dd_hi(n, n, A, lda, Ud, s) ||
      rocsolver_dpotrf(c->blas, rocblas_fill_upper, n, Ud, n, info) != rocblas_status_success)
    return fail(VCPHIP_ERR_DEVICE);
  int h[2] = {0, 0};
  if (hipMemcpyAsync(&h[0], flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&h[1], info, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return VCPHIP_ERR_DEVICE;
  if (h[0]) return VCPHIP_ERR_ARGUMENT;    // the reference's domain_error
  if (h[1]) return VCPHIP_ERR_SINGULAR;    // not positive definite: lapack_error("dpotrf")
  if (vcphip_dd_from_upper(n, Ud, n, U, ldu, s)) return fail(VCPHIP_ERR_DEVICE);
  T.mark("dd_chol_potrf");
  const double one = 1.0;
  for (int it = 0; it < itep; ++it) {
    // E = A - U^T U with the dd product (:510-518)
    int rc = dd_gemm_dev(c, true, n, n, n, U, ldu, U, ldu, P, n);
    if (rc) return fail(rc);
    // G = Ud^-T E Ud^-1 in double precision, Ud = double(U) (:520-529)
    if (vcphip_dd_sub_from(n, n, A, lda, P, n, s) || vcphip_dd_hi(n, n, P, n, G, s) ||
        vcphip_dd_hi(n, n, U, ldu, Ud, s) ||
        rocblas_dtrsm(c->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                      rocblas_diagonal_non_unit, n, n, &one, Ud, n, G, n) != rocblas_status_success ||
        rocblas_dtrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                      rocblas_diagonal_non_unit, n, n, &one, Ud, n, G, n) != rocblas_status_success ||
        vcphip_dd_w_from_g(n, G, W, s))
      return fail(VCPHIP_ERR_DEVICE);
    // U = U + W U (:531-550)
    rc = dd_gemm_dev(c, false, n, n, n, W, n, U, ldu, P, n);
    if (rc) return fail(rc);
    if (vcphip_dd_add_to(n, P, n, U, ldu, 1, s)) return fail(VCPHIP_ERR_DEVICE);
  }
  T.mark("dd_chol_refine");
  return hipStreamSynchronize(s) == hipSuccess ? VCPHIP_OK : VCPHIP_ERR_DEVICE;
}

// ===========================================================================
// dd symmetric eigenproblems: pddblas::eigsym / eigsymge (vcp/pddblas.hpp:
// 219-314, 566-659) on the device
// ===========================================================================
namespace {

// B == nullptr: A x = lam x; else A x = lam B x (B symmetric positive definite)
int dd_eig_dev(vcphip_ctx* c, int n, const double* A, int lda, const double* B, int ldb, double* lam, double* X,
               int ldx, int itep) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || lda < (n > 1 ? n : 1) || ldx < (n > 1 ? n : 1) || (B && ldb < (n > 1 ? n : 1)))
    return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  if (!A || !lam || !X || itep < 0) return VCPHIP_ERR_ARGUMENT;
  if (n > VCPHIP_OZ_MAXK) return VCPHIP_ERR_UNSUPPORTED;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  // workspace: AX, S, G (n x n dd each), B X for the generalized problem,
  // then D and the tridiagonal work (n each), lam (dd), row sums, delta, flags
  auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
  const size_t nn = (size_t)n * n, m2 = al(16 * nn);
  const size_t o_s = m2, o_g = 2 * m2, o_bx = 3 * m2, o_d = o_bx + (B ? m2 : 0);
  const size_t o_e = o_d + al(8 * (size_t)n), o_lam = o_e + al(8 * (size_t)n), o_rows = o_lam + al(16 * (size_t)n);
  const size_t o_misc = o_rows + al(16 * (size_t)n), total = o_misc + 256;
  void* base;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_DDS, total, &base));
  char* b8 = (char*)base;
  double *AX = (double*)b8, *S = (double*)(b8 + o_s), *G = (double*)(b8 + o_g);
  double* BX = B ? (double*)(b8 + o_bx) : nullptr;
  double *D = (double*)(b8 + o_d), *Ew = (double*)(b8 + o_e), *lw = (double*)(b8 + o_lam);
  double* rows = (double*)(b8 + o_rows);
  double* delta = (double*)(b8 + o_misc);
  rocblas_int* info = (rocblas_int*)(b8 + o_misc + 64);
  int* flag = (int*)(b8 + o_misc + 128);
  auto fail = [&](int code) {
    hipStreamSynchronize(s);
    return code;
  };
  VCPHIP_TRY(blas_on_stream(c));
  StageTimer T(c);
  // the reference's symmetry checks, then dsyev / dsygv of the high parts
  // (:567-580, :612-630); Ad and Bd live in the AX and S buffers until the
  // first step
  double *Ad = AX, *Bd = S;
  if (hipMemsetAsync(flag, 0, sizeof(int), s) != hipSuccess || vcphip_dd_symmetric(n, A, lda, flag, s) ||
      (B && vcphip_dd_symmetric(n, B, ldb, flag, s)) || vcphip_dd_hi(n, n, A, lda, Ad, s) ||
      (B && vcphip_dd_hi(n, n, B, ldb, Bd, s)))
    return fail(VCPHIP_ERR_DEVICE);
  const rocblas_status st =
      B ? rocsolver_dsygvd(c->blas, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n, Ad, n, Bd, n, D,
                           Ew, info)
        : rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, n, Ad, n, D, Ew, info);
  if (st != rocblas_status_success) return fail(VCPHIP_ERR_DEVICE);
  int h[2] = {0, 0};
  if (hipMemcpyAsync(&h[0], flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&h[1], info, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return VCPHIP_ERR_DEVICE;
  if (h[0]) return VCPHIP_ERR_ARGUMENT;   // the reference's domain_error
  if (h[1]) return VCPHIP_ERR_SINGULAR;   // no convergence, or B not positive definite: lapack_error
  if (vcphip_dd_from_d(n, n, Ad, n, X, ldx, s) || vcphip_dd_from_d(n, 1, D, n, lam, n, s))
    return fail(VCPHIP_ERR_DEVICE);
  T.mark("dd_eig_start");
  // max(itep, 2) Ogita-Aishima steps (:586-603, :640-656)
  const int steps = itep > 2 ? itep : 2;
  for (int it = 0; it < steps; ++it) {
    int rc = dd_gemm_dev(c, false, n, n, n, A, lda, X, ldx, AX, n);                    // A X
    if (!rc) rc = dd_gemm_dev(c, false, n, n, n, X, ldx, AX, n, S, n, true);           // X^T (A X)
    if (!rc && B) rc = dd_gemm_dev(c, false, n, n, n, B, ldb, X, ldx, BX, n);          // B X
    if (!rc) rc = B ? dd_gemm_dev(c, false, n, n, n, X, ldx, BX, n, G, n, true)         // X^T (B X)
                    : dd_gemm_dev(c, true, n, n, n, X, ldx, X, ldx, G, n);             // X^T X
    if (rc) return fail(rc);
    // lam, the norms, delta and E (into AX); X += X E (X E into S)
    if (vcphip_oa_elementwise(n, S, n, G, n, lw, rows, delta, AX, s)) return fail(VCPHIP_ERR_DEVICE);
    rc = dd_gemm_dev(c, false, n, n, n, X, ldx, AX, n, S, n);
    if (rc) return fail(rc);
    if (vcphip_dd_add_to(n, S, n, X, ldx, 0, s)) return fail(VCPHIP_ERR_DEVICE);
  }
  // the eigenvalues of the last step (computed from X before its update)
  if (hipMemcpyAsync(lam, lw, 16 * (size_t)n, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail(VCPHIP_ERR_DEVICE);
  T.mark("dd_eig_refine");
  return hipStreamSynchronize(s) == hipSuccess ? VCPHIP_OK : VCPHIP_ERR_DEVICE;
}

}  // namespace

extern "C" int vcphip_dd_eigsym_dev(vcphip_ctx* c, int n, const double* A, int lda, double* lam, double* X, int ldx,
                                    int itep) {
  return dd_eig_dev(c, n, A, lda, nullptr, 0, lam, X, ldx, itep);
}

extern "C" int vcphip_dd_eigsymge_dev(vcphip_ctx* c, int n, const double* A, int lda, const double* B, int ldb,
                                      double* lam, double* X, int ldx, int itep) {
  if (c && n > 0 && !B) return VCPHIP_ERR_ARGUMENT;
  return dd_eig_dev(c, n, A, lda, B, ldb, lam, X, ldx, itep);
}

// ===========================================================================
// verified linear solve (imats::linearsolve, vcp/imats.hpp:292-362)
// ===========================================================================
extern "C" int vcphip_lss_dev(vcphip_ctx* c, int n, int nrhs, const double* IA, int lda, const double* Ib, int ldb,
                              const double* R, int ldr, double* X, int ldx, double* normub) {
  if (!c || !IA || !Ib || !X) return VCPHIP_ERR_ARGUMENT;
  if (n < 2 || nrhs < 1 || lda < n || ldb < n || (R && ldr < n) || ldx < n) return VCPHIP_ERR_DIMENSION;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const size_t nr = (size_t)n * nrhs, al = 32;   // 256-byte alignment in doubles
  auto up = [&](size_t x) { return (x + al - 1) / al * al; };
  const size_t o_mb = 0, o_mx = o_mb + up(nr), o_x = o_mx + up(nr), o_y = o_x + up(2 * nr), o_t = o_y + up(2 * nr);
  const size_t o_w = o_t + up(2 * nr), o_c = o_w + up(2 * nr), o_r = o_c + up(2 * (size_t)n * n);
  const size_t np = (size_t)inv_order(n);
  const size_t o_s = o_r + up((size_t)n), o_ri = o_s + al, total = o_ri + (R ? 0 : up(inv_work_doubles(n)));
  void* base;
  const int wrc = ws_get(c, vcphip_ctx::WS_LSS, total * sizeof(double), &base);
  if (wrc) return wrc;
  double* w = (double*)base;
  double *mb = w + o_mb, *mx = w + o_mx, *Xi = w + o_x, *Y = w + o_y, *Tm = w + o_t, *W = w + o_w, *C = w + o_c;
  double *rows = w + o_r, *norms = w + o_s;   // norms: g, T, G lower, G upper
  StageTimer T(c);
  // R = inv(mid(IA)) on the device when the caller gives none (:298-301)
  InvBatch inv;
  if (!R) {
    const InvJob job{n, IA, lda, w + o_ri, (int)np, nullptr, 0};
    const int irc = inv_mid_enqueue(c, {job}, inv);
    if (irc) {
      hipStreamSynchronize(s);
      return irc;
    }
    R = inv_result(job);   // inv(mid(IA)) is the leading n x n block of the bordered inverse
    ldr = (int)np;
    T.mark("lss_inv");
  }
  int rc = 0;
  // mx = R * mid(b), nearest rounding (:299-302)
  rc |= vcphip_lss_launch_mid(n, nrhs, Ib, ldb, mb, n, s);
  rc |= vcphip_lss_launch_gemv_rn(n, nrhs, n, R, ldr, mb, n, mx, n, s);
  // y = R * (IA*mx - b) and |y| (:305-315)
  rc |= vcphip_launch_drgemm(VCPHIP_DR_IP, n, nrhs, n, IA, lda, mx, n, Xi, n, nullptr, 0, s);
  rc |= vcphip_lss_launch_sub(n, nrhs, Xi, n, Ib, ldb, s);
  rc |= vcphip_launch_drgemm(VCPHIP_DR_PI, n, nrhs, n, R, ldr, Xi, n, Y, n, nullptr, 0, s);
  rc |= vcphip_lss_launch_abs(n, nrhs, Y, n, s);
  T.mark("lss_rhs");
  // R*IA - I and g = ||R*IA - I||_inf (:318-329)
  rc |= planes_product(c, VCPHIP_DR_PI, n, n, n, R, ldr, IA, lda, C, n) ? 1 : 0;   // on s, the context stream
  rc |= vcphip_lss_launch_diag(n, C, n, s);
  rc |= vcphip_lss_launch_norm(n, n, C, n, rows, norms, s);
  T.mark("lss_resid");
  // T = ||y||_inf, G = T / (1 - g), |R*IA - I| * G (:334-350)
  rc |= vcphip_lss_launch_norm(n, nrhs, Y, n, rows, norms + 1, s);
  rc |= vcphip_lss_launch_abs(n, n, C, n, s);
  rc |= vcphip_lss_launch_fill_g(n, nrhs, norms, Tm, n, norms + 2, s);
  rc |= vcphip_launch_drgemm(VCPHIP_DR_II, n, nrhs, n, C, n, Tm, n, W, n, nullptr, 0, s);
  // X = mx + [-u, u], u = |y| + W (:352-361)
  rc |= vcphip_lss_launch_final(n, nrhs, Y, W, mx, X, ldx, s);
  T.mark("lss_bound");
  if (rc) {
    hipStreamSynchronize(s);
    return VCPHIP_ERR_DEVICE;
  }
  double g = 0.0;
  if (hipMemcpyAsync(&g, norms, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return VCPHIP_ERR_DEVICE;
  if (normub) *normub = g;
  if (!inv.ord.empty()) {
    std::vector<char> singular;
    VCPHIP_TRY(inv_mid_singular(c, inv, singular));
    if (singular[0]) return VCPHIP_ERR_SINGULAR;
  }
  return (g < 1.0) ? VCPHIP_OK : VCPHIP_ERR_VERIFICATION;
}

// Many systems of uneven size: R_p IA_p - I and its norm through the batched
// residual (its persistent launch and graph), the other steps as one launch
// each over the (row, column) items of all problems.
extern "C" int vcphip_lss_batched_dev(vcphip_ctx* c, int count, const int* n, const int* nrhs,
                                      const double* const* IA, const int* lda, const double* const* Ib,
                                      const int* ldb, const double* const* R, const int* ldr, double* const* X,
                                      const int* ldx, double* normub) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count < 0 || (count > 0 && (!n || !nrhs || !IA || !lda || !Ib || !ldb || !X || !ldx || !normub)))
    return VCPHIP_ERR_ARGUMENT;
  if (count == 0) return VCPHIP_OK;
  // R (the array) or R[p] may be NULL: that problem's R = inv(mid(IA_p)) is
  // formed on the device (imats.hpp:298-301)
  auto has_r = [&](int p) { return R && R[p]; };
  for (int p = 0; p < count; ++p) {
    const int np = n[p];
    if (!IA[p] || !Ib[p] || !X[p]) return VCPHIP_ERR_ARGUMENT;
    if (np < 2 || nrhs[p] < 1 || lda[p] < np || ldb[p] < np || (has_r(p) && (!ldr || ldr[p] < np)) || ldx[p] < np)
      return VCPHIP_ERR_DIMENSION;
  }
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  // workspace: descriptors, item offsets, g and norms, then per problem
  // mb, mx, xi, y (n x nrhs) and C (n x n intervals); 256-byte aligned
  auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
  const size_t o_pb = al(sizeof(VcphipLssProblem) * count);
  const size_t o_g = o_pb + al(sizeof(long) * (count + 1));
  const size_t o_nm = o_g + al(sizeof(double) * count);
  size_t off = o_nm + al(sizeof(double) * 4 * (size_t)count);
  std::vector<size_t> o_prob((size_t)count);
  std::vector<long> pb((size_t)count + 1, 0);
  std::vector<size_t> o_rinv((size_t)count, 0);
  for (int p = 0; p < count; ++p) {
    const size_t nr = (size_t)n[p] * nrhs[p];
    o_prob[p] = off;
    off += al(8 * nr) * 2 + al(16 * nr) * 2 + al(16 * (size_t)n[p] * n[p]);
    if (!has_r(p)) {
      o_rinv[p] = off;
      off += al(8 * inv_work_doubles(n[p]));
    }
    pb[p + 1] = pb[p] + (long)nr;
  }
  void* base;
  const int wrc = ws_get(c, vcphip_ctx::WS_LSSB, off, &base);
  if (wrc) return wrc;
  char* b8 = (char*)base;
  double* gdev = (double*)(b8 + o_g);
  std::vector<VcphipLssProblem> probs((size_t)count);
  std::vector<double*> Cp((size_t)count);
  std::vector<int> ldc((size_t)count), ldr_use((size_t)count);
  std::vector<const double*> R_use((size_t)count);
  std::vector<InvJob> jobs;
  for (int p = 0; p < count; ++p) {
    if (has_r(p)) {
      R_use[p] = R[p];
      ldr_use[p] = ldr[p];
    } else {
      const InvJob job{n[p], IA[p], lda[p], (double*)(b8 + o_rinv[p]), inv_order(n[p]), nullptr, 0};
      R_use[p] = inv_result(job);   // the leading n x n block of the bordered inverse
      ldr_use[p] = job.npad;
      jobs.push_back(job);
    }
  }
  for (int p = 0; p < count; ++p) {
    const size_t nr = (size_t)n[p] * nrhs[p];
    char* q = b8 + o_prob[p];
    VcphipLssProblem& P = probs[p];
    P.A = IA[p];
    P.b = Ib[p];
    P.R = R_use[p];
    P.X = X[p];
    P.lda = lda[p];
    P.ldb = ldb[p];
    P.ldr = ldr_use[p];
    P.ldx = ldx[p];
    P.n = n[p];
    P.nrhs = nrhs[p];
    P.mb = (double*)q;
    P.mx = (double*)(q + al(8 * nr));
    P.xi = (double*)(q + 2 * al(8 * nr));
    P.y = (double*)(q + 2 * al(8 * nr) + al(16 * nr));
    Cp[p] = (double*)(q + 2 * al(8 * nr) + 2 * al(16 * nr));
    P.C = Cp[p];
    P.normub = gdev + p;
    P.norms = (double*)(b8 + o_nm) + 4 * (size_t)p;
    ldc[p] = n[p];
  }
  // the host vectors above are the sources of asynchronous copies: every
  // return from here on waits for the stream first
  auto fail = [&](int code) {
    hipStreamSynchronize(s);
    if (c->aux) hipStreamSynchronize(c->aux);
    return code;
  };
  if (hipMemcpyAsync(b8, probs.data(), sizeof(VcphipLssProblem) * count, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b8 + o_pb, pb.data(), sizeof(long) * (count + 1), hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(VCPHIP_ERR_DEVICE);
  if (!c->aux) {
    if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
      return fail(VCPHIP_ERR_DEVICE);
  }
  const long* dpb = (const long*)(b8 + o_pb);
  const VcphipLssProblem* dprobs = (const VcphipLssProblem*)b8;
  // the missing approximate inverses first, on the context stream
  InvBatch inv;
  if (!jobs.empty()) {
    StageTimer T(c);
    const int irc = inv_mid_enqueue(c, jobs, inv);
    if (irc) return fail(irc);
    T.mark("lss_inv");
  }
  // right-hand side on the second stream, after the descriptor upload
  if (hipEventRecord(c->ev_fork, s) != hipSuccess || hipStreamWaitEvent(c->aux, c->ev_fork, 0) != hipSuccess ||
      vcphip_lss_launch_batched_rhs(count, pb[count], dpb, dprobs, c->aux) ||
      hipEventRecord(c->ev_join, c->aux) != hipSuccess)
    return fail(VCPHIP_ERR_DEVICE);
  // R_p IA_p - I and g_p (imats.hpp:318-329) meanwhile on the context stream
  int rc = vcphip_residual_batched_dev(c, count, n, R_use.data(), ldr_use.data(), IA, lda, Cp.data(), ldc.data(), gdev);
  if (rc) return fail(rc);
  StageTimer T(c);
  if (hipStreamWaitEvent(s, c->ev_join, 0) != hipSuccess ||
      vcphip_lss_launch_batched_bound(count, pb[count], dpb, dprobs, s))
    return fail(VCPHIP_ERR_DEVICE);
  T.mark("lss_batched");
  if (hipMemcpyAsync(normub, gdev, sizeof(double) * count, hipMemcpyDeviceToHost, s) != hipSuccess)
    return fail(VCPHIP_ERR_DEVICE);
  if (hipStreamSynchronize(s) != hipSuccess) return VCPHIP_ERR_DEVICE;
  if (!jobs.empty()) {
    std::vector<char> singular;
    VCPHIP_TRY(inv_mid_singular(c, inv, singular));
    for (size_t j = 0; j < jobs.size(); ++j)
      if (singular[j]) return VCPHIP_ERR_SINGULAR;
  }
  for (int p = 0; p < count; ++p)
    if (!(normub[p] < 1.0)) return VCPHIP_ERR_VERIFICATION;
  return VCPHIP_OK;
}

// Galerkin matrix of Legendre_Bases_Generator::uhphiphi (vcp/ldbase.hpp:955-1013)
extern "C" int vcphip_uhphiphi_dev(vcphip_ctx* c, int m, int npts, const double* phi, int ldp, const double* w,
                                   const double* u, double* A, int lda) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (m < 0 || npts < 0 || ldp < (m > 1 ? m : 1) || lda < (m > 1 ? m : 1)) return VCPHIP_ERR_DIMENSION;
  if (m == 0) return VCPHIP_OK;
  if (!phi || !w || !u || !A) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  StageTimer T(c);
  const int rc = check_launch(vcphip_launch_uhphiphi(m, npts, phi, ldp, w, u, A, lda, c->stream));
  T.mark("uhphiphi");
  return rc;
}

// ===========================================================================
// host-pointer entry points: stage through the context workspace
// ===========================================================================
namespace {

// copy a column-major matrix of `rows` x `cols` elements of `esize` bytes
// Large host <-> device copies go through two pinned buffers of PIN_BYTES:
// the host threads pack a block of columns into one buffer while the DMA of
// the previous block runs from the other.  Copies from pageable memory run
// at a fifth of the PCIe rate otherwise.
constexpr size_t PIN_BYTES = (size_t)128 << 20;
constexpr size_t PIN_MIN = (size_t)8 << 20;   // below this, one direct copy

bool pin_ready(vcphip_ctx* c) {
  for (int h = 0; h < 2; ++h) {
    if (!c->pin[h] && hipHostMalloc(&c->pin[h], PIN_BYTES, hipHostMallocDefault) != hipSuccess) {
      hipGetLastError();
      c->pin[h] = nullptr;
      return false;
    }
    if (!c->pin_ev[h] && hipEventCreateWithFlags(&c->pin_ev[h], hipEventDisableTiming) != hipSuccess) {
      hipGetLastError();
      c->pin_ev[h] = nullptr;
      return false;
    }
  }
  return true;
}

// columns [0, ncols) of `rowb` bytes each, from stride `sld` to stride `dld`
// (bytes), split over host threads
void copy_cols(char* dst, size_t dld, const char* src, size_t sld, size_t rowb, int ncols) {
  const size_t total = rowb * (size_t)ncols;
  unsigned nt = std::thread::hardware_concurrency();
  nt = nt < 1 ? 1 : (nt > 8 ? 8 : nt);   // eight memcpy streams already outrun PCIe
  const bool dense = (dld == rowb && sld == rowb);
  if (total < ((size_t)4 << 20) || nt == 1 || (!dense && ncols < 2)) {
    if (dense) memcpy(dst, src, total);
    else
      for (int j = 0; j < ncols; ++j) memcpy(dst + dld * j, src + sld * j, rowb);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nt);
  if (dense) {   // one contiguous block, split by bytes
    for (unsigned t = 0; t < nt; ++t) {
      const size_t b0 = total * t / nt, b1 = total * (t + 1) / nt;
      pool.emplace_back([=]() { memcpy(dst + b0, src + b0, b1 - b0); });
    }
  } else {
    if ((int)nt > ncols) nt = (unsigned)ncols;
    for (unsigned t = 0; t < nt; ++t) {
      const int j0 = (int)((long)ncols * t / nt), j1 = (int)((long)ncols * (t + 1) / nt);
      pool.emplace_back([=]() {
        for (int j = j0; j < j1; ++j) memcpy(dst + dld * j, src + sld * j, rowb);
      });
    }
  }
  for (auto& th : pool) th.join();
}

// Many host matrices into one packed device region (the batched entry
// points): piece q is ncols columns of rowb bytes at source stride sld, to be
// stored dense at byte offset doff of `dev`.  Offsets increase with q and do
// not overlap.  Large inputs are packed into the pinned buffers, a device
// range at a time, one DMA per range; below PIN_MIN, one copy per piece.
struct HostPiece {
  const char* src;
  size_t sld, rowb;
  int ncols;
  size_t doff;
};

// pieces [i, j) between their host memory and a pinned block that mirrors
// the device range from byte `start`; the pieces are spread over host
// threads by bytes (each piece is copied by one thread)
void copy_pieces(const std::vector<HostPiece>& pieces, size_t i, size_t j, char* block, size_t start, bool to_block) {
  auto one = [&](size_t q) {
    const HostPiece& p = pieces[q];
    if (to_block) copy_cols(block + (p.doff - start), p.rowb, p.src, p.sld, p.rowb, p.ncols);
    else copy_cols((char*)p.src, p.sld, block + (p.doff - start), p.rowb, p.rowb, p.ncols);
  };
  size_t bytes = 0;
  for (size_t q = i; q < j; ++q) bytes += pieces[q].rowb * (size_t)pieces[q].ncols;
  unsigned nt = std::thread::hardware_concurrency();
  nt = nt < 1 ? 1 : (nt > 8 ? 8 : nt);
  if (bytes < ((size_t)4 << 20) || nt == 1 || j - i < 2) {
    for (size_t q = i; q < j; ++q) one(q);
    return;
  }
  // thread t takes the pieces whose first byte falls in its share
  std::vector<std::thread> pool;
  size_t q = i, acc = 0;
  for (unsigned t = 0; t < nt && q < j; ++t) {
    const size_t limit = bytes * (t + 1) / nt;
    const size_t q0 = q;
    while (q < j && (acc < limit || t + 1 == nt)) {
      acc += pieces[q].rowb * (size_t)pieces[q].ncols;
      ++q;
    }
    if (q > q0)
      pool.emplace_back([&one, q0, q]() {
        for (size_t r = q0; r < q; ++r) one(r);
      });
  }
  for (auto& th : pool) th.join();
}

int pieces_to_dev(vcphip_ctx* c, char* dev, const std::vector<HostPiece>& pieces) {
  size_t total = 0;
  for (const auto& q : pieces) total += q.rowb * (size_t)q.ncols;
  bool pinned = total >= PIN_MIN && pin_ready(c);
  for (const auto& q : pieces) pinned = pinned && q.rowb * (size_t)q.ncols <= PIN_BYTES;
  if (!pinned) {
    for (const auto& q : pieces)
      if (q.ncols > 0 && q.rowb > 0 &&
          hipMemcpy2DAsync(dev + q.doff, q.rowb, q.src, q.sld, q.rowb, q.ncols, hipMemcpyHostToDevice, c->stream) !=
              hipSuccess)
        return VCPHIP_ERR_DEVICE;
    return VCPHIP_OK;
  }
  size_t i = 0;
  int blk = 0;
  while (i < pieces.size()) {
    // the device range [start, end) of the pieces that fit one buffer
    const size_t start = pieces[i].doff;
    size_t j = i, end = start;
    while (j < pieces.size() && pieces[j].doff + pieces[j].rowb * (size_t)pieces[j].ncols - start <= PIN_BYTES) {
      end = pieces[j].doff + pieces[j].rowb * (size_t)pieces[j].ncols;
      ++j;
    }
    const int b = blk & 1;
    if (hipEventSynchronize(c->pin_ev[b]) != hipSuccess) return VCPHIP_ERR_DEVICE;
    copy_pieces(pieces, i, j, (char*)c->pin[b], start, true);
    if (hipMemcpyAsync(dev + start, c->pin[b], end - start, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipEventRecord(c->pin_ev[b], c->stream) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
    i = j;
    ++blk;
  }
  return VCPHIP_OK;
}

// the reverse: the packed device region to the host pieces (dst, stride);
// synchronises the stream before returning
int pieces_to_host(vcphip_ctx* c, const char* dev, const std::vector<HostPiece>& pieces) {
  size_t total = 0;
  for (const auto& q : pieces) total += q.rowb * (size_t)q.ncols;
  bool pinned = total >= PIN_MIN && pin_ready(c);
  for (const auto& q : pieces) pinned = pinned && q.rowb * (size_t)q.ncols <= PIN_BYTES;
  if (!pinned) {
    for (const auto& q : pieces)
      if (q.ncols > 0 && q.rowb > 0 &&
          hipMemcpy2DAsync((void*)q.src, q.sld, dev + q.doff, q.rowb, q.rowb, q.ncols, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess)
        return VCPHIP_ERR_DEVICE;
    return hipStreamSynchronize(c->stream) == hipSuccess ? VCPHIP_OK : VCPHIP_ERR_DEVICE;
  }
  // ranges as above; range r lands in buffer r % 2, two in flight
  std::vector<std::pair<size_t, size_t>> range;   // piece index range [first, last)
  for (size_t i = 0; i < pieces.size();) {
    const size_t start = pieces[i].doff;
    size_t j = i;
    while (j < pieces.size() && pieces[j].doff + pieces[j].rowb * (size_t)pieces[j].ncols - start <= PIN_BYTES) ++j;
    range.emplace_back(i, j);
    i = j;
  }
  auto issue = [&](size_t r) {
    const size_t start = pieces[range[r].first].doff;
    const HostPiece& last = pieces[range[r].second - 1];
    const size_t end = last.doff + last.rowb * (size_t)last.ncols;
    const int b = (int)(r & 1);
    return hipMemcpyAsync(c->pin[b], dev + start, end - start, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
           hipEventRecord(c->pin_ev[b], c->stream) == hipSuccess;
  };
  if ((range.size() > 0 && !issue(0)) || (range.size() > 1 && !issue(1))) {
    hipStreamSynchronize(c->stream);
    return VCPHIP_ERR_DEVICE;
  }
  for (size_t r = 0; r < range.size(); ++r) {
    const int b = (int)(r & 1);
    if (hipEventSynchronize(c->pin_ev[b]) != hipSuccess) return VCPHIP_ERR_DEVICE;
    const size_t start = pieces[range[r].first].doff;
    copy_pieces(pieces, range[r].first, range[r].second, (char*)c->pin[b], start, false);
    if (r + 2 < range.size() && !issue(r + 2)) {
      hipStreamSynchronize(c->stream);
      return VCPHIP_ERR_DEVICE;
    }
  }
  return hipStreamSynchronize(c->stream) == hipSuccess ? VCPHIP_OK : VCPHIP_ERR_DEVICE;
}

int to_dev(vcphip_ctx* c, int slot, const double* h, int ld, int rows, int cols, size_t esize, double** d) {
  void* p;
  const size_t bytes = esize * (size_t)(rows > 0 ? rows : 1) * (size_t)(cols > 0 ? cols : 1);
  int rc = ws_get(c, slot, bytes, &p);
  if (rc) return rc;
  *d = (double*)p;
  if (rows <= 0 || cols <= 0) return VCPHIP_OK;
  const size_t rowb = esize * (size_t)rows;
  if (bytes < PIN_MIN || rowb > PIN_BYTES || !pin_ready(c)) {
    if (hipMemcpy2DAsync(p, rowb, h, esize * ld, rowb, cols, hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
    return VCPHIP_OK;
  }
  const int per = (int)(PIN_BYTES / rowb);
  for (int j0 = 0, i = 0; j0 < cols; j0 += per, ++i) {
    const int nc = std::min(per, cols - j0), b = i & 1;
    if (hipEventSynchronize(c->pin_ev[b]) != hipSuccess) return VCPHIP_ERR_DEVICE;   // its last DMA is done
    copy_cols((char*)c->pin[b], rowb, (const char*)h + esize * (size_t)ld * j0, esize * (size_t)ld, rowb, nc);
    if (hipMemcpyAsync((char*)p + rowb * j0, c->pin[b], rowb * nc, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipEventRecord(c->pin_ev[b], c->stream) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
  }
  return VCPHIP_OK;
}

int to_host(vcphip_ctx* c, const double* d, double* h, int ld, int rows, int cols, size_t esize) {
  if (rows > 0 && cols > 0) {
    const size_t rowb = esize * (size_t)rows, bytes = rowb * (size_t)cols;
    if (bytes < PIN_MIN || rowb > PIN_BYTES || !pin_ready(c)) {
      if (hipMemcpy2DAsync(h, esize * ld, d, rowb, rowb, cols, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return VCPHIP_ERR_DEVICE;
    } else {
      // two blocks in flight: block i lands in buffer i % 2; the host threads
      // copy it out, then block i + 2 is issued into the same buffer (stream
      // order keeps it behind any earlier DMA from that buffer)
      const int per = (int)(PIN_BYTES / rowb), nblk = (cols + per - 1) / per;
      auto issue = [&](int i) {
        const int j0 = i * per, nc = std::min(per, cols - j0), b = i & 1;
        return hipMemcpyAsync(c->pin[b], (const char*)d + rowb * j0, rowb * nc, hipMemcpyDeviceToHost, c->stream) ==
                   hipSuccess &&
               hipEventRecord(c->pin_ev[b], c->stream) == hipSuccess;
      };
      if (!issue(0) || (nblk > 1 && !issue(1))) {
        hipStreamSynchronize(c->stream);
        return VCPHIP_ERR_DEVICE;
      }
      for (int i = 0; i < nblk; ++i) {
        const int j0 = i * per, nc = std::min(per, cols - j0), b = i & 1;
        if (hipEventSynchronize(c->pin_ev[b]) != hipSuccess) return VCPHIP_ERR_DEVICE;
        copy_cols((char*)h + esize * (size_t)ld * j0, esize * (size_t)ld, (const char*)c->pin[b], rowb, rowb, nc);
        if (i + 2 < nblk && !issue(i + 2)) {
          hipStreamSynchronize(c->stream);
          return VCPHIP_ERR_DEVICE;
        }
      }
    }
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return VCPHIP_ERR_DEVICE;
  return VCPHIP_OK;
}

}  // namespace

extern "C" int vcphip_igemm(vcphip_ctx* c, int form, int m, int n, int k, const double* IA, int lda, const double* IB,
                            int ldb, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double *dA, *dB, *dC;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, IA, lda, m, k, 16, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, IB, ldb, k, n, 16, &dB));
  void* p;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)m * n, &p));
  dC = (double*)p;
  VCPHIP_TRY(vcphip_igemm_dev(c, form, m, n, k, dA, m > 0 ? m : 1, dB, k > 0 ? k : 1, dC, m));
  return to_host(c, dC, C, ldc, m, n, 16);
}

extern "C" int vcphip_igemm_ip(vcphip_ctx* c, int m, int n, int k, const double* IA, int lda, const double* B, int ldb,
                               double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double *dA, *dB;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, IA, lda, m, k, 16, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, B, ldb, k, n, 8, &dB));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)m * n, &p));
  VCPHIP_TRY(vcphip_igemm_ip_dev(c, m, n, k, dA, m, dB, k > 0 ? k : 1, (double*)p, m));
  return to_host(c, (double*)p, C, ldc, m, n, 16);
}

extern "C" int vcphip_igemm_pi(vcphip_ctx* c, int m, int n, int k, const double* B, int ldb, const double* IA, int lda,
                               double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || ldb < (m > 1 ? m : 1) || lda < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double *dA, *dB;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, B, ldb, m, k, 8, &dB));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, IA, lda, k, n, 16, &dA));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)m * n, &p));
  VCPHIP_TRY(vcphip_igemm_pi_dev(c, m, n, k, dB, m, dA, k > 0 ? k : 1, (double*)p, m));
  return to_host(c, (double*)p, C, ldc, m, n, 16);
}

extern "C" int vcphip_igemm_tn(vcphip_ctx* c, int k, int n, const double* IA, int lda, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (k < 0 || n < 0 || lda < (k > 1 ? k : 1) || ldc < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double* dA;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, IA, lda, k, n, 16, &dA));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)n * n, &p));
  VCPHIP_TRY(vcphip_igemm_tn_dev(c, k, n, dA, k > 0 ? k : 1, (double*)p, n));
  return to_host(c, (double*)p, C, ldc, n, n, 16);
}

extern "C" int vcphip_vgemm_tn(vcphip_ctx* c, int k, int n, const double* A, int lda, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (k < 0 || n < 0 || lda < (k > 1 ? k : 1) || ldc < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double* dA;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, k, n, 8, &dA));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)n * n, &p));
  VCPHIP_TRY(vcphip_vgemm_tn_dev(c, k, n, dA, k > 0 ? k : 1, (double*)p, n));
  return to_host(c, (double*)p, C, ldc, n, n, 16);
}

extern "C" int vcphip_gemm_rurd(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                                double* Cup, double* Cdn, int ldc, int accumulate) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double *dA, *dB, *dU, *dD;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, m, k, 8, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, B, ldb, k, n, 8, &dB));
  if (accumulate) {
    VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HC, Cup, ldc, m, n, 8, &dU));
    VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HC2, Cdn, ldc, m, n, 8, &dD));
  } else {
    void *p, *q;
    VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 8 * (size_t)m * n, &p));
    VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC2, 8 * (size_t)m * n, &q));
    dU = (double*)p;
    dD = (double*)q;
  }
  VCPHIP_TRY(vcphip_gemm_rurd_dev(c, m, n, k, dA, m, dB, k > 0 ? k : 1, dU, dD, m, accumulate));
  VCPHIP_TRY(to_host(c, dU, Cup, ldc, m, n, 8));
  return to_host(c, dD, Cdn, ldc, m, n, 8);
}

extern "C" int vcphip_ddgemm(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                             double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double *dA, *dB;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, m, k, 16, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, B, ldb, k, n, 16, &dB));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)m * n, &p));
  VCPHIP_TRY(vcphip_ddgemm_dev(c, m, n, k, dA, m, dB, k > 0 ? k : 1, (double*)p, m));
  return to_host(c, (double*)p, C, ldc, m, n, 16);
}

extern "C" int vcphip_dd_residual(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* X,
                                  int ldx, const double* B, int ldb, double* R, int ldr) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (m > 1 ? m : 1) || ldx < (k > 1 ? k : 1) || ldb < (m > 1 ? m : 1) ||
      ldr < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double *dA, *dX, *dB, *dR;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, m, k, 16, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, X, ldx, k, n, 16, &dX));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HR, B, ldb, m, n, 16, &dB));
  void* p;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)m * n, &p));
  dR = (double*)p;
  VCPHIP_TRY(vcphip_dd_residual_dev(c, m, n, k, dA, m > 0 ? m : 1, dX, k > 0 ? k : 1, dB, m, dR, m));
  return to_host(c, dR, R, ldr, m, n, 16);
}

extern "C" int vcphip_ddgemm_ta(vcphip_ctx* c, int m, int n, int k, const double* A, int lda, const double* B,
                                int ldb, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (!dims_ok(m, n, k) || lda < (k > 1 ? k : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return VCPHIP_ERR_DIMENSION;
  if (m == 0 || n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double *dA, *dB;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, k, m, 16, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, B, ldb, k, n, 16, &dB));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)m * n, &p));
  VCPHIP_TRY(vcphip_ddgemm_ta_dev(c, m, n, k, dA, k > 0 ? k : 1, dB, k > 0 ? k : 1, (double*)p, m));
  return to_host(c, (double*)p, C, ldc, m, n, 16);
}

extern "C" int vcphip_ddgemm_tn(vcphip_ctx* c, int k, int n, const double* A, int lda, double* C, int ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (k < 0 || n < 0 || lda < (k > 1 ? k : 1) || ldc < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  double* dA;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, k, n, 16, &dA));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)n * n, &p));
  VCPHIP_TRY(dd_gemm_dev(c, true, n, n, k, dA, k > 0 ? k : 1, dA, k > 0 ? k : 1, (double*)p, n));
  return to_host(c, (double*)p, C, ldc, n, n, 16);
}

extern "C" int vcphip_lss(vcphip_ctx* c, int n, int nrhs, const double* IA, int lda, const double* Ib, int ldb,
                          const double* R, int ldr, double* X, int ldx, double* normub) {
  if (!c || !IA || !Ib || !X) return VCPHIP_ERR_ARGUMENT;
  if (n < 2 || nrhs < 1 || lda < n || ldb < n || (R && ldr < n) || ldx < n) return VCPHIP_ERR_DIMENSION;
  hipSetDevice(c->device);
  double *dA, *dB, *dR = nullptr, *dX;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, IA, lda, n, n, 16, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, Ib, ldb, n, nrhs, 16, &dB));
  if (R) VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HR, R, ldr, n, n, 8, &dR));
  void* px;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC2, 16 * (size_t)n * nrhs, &px));
  dX = (double*)px;
  const int rc = vcphip_lss_dev(c, n, nrhs, dA, n, dB, n, dR, n, dX, n, normub);
  if (rc != VCPHIP_OK && rc != VCPHIP_ERR_VERIFICATION && rc != VCPHIP_ERR_SINGULAR) return rc;
  VCPHIP_TRY(to_host(c, dX, X, ldx, n, nrhs, 16));
  return rc;
}

extern "C" int vcphip_igemm_batched(vcphip_ctx* c, int count, const int* m, const int* n, const int* k,
                                    const double* const* IA, const int* lda, const double* const* IB, const int* ldb,
                                    double* const* C, const int* ldc) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count < 0 || (count > 0 && (!m || !n || !k || !IA || !lda || !IB || !ldb || !C || !ldc)))
    return VCPHIP_ERR_ARGUMENT;
  if (count == 0) return VCPHIP_OK;
  for (int p = 0; p < count; ++p) {
    if (!dims_ok(m[p], n[p], k[p]) || lda[p] < (m[p] > 1 ? m[p] : 1) || ldb[p] < (k[p] > 1 ? k[p] : 1) ||
        ldc[p] < (m[p] > 1 ? m[p] : 1))
      return VCPHIP_ERR_DIMENSION;
    if (m[p] > 0 && n[p] > 0 && (!C[p] || (k[p] > 0 && (!IA[p] || !IB[p])))) return VCPHIP_ERR_ARGUMENT;
  }
  hipSetDevice(c->device);
  // inputs packed into one staging buffer (A, B per problem), outputs into another
  auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
  std::vector<size_t> oi((size_t)count), ox((size_t)count);
  size_t in_bytes = 0, out_bytes = 0;
  for (int p = 0; p < count; ++p) {
    oi[p] = in_bytes;
    in_bytes += al(16 * (size_t)m[p] * k[p]) + al(16 * (size_t)k[p] * n[p]);
    ox[p] = out_bytes;
    out_bytes += al(16 * (size_t)m[p] * n[p]);
  }
  void *pin, *pout;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HA, in_bytes, &pin));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC2, out_bytes, &pout));
  std::vector<const double*> dA((size_t)count), dB((size_t)count);
  std::vector<double*> dC((size_t)count);
  std::vector<int> dlda((size_t)count), dldb((size_t)count), dldc((size_t)count);
  std::vector<HostPiece> in, out;
  for (int p = 0; p < count; ++p) {
    char* q = (char*)pin + oi[p];
    dA[p] = (const double*)q;
    dB[p] = (const double*)(q + al(16 * (size_t)m[p] * k[p]));
    dC[p] = (double*)((char*)pout + ox[p]);
    dlda[p] = m[p] > 1 ? m[p] : 1;
    dldb[p] = k[p] > 1 ? k[p] : 1;
    dldc[p] = m[p] > 1 ? m[p] : 1;
    if (m[p] > 0 && k[p] > 0)
      in.push_back(HostPiece{(const char*)IA[p], 16 * (size_t)lda[p], 16 * (size_t)m[p], k[p], oi[p]});
    if (k[p] > 0 && n[p] > 0)
      in.push_back(HostPiece{(const char*)IB[p], 16 * (size_t)ldb[p], 16 * (size_t)k[p], n[p],
                             oi[p] + al(16 * (size_t)m[p] * k[p])});
    if (m[p] > 0 && n[p] > 0)
      out.push_back(HostPiece{(const char*)C[p], 16 * (size_t)ldc[p], 16 * (size_t)m[p], n[p], ox[p]});
  }
  VCPHIP_TRY(pieces_to_dev(c, (char*)pin, in));
  VCPHIP_TRY(vcphip_igemm_batched_dev(c, count, m, n, k, dA.data(), dlda.data(), dB.data(), dldb.data(), dC.data(),
                                      dldc.data()));
  return pieces_to_host(c, (const char*)pout, out);
}

extern "C" int vcphip_dd_solve(vcphip_ctx* c, int n, int nrhs, const double* A, int lda, const double* b, int ldb,
                               double* x, int ldx, int itmax, int* iters) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || nrhs < 0 || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1) || ldx < (n > 1 ? n : 1))
    return VCPHIP_ERR_DIMENSION;
  if (iters) *iters = 0;
  if (n == 0 || nrhs == 0) return VCPHIP_OK;
  if (!A || !b || !x || itmax < 0) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  double *dA, *dB;
  void* px;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, n, n, 16, &dA));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, b, ldb, n, nrhs, 16, &dB));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC2, 16 * (size_t)n * nrhs, &px));
  VCPHIP_TRY(vcphip_dd_solve_dev(c, n, nrhs, dA, n, dB, n, (double*)px, n, itmax, iters));
  return to_host(c, (double*)px, x, ldx, n, nrhs, 16);
}

extern "C" int vcphip_dd_cholesky(vcphip_ctx* c, int n, const double* A, int lda, double* U, int ldu, int itep) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || lda < (n > 1 ? n : 1) || ldu < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  if (!A || !U || itep < 0) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  double* dA;
  void* pu;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, n, n, 16, &dA));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC2, 16 * (size_t)n * n, &pu));
  VCPHIP_TRY(vcphip_dd_cholesky_dev(c, n, dA, n, (double*)pu, n, itep));
  return to_host(c, (double*)pu, U, ldu, n, n, 16);
}

namespace {
int dd_eig_host(vcphip_ctx* c, int n, const double* A, int lda, const double* B, int ldb, double* lam, double* X,
                int ldx, int itep) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || lda < (n > 1 ? n : 1) || ldx < (n > 1 ? n : 1) || (B && ldb < (n > 1 ? n : 1)))
    return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  if (!A || !lam || !X || itep < 0) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  double *dA, *dB = nullptr;
  void *px, *pl;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, A, lda, n, n, 16, &dA));
  if (B) VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, B, ldb, n, n, 16, &dB));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC2, 16 * (size_t)n * n, &px));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HR, 16 * (size_t)n, &pl));
  VCPHIP_TRY(dd_eig_dev(c, n, dA, n, dB, n, (double*)pl, (double*)px, n, itep));
  VCPHIP_TRY(to_host(c, (double*)pl, lam, n, n, 1, 16));
  return to_host(c, (double*)px, X, ldx, n, n, 16);
}
}  // namespace

extern "C" int vcphip_dd_eigsym(vcphip_ctx* c, int n, const double* A, int lda, double* lam, double* X, int ldx,
                                int itep) {
  return dd_eig_host(c, n, A, lda, nullptr, 0, lam, X, ldx, itep);
}

extern "C" int vcphip_dd_eigsymge(vcphip_ctx* c, int n, const double* A, int lda, const double* B, int ldb,
                                  double* lam, double* X, int ldx, int itep) {
  if (c && n > 0 && !B) return VCPHIP_ERR_ARGUMENT;
  return dd_eig_host(c, n, A, lda, B, ldb, lam, X, ldx, itep);
}

extern "C" int vcphip_inv_mid(vcphip_ctx* c, int n, const double* IA, int lda, double* R, int ldr) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (n < 0 || lda < (n > 1 ? n : 1) || ldr < (n > 1 ? n : 1)) return VCPHIP_ERR_DIMENSION;
  if (n == 0) return VCPHIP_OK;
  if (!IA || !R) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  double* dA;
  void* p;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, IA, lda, n, n, 16, &dA));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 8 * (size_t)n * n, &p));
  const int rc = vcphip_inv_mid_dev(c, n, dA, n, (double*)p, n);
  if (rc) return rc;
  return to_host(c, (double*)p, R, ldr, n, n, 8);
}

extern "C" int vcphip_lss_batched(vcphip_ctx* c, int count, const int* n, const int* nrhs, const double* const* IA,
                                  const int* lda, const double* const* Ib, const int* ldb, const double* const* R,
                                  const int* ldr, double* const* X, const int* ldx, double* normub) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count < 0 || (count > 0 && (!n || !nrhs || !IA || !lda || !Ib || !ldb || !X || !ldx || !normub)))
    return VCPHIP_ERR_ARGUMENT;
  if (count == 0) return VCPHIP_OK;
  auto has_r = [&](int p) { return R && R[p]; };
  for (int p = 0; p < count; ++p) {
    if (!IA[p] || !Ib[p] || !X[p]) return VCPHIP_ERR_ARGUMENT;
    if (n[p] < 2 || nrhs[p] < 1 || lda[p] < n[p] || ldb[p] < n[p] || (has_r(p) && (!ldr || ldr[p] < n[p])) ||
        ldx[p] < n[p])
      return VCPHIP_ERR_DIMENSION;
  }
  hipSetDevice(c->device);
  // inputs packed into one staging buffer (A, b, R per problem), outputs into another
  std::vector<size_t> oi((size_t)count), ox((size_t)count);
  size_t in_bytes = 0, out_bytes = 0;
  for (int p = 0; p < count; ++p) {
    const size_t np = (size_t)n[p], nr = np * nrhs[p];
    oi[p] = in_bytes;
    in_bytes += ((16 * np * np + 16 * nr + (has_r(p) ? 8 * np * np : 0)) + 255) & ~(size_t)255;
    ox[p] = out_bytes;
    out_bytes += (16 * nr + 255) & ~(size_t)255;
  }
  void *pin, *pout;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HA, in_bytes, &pin));
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC2, out_bytes, &pout));
  std::vector<const double*> dA((size_t)count), dB((size_t)count), dR((size_t)count);
  std::vector<double*> dX((size_t)count);
  std::vector<int> pn((size_t)count);
  std::vector<HostPiece> in, out;
  for (int p = 0; p < count; ++p) {
    const size_t np = (size_t)n[p], nr = np * nrhs[p];
    char* q = (char*)pin + oi[p];
    dA[p] = (const double*)q;
    dB[p] = (const double*)(q + 16 * np * np);
    dR[p] = has_r(p) ? (const double*)(q + 16 * np * np + 16 * nr) : nullptr;
    dX[p] = (double*)((char*)pout + ox[p]);
    pn[p] = n[p];
    in.push_back(HostPiece{(const char*)IA[p], 16 * (size_t)lda[p], 16 * np, n[p], oi[p]});
    in.push_back(HostPiece{(const char*)Ib[p], 16 * (size_t)ldb[p], 16 * np, nrhs[p], oi[p] + 16 * np * np});
    if (has_r(p))
      in.push_back(HostPiece{(const char*)R[p], 8 * (size_t)ldr[p], 8 * np, n[p], oi[p] + 16 * np * np + 16 * nr});
    out.push_back(HostPiece{(const char*)X[p], 16 * (size_t)ldx[p], 16 * np, nrhs[p], ox[p]});
  }
  VCPHIP_TRY(pieces_to_dev(c, (char*)pin, in));
  const int rc = vcphip_lss_batched_dev(c, count, n, nrhs, dA.data(), pn.data(), dB.data(), pn.data(), dR.data(),
                                        pn.data(), dX.data(), pn.data(), normub);
  if (rc != VCPHIP_OK && rc != VCPHIP_ERR_VERIFICATION && rc != VCPHIP_ERR_SINGULAR) return rc;
  VCPHIP_TRY(pieces_to_host(c, (const char*)pout, out));
  return rc;
}

extern "C" int vcphip_uhphiphi(vcphip_ctx* c, int m, int npts, const double* phi, int ldp, const double* w,
                               const double* u, double* A, int lda) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (m < 0 || npts < 0 || ldp < (m > 1 ? m : 1) || lda < (m > 1 ? m : 1)) return VCPHIP_ERR_DIMENSION;
  if (m == 0) return VCPHIP_OK;
  if (!phi || !w || !u || !A) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  double *dP, *dW, *dU;
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HA, phi, ldp, m, npts, 16, &dP));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HB, w, 1, 1, npts, 16, &dW));
  VCPHIP_TRY(to_dev(c, vcphip_ctx::WS_HR, u, 1, 1, npts, 16, &dU));
  void* p;
  VCPHIP_TRY(ws_get(c, vcphip_ctx::WS_HC, 16 * (size_t)m * m, &p));
  VCPHIP_TRY(vcphip_uhphiphi_dev(c, m, npts, dP, m, dW, dU, (double*)p, m));
  return to_host(c, (double*)p, A, lda, m, m, 16);
}

extern "C" int vcphip_residual_batched(vcphip_ctx* c, int count, const int* n, const double* const* R, const int* ldr,
                                       const double* const* IA, const int* lda, double* const* C, const int* ldc,
                                       double* normub) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count < 0 || (count > 0 && (!n || !R || !ldr || !IA || !lda || !C || !ldc || !normub))) return VCPHIP_ERR_ARGUMENT;
  if (count == 0) return VCPHIP_OK;
  hipSetDevice(c->device);
  // one staging buffer: R_i (8 n^2), A_i (16 n^2), C_i (16 n^2) packed, then normub
  std::vector<size_t> offR(count), offA(count), offC(count);
  size_t tot = 0;
  for (int i = 0; i < count; ++i) {
    if (n[i] < 0 || ldr[i] < (n[i] > 1 ? n[i] : 1) || lda[i] < (n[i] > 1 ? n[i] : 1) || ldc[i] < (n[i] > 1 ? n[i] : 1))
      return VCPHIP_ERR_DIMENSION;
    const size_t nn = (size_t)n[i] * n[i];
    offR[i] = tot; tot += ((8 * nn + 255) & ~(size_t)255);
    offA[i] = tot; tot += ((16 * nn + 255) & ~(size_t)255);
    offC[i] = tot; tot += ((16 * nn + 255) & ~(size_t)255);
  }
  const size_t offN = tot;
  tot += sizeof(double) * count + 256;
  void* base;
  int rc = ws_get(c, vcphip_ctx::WS_HA, tot, &base);
  if (rc) return rc;
  char* b8 = (char*)base;
  std::vector<const double*> dR(count), dA(count);
  std::vector<double*> dC(count);
  std::vector<int> ld(count);
  for (int i = 0; i < count; ++i) {
    const int ni = n[i];
    dR[i] = (const double*)(b8 + offR[i]);
    dA[i] = (const double*)(b8 + offA[i]);
    dC[i] = (double*)(b8 + offC[i]);
    ld[i] = ni > 1 ? ni : 1;
    if (ni > 0) {
      if (hipMemcpy2DAsync(b8 + offR[i], 8 * ni, R[i], 8 * (size_t)ldr[i], 8 * ni, ni, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
          hipMemcpy2DAsync(b8 + offA[i], 16 * ni, IA[i], 16 * (size_t)lda[i], 16 * ni, ni, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return VCPHIP_ERR_DEVICE;
    }
  }
  rc = vcphip_residual_batched_dev(c, count, n, dR.data(), ld.data(), dA.data(), ld.data(), dC.data(), ld.data(),
                                   (double*)(b8 + offN));
  if (rc) return rc;
  for (int i = 0; i < count; ++i)
    if (n[i] > 0 && hipMemcpy2DAsync(C[i], 16 * (size_t)ldc[i], dC[i], 16 * n[i], 16 * n[i], n[i], hipMemcpyDeviceToHost, c->stream) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
  if (hipMemcpyAsync(normub, b8 + offN, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    return VCPHIP_ERR_DEVICE;
  return hipStreamSynchronize(c->stream) == hipSuccess ? VCPHIP_OK : VCPHIP_ERR_DEVICE;
}

// ===========================================================================
// autotuner
// ===========================================================================
namespace {

constexpr int TUNE_VARIANTS[] = {6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18};

// time the int8 MFMA stage of one Ozaki product shape with every tile variant
// (operand planes of arbitrary content: the integer MFMA time does not depend
// on the values); returns the fastest
int tune_gemm_stage(vcphip_ctx* c, bool dd, int m, int n, int k, int* best_variant, double* best_ms) {
  if (m <= 0 || n <= 0 || k <= 0) return VCPHIP_ERR_DIMENSION;
  if (k > VCPHIP_OZ_MAXK) return VCPHIP_ERR_UNSUPPORTED;
  const vcphip_ctx::Tune* tuned = tune_find(c, dd ? TUNE_DDGEMM : TUNE_IGEMM, m, n, k);
  int nmod;
  VcphipOzParams P;
  int rc;
  if (dd) {
    nmod = c->dd_nmod ? (int)c->dd_nmod : default_dd_nmod(k);
    rc = dd_make_params(k, nmod, &P);
  } else {
    nmod = c->oz_nmod ? (int)c->oz_nmod : (tuned && tuned->nmod ? tuned->nmod : default_nmod(k));
    rc = oz_make_params(k, nmod, &P);
  }
  if (rc) return rc;
  if (!dd) oz_set_radius(&P, oz_radius_scheme(c, tuned, k));
  const int nrad = dd ? 0 : P.rnz;
  const int mpad = (m + VCPHIP_OZ_PADMN - 1) / VCPHIP_OZ_PADMN * VCPHIP_OZ_PADMN;
  const int npad = (n + VCPHIP_OZ_PADMN - 1) / VCPHIP_OZ_PADMN * VCPHIP_OZ_PADMN;
  const int kpad = (k + VCPHIP_OZ_PADK - 1) / VCPHIP_OZ_PADK * VCPHIP_OZ_PADK;
  void *pa, *pb, *pu, *pz = nullptr;
  const size_t nplane_a = (size_t)nmod + (dd ? 0 : VCPHIP_OZ_NDIGIT_A);
  const size_t nplane_b = (size_t)nmod + (dd ? 0 : VCPHIP_OZ_NDIGIT_B);
  if ((rc = ws_get(c, vcphip_ctx::WS_PA, nplane_a * mpad * kpad, &pa))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_PB, nplane_b * npad * kpad, &pb))) return rc;
  if ((rc = ws_get(c, vcphip_ctx::WS_U, (size_t)nmod * mpad * npad, &pu))) return rc;
  if (!dd && (rc = ws_get(c, vcphip_ctx::WS_Z, (size_t)nrad * sizeof(int32_t) * (size_t)mpad * npad, &pz)))
    return rc;
  // operand planes of random bytes: timings on the zeros of fresh memory do
  // not carry over to real operands
  if ((rc = check_launch(vcphip_fill_random(pa, nplane_a * mpad * kpad, 0x5EED0001u, c->stream)))) return rc;
  if ((rc = check_launch(vcphip_fill_random(pb, nplane_b * npad * kpad, 0x5EED0002u, c->stream)))) return rc;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return VCPHIP_ERR_DEVICE;
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return VCPHIP_ERR_DEVICE;
  }
  const double flops = 2.0 * mpad * npad * (double)kpad * (double)(nmod + (dd ? 0 : oz_radius_products(P.rscheme)));
  const int reps = flops > 1e13 ? 2 : (flops > 1e11 ? 4 : 10);
  int best = 0;
  double best_t = 0.0;
  rc = VCPHIP_OK;
  for (int v : TUNE_VARIANTS) {
    auto launch = [&]() {
      return vcphip_oz_gemm(mpad, npad, kpad, nmod, nrad, (const int8_t*)pa, (const int8_t*)pb, P, (uint8_t*)pu,
                            (int32_t*)pz, v, c->stream);
    };
    if (launch() != 0) {   // variant not applicable to this shape
      hipGetLastError();
      continue;
    }
    // the least of three timings: variants differ by a few per cent, about
    // as much as one timing varies
    double t = 0.0;
    for (int round = 0; round < 3 && rc == VCPHIP_OK; ++round) {
      hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps; ++r) launch();
      hipEventRecord(e1, c->stream);
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = VCPHIP_ERR_DEVICE;
        break;
      }
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (round == 0 || ms / reps < t) t = ms / reps;
    }
    if (rc) break;
    if (!best || t < best_t) {
      best = v;
      best_t = t;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  if (!best) return VCPHIP_ERR_UNSUPPORTED;
  tune_put(c, dd ? TUNE_DDGEMM : TUNE_IGEMM, size_class(m), size_class(n), size_class(k), best, 0);
  if (best_variant) *best_variant = best;
  if (best_ms) *best_ms = best_t;
  return VCPHIP_OK;
}

}  // namespace

extern "C" int vcphip_tune_igemm_tile(vcphip_ctx* c, int m, int n, int k, int* variant, double* ms) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  return tune_gemm_stage(c, false, m, n, k, variant, ms);
}

extern "C" int vcphip_tune_ddgemm_tile(vcphip_ctx* c, int m, int n, int k, int* variant, double* ms) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  hipSetDevice(c->device);
  return tune_gemm_stage(c, true, m, n, k, variant, ms);
}

extern "C" int vcphip_tune_igemm_moduli_dev(vcphip_ctx* c, int m, int n, int k, const double* IA, int lda,
                                            const double* IB, int ldb, double max_ratio, int sample_rows, int* nmod,
                                            int* radius, double* ratio) {
  if (!c || !IA || !IB || !(max_ratio > 0.0)) return VCPHIP_ERR_ARGUMENT;
  if (m <= 0 || n <= 0 || k <= 0 || lda < m || ldb < k) return VCPHIP_ERR_DIMENSION;
  if (k > VCPHIP_OZ_MAXK) return VCPHIP_ERR_UNSUPPORTED;
  hipSetDevice(c->device);
  // the sampled maximum of the ratio underestimates the full one, and the
  // more so the smaller the sample: by default max(1024, m/8) rows (at
  // n = 8192 on exponents over 41 binades, 256 rows accepted a choice whose
  // full-matrix maximum was 2.16 against a target of 2; 1024 rows did not)
  const int s = std::max(1, std::min(m, sample_rows > 0 ? sample_rows : std::max(1024, m / 8)));
  // sample rows, the plain and the Ozaki products of the sample, block maxima
  const size_t bs = sizeof(double) * 2 * (size_t)s * k, bc = sizeof(double) * 2 * (size_t)s * n;
  const int nblocks = 1024;
  const size_t o_p = (bs + 255) & ~(size_t)255, o_o = o_p + ((bc + 255) & ~(size_t)255);
  const size_t o_m = o_o + ((bc + 255) & ~(size_t)255);
  void* base;
  int rc = ws_get(c, vcphip_ctx::WS_HC2, o_m + sizeof(double) * nblocks, &base);
  if (rc) return rc;
  char* b8 = (char*)base;
  double* As = (double*)b8;
  double* Cp = (double*)(b8 + o_p);
  double* Co = (double*)(b8 + o_o);
  double* part = (double*)(b8 + o_m);
  if ((rc = check_launch(vcphip_gather_rows(IA, lda, m, k, s, As, c->stream)))) return rc;
  if ((rc = check_launch(vcphip_launch_drgemm(VCPHIP_DR_II, s, n, k, As, s, IB, ldb, Cp, s, nullptr, 0, c->stream))))
    return rc;
  const long long saved = c->oz_nmod, saved_rad = c->oz_radius;
  const long long saved_timing = c->timing;
  c->timing = 0;
  std::vector<double> hpart(nblocks);
  // candidates in order of their plane products (N + 6 folded, N + 8 full);
  // the first cost with a candidate within max_ratio decides, the smaller
  // sampled ratio winning a tie of cost
  int chosen = 0, chosen_rad = 0, last = 0, last_rad = 0;
  double chosen_ratio = 0.0, last_ratio = 0.0;
  for (int cost = 12 + 2; cost <= VCPHIP_OZ_MAXMOD_IV + 8 && !chosen && !rc; ++cost) {
    for (const int rad : {VCPHIP_OZ_RADIUS_CEIL, VCPHIP_OZ_RADIUS_HALF, VCPHIP_OZ_RADIUS_FOLDED,
                          VCPHIP_OZ_RADIUS_FULL}) {
      const int t = cost - oz_radius_products(rad);
      VcphipOzParams P;
      if (t < 12 || t > VCPHIP_OZ_MAXMOD_IV || oz_make_params(k, t, &P)) continue;   // too few moduli for this k
      c->oz_nmod = t;
      c->oz_radius = rad;
      rc = oz_igemm_dev(c, s, n, k, As, s, IB, ldb, Co, s);
      if (rc) break;
      if ((rc = check_launch(vcphip_ratio_max(Co, Cp, (long)s * n, part, nblocks, c->stream)))) break;
      if (hipMemcpyAsync(hpart.data(), part, sizeof(double) * nblocks, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = VCPHIP_ERR_DEVICE;
        break;
      }
      double mx = 0.0;
      for (double v : hpart) mx = std::max(mx, v);
      last = t;
      last_rad = rad;
      last_ratio = mx;
      if (mx <= max_ratio && (!chosen || mx < chosen_ratio)) {
        chosen = t;
        chosen_rad = rad;
        chosen_ratio = mx;
      }
    }
  }
  c->oz_nmod = saved;
  c->oz_radius = saved_rad;
  c->timing = saved_timing;
  if (rc) return rc;
  if (!chosen) {   // none within max_ratio: the most moduli, full scheme
    chosen = last;
    chosen_rad = last_rad;
    chosen_ratio = last_ratio;
  }
  if (!chosen) return VCPHIP_ERR_UNSUPPORTED;
  tune_put(c, TUNE_IGEMM, size_class(m), size_class(n), size_class(k), 0, chosen, chosen_rad);
  if (nmod) *nmod = chosen;
  if (radius) *radius = chosen_rad;
  if (ratio) *ratio = chosen_ratio;
  return VCPHIP_OK;
}

namespace {

// padded fraction of a tile set's multiply-adds
double padded_fraction(const TileSet& s) { return s.tile_work > 0.0 ? 1.0 - s.useful / s.tile_work : 0.0; }

}  // namespace

extern "C" int vcphip_tune_residual_dev(vcphip_ctx* c, int count, const int* n, const double* const* R,
                                        const int* ldr, const double* const* IA, const int* lda, double* const* C,
                                        const int* ldc, double* normub_dev, int* per_cu, double* ms, double* padded) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count <= 0) return VCPHIP_ERR_ARGUMENT;
  // build (or reuse) the descriptors of this problem set
  int rc = vcphip_residual_batched_dev(c, count, n, R, ldr, IA, lda, C, ldc, normub_dev);
  if (rc) return rc;
  vcphip_ctx::Resid& rs = c->resid;
  if (rs.probs.empty() || !rs.dp) return VCPHIP_ERR_UNSUPPORTED;
  const std::vector<VcphipResidProblem> probs = rs.probs;
  const VcphipResidProblem* dp = rs.dp;
  // the tile launches of the set under the current shape table and a grid of
  // `per` workgroups per CU, from a tile array of the tuner's own
  int per = c->resid_per_cu;
  auto time_set = [&](double* t) -> int {
    TileSet ts = resid_tiles(c, probs);
    const int n64 = (int)ts.t64.size(), n32 = (int)ts.t32.size();
    const int g64 = batch_grid(c, ts.t64.size(), per), g32 = batch_grid(c, ts.t32.size(), per);
    const std::vector<VcphipResidTile> tiles = tile_array(ts);
    void* base;
    int wrc = ws_get(c, vcphip_ctx::WS_TT, sizeof(VcphipResidTile) * tiles.size(), &base);
    if (wrc) return wrc;
    const VcphipResidTile* dt = (const VcphipResidTile*)base;
    if (!tiles.empty() && hipMemcpyAsync(base, tiles.data(), sizeof(VcphipResidTile) * tiles.size(),
                                         hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
    // time_runs synchronises the stream before `tiles` goes out of scope
    unsigned* queue = batch_queue(c);
    if (!queue) return VCPHIP_ERR_OOM;
    return time_runs(c, [&]() { return vcphip_launch_resid_tiles(n64, n32, dt, dp, g64, g32, queue, c->stream); }, 5,
                     t);
  };
  std::vector<int> rows((size_t)count);
  std::vector<double> work((size_t)count);
  for (int i = 0; i < count; ++i) {
    rows[i] = n[i];
    work[i] = (double)n[i] * n[i] * n[i];
  }
  double best_t = 0.0;
  rc = tune_shapes(c, TS_RESID, rows, work, time_set, &best_t);
  // then the persistent grid with the chosen shapes: 1-3 workgroups per CU,
  // as many as are co-resident at the tile kernels' three waves per SIMD (a
  // fourth only queues behind them; alone that measured within noise of
  // three, and inside the batched solve it took CUs from the right-hand-side
  // kernels running beside it: 6.3 against 5.55 ms)
  int best = c->resid_per_cu;
  for (per = 1; per <= 3 && rc == VCPHIP_OK; ++per) {
    if (per == c->resid_per_cu) continue;
    double t = 0.0;
    rc = time_set(&t);
    if (rc == VCPHIP_OK && t < 0.99 * best_t) {
      best = per;
      best_t = t;
    }
  }
  if (rc) return rc;
  c->resid_per_cu = best;
  // drop the cached graph: the next call captures the launches with the tuned choices
  if (rs.exec) {
    hipGraphExecDestroy(rs.exec);
    rs.exec = nullptr;
  }
  if (rs.graph) {
    hipGraphDestroy(rs.graph);
    rs.graph = nullptr;
  }
  rs.probs.clear();
  if (per_cu) *per_cu = best;
  if (ms) *ms = best_t;
  if (padded) *padded = padded_fraction(resid_tiles(c, probs));
  return VCPHIP_OK;
}

extern "C" int vcphip_tune_igemm_batched_dev(vcphip_ctx* c, int count, const int* m, const int* n, const int* k,
                                             const double* const* IA, const int* lda, const double* const* IB,
                                             const int* ldb, double* const* C, const int* ldc, double* ms,
                                             double* padded) {
  if (!c) return VCPHIP_ERR_ARGUMENT;
  if (count <= 0) return VCPHIP_ERR_ARGUMENT;
  int rc = vcphip_igemm_batched_dev(c, count, m, n, k, IA, lda, IB, ldb, C, ldc);
  if (rc) return rc;
  vcphip_ctx::Igb& g = c->igb;
  if (g.probs.empty() || !g.dp) return VCPHIP_ERR_UNSUPPORTED;
  const std::vector<VcphipIgemmProblem> probs = g.probs;
  const VcphipIgemmProblem* dp = g.dp;
  auto time_set = [&](double* t) -> int {
    TileSet ts = igemm_tiles(c, probs);
    const int n64 = (int)ts.t64.size(), n32 = (int)ts.t32.size();
    const int g64 = batch_grid(c, ts.t64.size(), 2), g32 = batch_grid(c, ts.t32.size(), 2);
    const std::vector<VcphipResidTile> tiles = tile_array(ts);
    void* base;
    int wrc = ws_get(c, vcphip_ctx::WS_TT, sizeof(VcphipResidTile) * tiles.size(), &base);
    if (wrc) return wrc;
    const VcphipResidTile* dt = (const VcphipResidTile*)base;
    if (!tiles.empty() && hipMemcpyAsync(base, tiles.data(), sizeof(VcphipResidTile) * tiles.size(),
                                         hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return VCPHIP_ERR_DEVICE;
    unsigned* queue = batch_queue(c);
    if (!queue) return VCPHIP_ERR_OOM;
    return time_runs(c, [&]() { return vcphip_launch_igemm_tiles(n64, n32, dt, dp, g64, g32, queue, c->stream); }, 5,
                     t);
  };
  std::vector<int> rows((size_t)count);
  std::vector<double> work((size_t)count);
  for (int i = 0; i < count; ++i) {
    rows[i] = m[i];
    work[i] = (double)m[i] * n[i] * k[i];
  }
  double best_t = 0.0;
  rc = tune_shapes(c, TS_IGEMM, rows, work, time_set, &best_t);
  if (rc) return rc;
  if (g.exec) {
    hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
  }
  if (g.graph) {
    hipGraphDestroy(g.graph);
    g.graph = nullptr;
  }
  g.probs.clear();
  if (ms) *ms = best_t;
  if (padded) *padded = padded_fraction(igemm_tiles(c, probs));
  return VCPHIP_OK;
}

extern "C" int vcphip_tune_save(vcphip_ctx* c, const char* path) {
  if (!c || !path) return VCPHIP_ERR_ARGUMENT;
  FILE* f = fopen(path, "w");
  if (!f) return VCPHIP_ERR_ARGUMENT;
  fprintf(f, "vcphip-tune 2\n");
  for (const auto& t : c->tune)
    fprintf(f, "%s %d %d %d %d %d %d\n", t.kind == TUNE_IGEMM ? "igemm" : "ddgemm", t.lm, t.ln, t.lk, t.variant,
            t.nmod, t.radius);
  fprintf(f, "residual %d\n", c->resid_per_cu);
  for (int kind = 0; kind < 2; ++kind)
    for (int l = 0; l < 32; ++l)
      if (c->tshape[kind][l]) fprintf(f, "tiles %s %d %d\n", kind == TS_RESID ? "residual" : "igemm", l, c->tshape[kind][l]);
  const bool ok = fclose(f) == 0;
  return ok ? VCPHIP_OK : VCPHIP_ERR_ARGUMENT;
}

extern "C" int vcphip_tune_load(vcphip_ctx* c, const char* path) {
  if (!c || !path) return VCPHIP_ERR_ARGUMENT;
  FILE* f = fopen(path, "r");
  if (!f) return VCPHIP_ERR_ARGUMENT;
  char line[256];
  int version = 0;
  if (!fgets(line, sizeof line, f) || sscanf(line, "vcphip-tune %d", &version) != 1 || version < 1 || version > 2) {
    fclose(f);
    return VCPHIP_ERR_ARGUMENT;
  }
  std::vector<vcphip_ctx::Tune> table;
  struct Shape {
    int kind, cls, mode;
  };
  std::vector<Shape> shapes;
  int per_cu = c->resid_per_cu;
  int rc = VCPHIP_OK;
  while (fgets(line, sizeof line, f)) {
    char kind[32];
    int lm, ln, lk, v, nm, p, rad = 0;
    const int nf = sscanf(line, "%31s %d %d %d %d %d %d", kind, &lm, &ln, &lk, &v, &nm, &rad);
    if (nf == 6 + (version >= 2 ? 1 : 0)) {   // version 1 lines carry no radius scheme
      const bool iv = strcmp(kind, "igemm") == 0, dd = strcmp(kind, "ddgemm") == 0;
      const bool ok = (iv || dd) && lm >= 0 && lm <= 31 && ln >= 0 && ln <= 31 && lk >= 0 && lk <= 31 &&
                      (v == 0 || (v >= 6 && v <= 18)) && (nm == 0 || (iv && nm >= 8 && nm <= VCPHIP_OZ_MAXMOD_IV) ||
                                  (dd && nm >= 24 && nm <= VCPHIP_OZ_MAXMOD)) &&
                      (rad == 0 || (iv && oz_radius_valid(rad)));
      if (!ok) {
        rc = VCPHIP_ERR_ARGUMENT;
        break;
      }
      table.push_back(vcphip_ctx::Tune{iv ? TUNE_IGEMM : TUNE_DDGEMM, lm, ln, lk, v, nm, rad});
    } else if (sscanf(line, "residual %d", &p) == 1 && p >= 1 && p <= 4) {
      per_cu = p;
    } else if (sscanf(line, "tiles %31s %d %d", kind, &lm, &v) == 3) {
      const int tk = strcmp(kind, "residual") == 0 ? TS_RESID : strcmp(kind, "igemm") == 0 ? TS_IGEMM : -1;
      if (tk < 0 || lm < 0 || lm > 31 || v < VCPHIP_BATCH_TILES_64 || v > VCPHIP_BATCH_TILES_32) {
        rc = VCPHIP_ERR_ARGUMENT;
        break;
      }
      shapes.push_back({tk, lm, v});
    } else if (line[0] != '\n' && line[0] != '#') {
      rc = VCPHIP_ERR_ARGUMENT;
      break;
    }
  }
  fclose(f);
  if (rc) return rc;
  for (const auto& t : table) tune_put(c, t.kind, t.lm, t.ln, t.lk, t.variant, t.nmod, t.radius);
  c->resid_per_cu = per_cu;
  for (const Shape& s : shapes) c->tshape[s.kind][s.cls] = (unsigned char)s.mode;
  if (!shapes.empty()) {   // cached tile lists are rebuilt on the next call
    c->resid.probs.clear();
    c->igb.probs.clear();
  }
  return VCPHIP_OK;
}
