// Helper kernels of the autotuner: an evenly spaced row sample of a
// column-major interval matrix, and the blockwise maximum of the radius ratio
// of two interval results.
#include <algorithm>

#include "vcphip_device.h"
#include "vcphip_kernels.h"

namespace {

// out (s x k, column-major pairs) = rows floor(r m / s), r < s, of A (m x k)
__global__ __launch_bounds__(256) void k_gather_rows(const double2* __restrict__ A, long lda, int m, int k, int s,
                                                     double2* __restrict__ out) {
  const long total = (long)s * k;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
    const int r = (int)(t % s);
    const long kk = t / s;
    const long i = (long)r * m / s;
    out[t] = A[i + lda * kk];
  }
}

// part[block] = max over its entries of w(X) / w(Y), w = sup - inf; an entry
// with w(Y) = 0 counts w(X) / DBL_MIN (0 when both are points)
__global__ __launch_bounds__(256) void k_ratio_max(const double2* __restrict__ X, const double2* __restrict__ Y,
                                                   long total, double* __restrict__ part) {
  __shared__ double red[256];
  double mx = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
    const double2 x = X[t], y = Y[t];
    const double wx = x.y - x.x, wy = y.y - y.x;
    const double r = wx / (wy > 2.2250738585072014e-308 ? wy : 2.2250738585072014e-308);
    mx = r > mx ? r : mx;
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// bytes of a hash of (seed, index): stand-in operand planes for timing (the
// integer MFMA draws more power, and may clock lower, on random data than on
// the zeros of fresh memory)
__global__ __launch_bounds__(256) void k_fill_random(uint4* __restrict__ p, long n16, uint32_t seed) {
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n16; t += (long)gridDim.x * 256) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t x = (uint32_t)(4 * t + q) * 0x9E3779B9u ^ seed;
      x ^= x >> 16;
      x *= 0x85EBCA6Bu;
      x ^= x >> 13;
      x *= 0xC2B2AE35u;
      x ^= x >> 16;
      w[q] = x;
    }
    p[t] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

}  // namespace

int vcphip_fill_random(void* p, size_t bytes, uint32_t seed, hipStream_t st) {
  const long n16 = (long)(bytes / 16);
  if (n16 <= 0) return 0;
  const int grid = (int)std::min<long>((n16 + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_random, dim3(grid), dim3(256), 0, st, (uint4*)p, n16, seed);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_gather_rows(const double* A, long lda, int m, int k, int s, double* out, hipStream_t st) {
  if (s <= 0 || k <= 0) return 0;
  const long total = (long)s * k;
  const int grid = (int)std::min<long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_gather_rows, dim3(grid), dim3(256), 0, st, (const double2*)A, lda, m, k, s, (double2*)out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_ratio_max(const double* X, const double* Y, long total, double* part, int nblocks, hipStream_t st) {
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_ratio_max, dim3(nblocks), dim3(256), 0, st, (const double2*)X, (const double2*)Y, total, part);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
