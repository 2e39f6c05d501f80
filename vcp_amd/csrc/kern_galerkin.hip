// Galerkin matrix producer: Legendre_Bases_Generator::uhphiphi
// (vcp/ldbase.hpp:955-1013) in interval arithmetic,
//   A(i, j) = sum over points k of ((w_k * u_k) * phi(i, k)) * phi(j, k),
// i <= j accumulated from [0, 0] in k order and mirrored to A(j, i).  Each
// product and each sum is a separately rounded interval operation (kv's
// interval *, +): a product's lower end is the smallest of the four
// downward-rounded endpoint products, its upper end the largest of the four
// upward-rounded ones.  The kernel runs in round-toward-+inf and carries
// lower ends negated (RD(x) = -RU(-x)).
//
// Tiles: BM x BM outputs of the upper triangle (diagonal tiles write their
// upper part and its mirror), 16 points per LDS stage; 64 x 64 when that
// gives enough workgroups to fill the chip, else 32 x 32.  The row operand is
// s(i, k) = (w_k * u_k) * phi(i, k), formed while staging; the column
// operand is phi(j, k).  Both come from the column-major phi (m x npts), so
// each stage reads 16 contiguous columns.
#include "vcphip_device.h"
#include "vcphip_kernels.h"

namespace {

constexpr int BK = 16, NT = 256;

// interval product in round-up mode: returns {-lower, upper}
VCPHIP_DEVINL double2 imul_neg(double al, double au, double bl, double bu) {
  const double up = __builtin_fmax(__builtin_fmax(al * bl, al * bu), __builtin_fmax(au * bl, au * bu));
  const double nlo = __builtin_fmax(__builtin_fmax((-al) * bl, (-al) * bu), __builtin_fmax((-au) * bl, (-au) * bu));
  return make_double2(nlo, up);
}

// An interval operand of the inner products takes the two-product form
// below when it is finite and lies strictly on one side of zero (lo > 0 or
// hi < 0); padding elements (outside the matrix or past the last point)
// are [0, 0] in both forms.
VCPHIP_DEVINL bool sign_definite(double lo, double hi) {
  return (lo > 0.0 || hi < 0.0) && __builtin_fabs(lo) < __builtin_inf() && __builtin_fabs(hi) < __builtin_inf();
}

// BM = 16 TM: a 16 x 16 thread grid, TM x TM outputs per thread: thread
// (tx, ty) holds rows ty TM .. ty TM + TM - 1 and columns tx + 16 j of its
// tile, so that a wave reads the column operands of a point from
// consecutive 16-byte LDS slots (no bank conflicts) and the row operands as
// broadcasts
template <int BM, int TM>
__global__ __launch_bounds__(NT, 2) void k_uhphiphi(int m, int npts, const double* __restrict__ phi, long ldp,
                                                    const double* __restrict__ w, const double* __restrict__ u,
                                                    double* __restrict__ A, long lda) {
  static_assert(BM == 16 * TM && (BM * BK) % NT == 0, "tile shape");
  constexpr int LA = BM + 1, RPP = NT / BM;   // rows per staging pass: BM rows x RPP points
  vcphip_mode_round_up();
  // per point: s(i, k) and phi(j, k) as {lower, upper} (general form) and
  // as {end of smaller magnitude, end of larger magnitude} (two-product form)
  __shared__ double2 sS[BK][LA], sP[BK][LA], fS[BK][LA], fP[BK][LA];
  // upper-triangle tile t -> (tm, tn), tm <= tn
  const int nt = (m + BM - 1) / BM;
  int t = blockIdx.x, tm = 0;
  while (t >= nt - tm) {
    t -= nt - tm;
    ++tm;
  }
  const int tn = tm + t;
  const int row0 = tm * BM, col0 = tn * BM;
  const int tid = threadIdx.x;
  const int lr = tid % BM, lk = tid / BM;   // staging: BM rows x RPP points per pass
  const int tx = tid % 16, ty = tid / 16;
  double up[TM][TM], dn[TM][TM];   // dn = -(lower end)
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TM; ++j) {
      up[i][j] = 0.0;
      dn[i][j] = -0.0;
    }
  // stage k0 in registers: the global loads of the next stage are issued
  // before the current stage's arithmetic and stored to LDS after it
  constexpr int NQ = BK / RPP;
  double2 ga[NQ], gb[NQ], gw[NQ], gu[NQ];
  auto load_stage = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int k = k0 + lk + RPP * q;
      const bool kok = k < npts;
      gw[q] = kok ? *reinterpret_cast<const double2*>(w + 2 * k) : make_double2(0.0, 0.0);
      gu[q] = kok ? *reinterpret_cast<const double2*>(u + 2 * k) : make_double2(0.0, 0.0);
      ga[q] = (kok && row0 + lr < m) ? *reinterpret_cast<const double2*>(phi + 2 * ((row0 + lr) + ldp * k))
                                     : make_double2(0.0, 0.0);
      gb[q] = (kok && col0 + lr < m) ? *reinterpret_cast<const double2*>(phi + 2 * ((col0 + lr) + ldp * k))
                                     : make_double2(0.0, 0.0);
    }
  };
  // a negative interval's end of smaller magnitude is its upper end
  auto small_large = [](double lo, double hi) { return hi < 0.0 ? make_double2(hi, lo) : make_double2(lo, hi); };
  // stores stage k0; returns whether this thread's elements all take the
  // two-product form
  auto store_stage = [&](int k0) {
    bool ok = true;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int kl = lk + RPP * q, k = k0 + kl;
      double2 sv = make_double2(-0.0, 0.0);
      if (k < npts && row0 + lr < m) {
        const double2 v = imul_neg(gw[q].x, gw[q].y, gu[q].x, gu[q].y);   // w_k * u_k
        sv = imul_neg(-v.x, v.y, ga[q].x, ga[q].y);                     // (w_k u_k) * phi(i, k)
        ok = ok && sign_definite(-sv.x, sv.y);
      }
      if (k < npts && col0 + lr < m) ok = ok && sign_definite(gb[q].x, gb[q].y);
      sS[kl][lr] = make_double2(-sv.x, sv.y);
      sP[kl][lr] = gb[q];
      fS[kl][lr] = small_large(-sv.x, sv.y);
      fP[kl][lr] = small_large(gb[q].x, gb[q].y);
    }
    return ok;
  };
  // two_product: every operand of the stage takes the two-product form
  bool two_product = false;
  if (npts > 0) {
    load_stage(0);
    two_product = __syncthreads_and(store_stage(0)) != 0;
  } else {
    __syncthreads();
  }
  for (int k0 = 0; k0 < npts; k0 += BK) {
    const bool more = k0 + BK < npts;
    if (more) load_stage(k0 + BK);
    if (two_product) {
      // s(i, k) * phi(j, k) for operands strictly on one side of zero: with
      // a, A the ends of s of smaller and larger magnitude and b, B those of
      // phi, the exact product's ends are a b and A B (both of the sign of
      // the product), so its rounded ends are max(RU(a b), RU(A B)) and
      // -max(RU(-a b), RU(-A B)): four products instead of eight and no
      // comparisons of products, with the bits of imul_neg (rounding is
      // monotone, and four products of one sign give no +0 / -0 ties)
      for (int kk = 0; kk < BK; ++kk) {
        double2 a[TM], b[TM];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          a[i] = fS[kk][ty * TM + i];
          b[i] = fP[kk][tx + 16 * i];
        }
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          // a row's products first, then their maxima and sums: independent
          // products in flight instead of one mul-max-add chain at a time
          double pu[TM][2], pd[TM][2];
#pragma unroll
          for (int j = 0; j < TM; ++j) {
            pu[j][0] = a[i].x * b[j].x;
            pu[j][1] = a[i].y * b[j].y;
            pd[j][0] = (-a[i].x) * b[j].x;
            pd[j][1] = (-a[i].y) * b[j].y;
          }
#pragma unroll
          for (int j = 0; j < TM; ++j) {
            up[i][j] = up[i][j] + __builtin_fmax(pu[j][0], pu[j][1]);
            dn[i][j] = dn[i][j] + __builtin_fmax(pd[j][0], pd[j][1]);
          }
        }
      }
    } else {
#pragma unroll 2
      for (int kk = 0; kk < BK; ++kk) {
        double2 a[TM], b[TM];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          a[i] = sS[kk][ty * TM + i];
          b[i] = sP[kk][tx + 16 * i];
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TM; ++j) {
            const double2 p = imul_neg(a[i].x, a[i].y, b[j].x, b[j].y);   // s(i, k) * phi(j, k)
            up[i][j] = up[i][j] + p.y;
            dn[i][j] = dn[i][j] + p.x;
          }
      }
    }
    __syncthreads();
    if (more) two_product = __syncthreads_and(store_stage(k0 + BK)) != 0;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TM; ++j) {
      const int r = row0 + ty * TM + i, c = col0 + tx + 16 * j;
      if (r < m && c < m && r <= c) {
        const double2 val = make_double2(-dn[i][j], up[i][j]);
        *reinterpret_cast<double2*>(A + 2 * (r + lda * (long)c)) = val;
        *reinterpret_cast<double2*>(A + 2 * (c + lda * (long)r)) = val;
      }
    }
}

}  // namespace

int vcphip_launch_uhphiphi(int m, int npts, const double* phi, long ldp, const double* w, const double* u, double* A,
                           long lda, hipStream_t s) {
  if (m <= 0) return 0;
  const int nt64 = (m + 63) / 64;
  if (nt64 * (nt64 + 1) / 2 >= 1024) {
    hipLaunchKernelGGL((k_uhphiphi<64, 4>), dim3(nt64 * (nt64 + 1) / 2), dim3(NT), 0, s, m, npts, phi, ldp, w, u, A,
                       lda);
  } else {
    const int nt = (m + 31) / 32;
    hipLaunchKernelGGL((k_uhphiphi<32, 2>), dim3(nt * (nt + 1) / 2), dim3(NT), 0, s, m, npts, phi, ldp, w, u, A, lda);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
