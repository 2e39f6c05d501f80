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

// BM = 16 TM: a 16 x 16 thread grid, TM x TM outputs per thread
template <int BM, int TM>
__global__ __launch_bounds__(NT) void k_uhphiphi(int m, int npts, const double* __restrict__ phi, long ldp,
                                                 const double* __restrict__ w, const double* __restrict__ u,
                                                 double* __restrict__ A, long lda) {
  static_assert(BM == 16 * TM && (BM * BK) % NT == 0, "tile shape");
  constexpr int LA = BM + 2, RPP = NT / BM;   // rows per staging pass: BM rows x RPP points
  vcphip_mode_round_up();
  __shared__ __attribute__((aligned(16))) double sSl[BK][LA], sSu[BK][LA], sPl[BK][LA], sPu[BK][LA];
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
  auto store_stage = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int kl = lk + RPP * q, k = k0 + kl;
      double2 sv = make_double2(-0.0, 0.0);
      if (k < npts && row0 + lr < m) {
        const double2 v = imul_neg(gw[q].x, gw[q].y, gu[q].x, gu[q].y);   // w_k * u_k
        sv = imul_neg(-v.x, v.y, ga[q].x, ga[q].y);                     // (w_k u_k) * phi(i, k)
      }
      sSl[kl][lr] = -sv.x;
      sSu[kl][lr] = sv.y;
      sPl[kl][lr] = gb[q].x;
      sPu[kl][lr] = gb[q].y;
    }
  };
  if (npts > 0) {
    load_stage(0);
    store_stage(0);
  }
  __syncthreads();
  for (int k0 = 0; k0 < npts; k0 += BK) {
    const bool more = k0 + BK < npts;
    if (more) load_stage(k0 + BK);
#pragma unroll 2
    for (int kk = 0; kk < BK; ++kk) {
      double sl[TM], su[TM], pl[TM], pu[TM];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        sl[i] = sSl[kk][ty * TM + i];
        su[i] = sSu[kk][ty * TM + i];
        pl[i] = sPl[kk][tx * TM + i];
        pu[i] = sPu[kk][tx * TM + i];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TM; ++j) {
          const double2 p = imul_neg(sl[i], su[i], pl[j], pu[j]);   // s(i, k) * phi(j, k)
          up[i][j] = up[i][j] + p.y;
          dn[i][j] = dn[i][j] + p.x;
        }
    }
    __syncthreads();
    if (more) {
      store_stage(k0 + BK);
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TM; ++j) {
      const int r = row0 + ty * TM + i, c = col0 + tx * TM + j;
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
