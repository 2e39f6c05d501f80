// Plain FP64 directed-rounding GEMM family (vector FMA, LDS-tiled, gfx950).
//
// One kernel template covers the four products of the pidblas policy:
//   II  C = IA*IB        pidblas::mulmm     vcp/pidblas.hpp:239-294
//   IP  C = IA*B         pidblas::mul_im_m  vcp/pidblas.hpp:90-140
//   PI  C = B*IA         pidblas::mul_m_im  vcp/pidblas.hpp:142-196
//   PP  Cup, Cdn = A*B   rmatmul(+1/-1)     vblas/rmatmul.hpp:100-127,
//                        udmatmul           vblas/udmatmul.hpp:738-760
// Each output element carries the k-ascending chains of the reference's
// BLAS calls (the rmatmul contract, vblas/README.md:167-170) in registers:
//   up  = RU chain of mid(A)*mid(B)                       (:269-271)
//   dn  = RU chain of (-mid(A))*mid(B) = -(RD chain)      (:264-266)
//   r1  = RU chain of (|mid A|+rad A)*rad B               (:277-279)
//   r2  = RU chain of rad A*|mid B|                       (:281)
// and the epilogue applies the reference's fix-ups in its order
//   t = RU(r1 + r2); upper = RU(up + t); lower = RD(-dn - t) = -RU(dn + t).
// Two implementations share these chains: k_drgemm_planes (the single
// products), fed from midpoint-radius planes made once per operand, and the
// tile function dr_tile (k_drgemm for the thin products, the batched
// launches), which reads the AoS {inf, sup} operands and converts them to
// midpoint-radius on their way into LDS (pidblas_assist::midrad, :43-60).
//
// The whole kernel runs in round-toward-+inf (see vcphip_device.h).
#include "vcphip_device.h"
#include "vcphip_kernels.h"

namespace {

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int TM = 4, TN = 4;
constexpr int NT = 256;          // 16 x 16 threads, 4 waves
// thread (tx, ty) holds rows 4 ty .. 4 ty + 3 of a tile and four of its
// columns: 2 tx, 2 tx + 1, 2 tx + 32, 2 tx + 33 in dr_tile (dr_col), tx + 16 j
// in the syrk and inf-sup tiles.  Either way a wave reads B from consecutive
// LDS addresses (no bank conflicts) and A as a broadcast of four row
// quadruples
constexpr int PAD = 2;           // doubles of row padding (keeps 16-byte alignment)
constexpr int LA = BM + PAD;     // LDS row length of an A k-slice
constexpr int LB = BN + PAD;
constexpr int A_PER_T = BM * BK / NT;   // 4 elements of A per thread per tile
constexpr int B_PER_T = BK * BN / NT;   // 4 elements of B per thread per tile

template <bool AI, bool BI, int APT>
struct Stage {
  // register staging of one global tile (prefetch)
  double a0[APT], a1[APT];
  double b0[B_PER_T], b1[B_PER_T];
};

// LDS of one k-slice: midpoints and radii of A (rows) and B (columns).  One
// instance per kernel, shared by every tile shape the kernel runs; a tile of
// 32 rows uses the first 32 (+ padding) doubles of each A row.
template <bool AI, bool BI>
struct DrLds {
  double Am[BK][LA];
  double Ar[AI ? BK : 1][LA];
  double Bm[BK][LB];
  double Br[BI ? BK : 1][LB];
};

// column of output j (< 4) of thread column tx in a 64-wide tile: pairs
// 2 tx, 2 tx + 1 and 2 tx + 32, 2 tx + 33, so a wave reads B as 16-byte
// pieces at consecutive addresses
__device__ __forceinline__ int dr_col(int tx, int j) { return 2 * tx + (j & 1) + 32 * (j >> 1); }

// One (16 TMR) x BN output tile at (row0, col0): TMR = 4 gives the 64 x 64
// tile, TMR = 2 the 32 x 64 tile of the batched launches' short problems and
// edge rows (each output's chains are the same in both, so are its bits).
// RESID: the batched residual form R*A - I, whose diagonal elements get the
// interval subtraction of 1 (imats.hpp:324-326).  Runs in the caller's
// round-up mode.
template <bool AI, bool BI, bool RESID, int TMR = TM, class Lds = DrLds<AI, BI>>
__device__ __forceinline__ void dr_tile(Lds& L, int m, int n, int k, const double* __restrict__ A,
                                        long lda, const double* __restrict__ B, long ldb, double* __restrict__ C,
                                        long ldc, double* __restrict__ C2, int accumulate, int row0, int col0) {
  constexpr bool II = AI && BI;
  constexpr int TBM = 16 * TMR;              // tile rows
  constexpr int APT = TBM * BK / NT;         // elements of A per thread per slice
  constexpr int AKS = NT / TBM;              // k stride of a thread's A elements
  static_assert(TMR == 4 || TMR == 2, "tile rows");
  auto& sAm = L.Am;
  auto& sAr = L.Ar;
  auto& sBm = L.Bm;
  auto& sBr = L.Br;

  const int tid = threadIdx.x;

  // global-load mapping: A column-major, TBM consecutive rows per k column;
  // B column-major, 16 consecutive k per column.
  const int a_row = tid % TBM, a_kc = tid / TBM;   // k column a_kc + AKS q
  const int b_k = tid % BK, b_col = tid / BK;      // column b_col + 16q

  Stage<AI, BI, APT> st;
  // global loads of one k-slice; a slice whose rows, columns and k lie
  // inside the operands (every slice of an interior tile but a partial last
  // one) takes the unchecked path (the choice is uniform per workgroup).  The
  // interval x point form keeps the checked path only: the second path makes
  // it spill at three waves per SIMD.
  auto load_tile = [&](int k0) {
    constexpr bool FAST = II || !AI;
    const bool inside = FAST && row0 + TBM <= m && col0 + BN <= n && k0 + BK <= k;
    if (inside) {
#pragma unroll
      for (int q = 0; q < APT; ++q) {
        const int kk = k0 + a_kc + AKS * q, r = row0 + a_row;
        if (AI) {
          const double2 v = *reinterpret_cast<const double2*>(A + 2 * (r + lda * kk));
          st.a0[q] = v.x;
          st.a1[q] = v.y;
        } else {
          st.a0[q] = A[r + lda * kk];
        }
      }
#pragma unroll
      for (int q = 0; q < B_PER_T; ++q) {
        const int kk = k0 + b_k, c = col0 + b_col + 16 * q;
        if (BI) {
          const double2 v = *reinterpret_cast<const double2*>(B + 2 * (kk + ldb * c));
          st.b0[q] = v.x;
          st.b1[q] = v.y;
        } else {
          st.b0[q] = B[kk + ldb * c];
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < APT; ++q) {
      const int kk = k0 + a_kc + AKS * q, r = row0 + a_row;
      const bool ok = (r < m) && (kk < k);
      if (AI) {
        double2 v = ok ? *reinterpret_cast<const double2*>(A + 2 * (r + lda * kk)) : make_double2(0.0, 0.0);
        st.a0[q] = v.x;
        st.a1[q] = v.y;
      } else {
        st.a0[q] = ok ? A[r + lda * kk] : 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < B_PER_T; ++q) {
      const int kk = k0 + b_k, c = col0 + b_col + 16 * q;
      const bool ok = (kk < k) && (c < n);
      if (BI) {
        double2 v = ok ? *reinterpret_cast<const double2*>(B + 2 * (kk + ldb * c)) : make_double2(0.0, 0.0);
        st.b0[q] = v.x;
        st.b1[q] = v.y;
      } else {
        st.b0[q] = ok ? B[kk + ldb * c] : 0.0;
      }
    }
  };
  auto store_tile = [&]() {
#pragma unroll
    for (int q = 0; q < APT; ++q) {
      if (AI) {
        double mm, rr;
        vcphip_midrad_ru(st.a0[q], st.a1[q], mm, rr);
        sAm[a_kc + AKS * q][a_row] = mm;
        sAr[a_kc + AKS * q][a_row] = rr;
      } else {
        sAm[a_kc + AKS * q][a_row] = st.a0[q];
      }
    }
#pragma unroll
    for (int q = 0; q < B_PER_T; ++q) {
      if (BI) {
        double mm, rr;
        vcphip_midrad_ru(st.b0[q], st.b1[q], mm, rr);
        sBm[b_k][b_col + 16 * q] = mm;
        sBr[b_k][b_col + 16 * q] = rr;
      } else {
        sBm[b_k][b_col + 16 * q] = st.b0[q];
      }
    }
  };

  const int tx = tid % 16, ty = tid / 16;
  double up[TMR][TN], dn[TMR][TN], r1[TMR][TN], r2[TMR][TN];
#pragma unroll
  for (int i = 0; i < TMR; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      up[i][j] = 0.0;
      dn[i][j] = 0.0;
      r1[i][j] = 0.0;
      r2[i][j] = 0.0;
    }

  if (!AI && !BI && accumulate) {
    // udmatmul form: the chains continue from the values already in CU / CD
#pragma unroll
    for (int i = 0; i < TMR; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = row0 + ty * TMR + i, c = col0 + dr_col(tx, j);
        if (r < m && c < n) {
          up[i][j] = C[r + ldc * c];
          dn[i][j] = -C2[r + ldc * c];
        }
      }
  }

  const int ntiles = (k + BK - 1) / BK;
  load_tile(0);
  store_tile();
  __syncthreads();
  for (int t = 0; t < ntiles; ++t) {
    if (t + 1 < ntiles) load_tile((t + 1) * BK);
#pragma unroll 2
    for (int kk = 0; kk < BK; ++kk) {
      double am[TMR], ar[TMR], bm[TN], br[TN];
      {
#pragma unroll
        for (int h = 0; h < TMR / 2; ++h) {
          const double2 x = *reinterpret_cast<const double2*>(&sAm[kk][ty * TMR + 2 * h]);
          am[2 * h] = x.x;
          am[2 * h + 1] = x.y;
        }
        const double2 y0 = *reinterpret_cast<const double2*>(&sBm[kk][2 * tx]);
        const double2 y1 = *reinterpret_cast<const double2*>(&sBm[kk][2 * tx + 32]);
        bm[0] = y0.x; bm[1] = y0.y; bm[2] = y1.x; bm[3] = y1.y;
      }
      if (AI) {
#pragma unroll
        for (int h = 0; h < TMR / 2; ++h) {
          const double2 x = *reinterpret_cast<const double2*>(&sAr[kk][ty * TMR + 2 * h]);
          ar[2 * h] = x.x;
          ar[2 * h + 1] = x.y;
        }
      }
      if (BI) {
        const double2 y0 = *reinterpret_cast<const double2*>(&sBr[kk][2 * tx]);
        const double2 y1 = *reinterpret_cast<const double2*>(&sBr[kk][2 * tx + 32]);
        br[0] = y0.x; br[1] = y0.y; br[2] = y1.x; br[3] = y1.y;
      }
#pragma unroll
      for (int i = 0; i < TMR; ++i) {
        // (|mA| + rA), RU (:277): an add per k and row, not a third LDS array,
        // since the tile's LDS reads already match its FMA time
        double as = 0.0;
        if (II) as = __builtin_fabs(am[i]) + ar[i];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          up[i][j] = __builtin_fma(am[i], bm[j], up[i][j]);
          dn[i][j] = __builtin_fma(-am[i], bm[j], dn[i][j]);
          if (II) {
            r1[i][j] = __builtin_fma(as, br[j], r1[i][j]);   // (|mA| + rA) rB (:277-279)
            r2[i][j] = __builtin_fma(ar[i], __builtin_fabs(bm[j]), r2[i][j]);
          } else if (AI) {
            r1[i][j] = __builtin_fma(ar[i], __builtin_fabs(bm[j]), r1[i][j]);   // rA*|B|  (:129)
          } else if (BI) {
            r1[i][j] = __builtin_fma(__builtin_fabs(am[i]), br[j], r1[i][j]);   // |B|*rA  (:185)
          }
        }
      }
    }
    __syncthreads();
    if (t + 1 < ntiles) {
      store_tile();
      __syncthreads();
    }
  }

  // epilogue
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int c = col0 + dr_col(tx, j);
    if (c >= n) continue;
#pragma unroll
    for (int i = 0; i < TMR; ++i) {
      const int r = row0 + ty * TMR + i;
      if (r >= m) continue;
      if (AI || BI) {
        double t;
        if (II) t = r1[i][j] + r2[i][j];    // temp = (|mA|+rA) rB + rA |mB|   (:283)
        else t = r1[i][j];
        double hi = up[i][j] + t;             // uC += temp, RU                  (:284)
        double nlo = dn[i][j] + t;            // lC -= temp, RD = -(RU(-lC + t)) (:285-286)
        if (RESID && r == c) {                // b(i,i) -= 1 in interval arithmetic
          hi = hi - 1.0;
          nlo = nlo + 1.0;
        }
        *reinterpret_cast<double2*>(C + 2 * (r + ldc * c)) = make_double2(-nlo, hi);
      } else {
        C[r + ldc * c] = up[i][j];
        C2[r + ldc * c] = -dn[i][j];
      }
    }
  }
}

template <bool AI, bool BI>
// interval x interval tiles hold four chains per output (2 waves per SIMD);
// the other forms fit in 168 VGPRs (3 waves per SIMD)
__global__ __launch_bounds__(NT, (AI && BI) ? 2 : 3) void k_drgemm(int m, int n, int k, const double* __restrict__ A, long lda,
                                                   const double* __restrict__ B, long ldb, double* __restrict__ C,
                                                   long ldc, double* __restrict__ C2, int accumulate) {
  vcphip_mode_round_up();
  const int ntm = (m + BM - 1) / BM, ntn = (n + BN - 1) / BN;
  const int tile = vcphip_xcd_remap(blockIdx.x, ntm * ntn);
  int tm, tn;
  vcphip_tile_coords(tile, ntm, ntn, 16, tm, tn);
  __shared__ __attribute__((aligned(16))) DrLds<AI, BI> lds;
  dr_tile<AI, BI, false>(lds, m, n, k, A, lda, B, ldb, C, ldc, C2, accumulate, tm * BM, tn * BN);
}

// ---------------------------------------------------------------------------
// The single products from operand planes.
//
// pidblas_assist::midrad (:43-60) is taken once per operand element by two
// small kernels instead of once per element and tile inside the product, and
// both operands are laid out for whole, contiguous k-slices:
//   A (m x k, column-major) -> mid(A), rad(A) (interval) or A (point),
//                              mpad x kpad, column-major (ld mpad)
//   B (k x n, column-major) -> mid(B), rad(B) or B, kpad x npad, row-major
//                              (ld npad)
// zero beyond m, n and k (mpad, npad multiples of 64, kpad of 16).  Every
// k-slice of a tile (16 k x 64 of each plane) then lies in 16 runs of 512
// contiguous bytes, which the product loads straight into LDS with no bounds
// tests, no conversions and no register staging.  The chains and their order
// are those of dr_tile, so are the results' bits (padding adds the same zero
// terms the checked loads of dr_tile add past k).
// ---------------------------------------------------------------------------
template <bool AI>
__global__ __launch_bounds__(256) void k_planes_a(int m, int k, int mpad, const double* __restrict__ A, long lda,
                                                  double* __restrict__ P0, double* __restrict__ P1) {
  vcphip_mode_round_up();
  const int r = blockIdx.x * 256 + threadIdx.x, kk = blockIdx.y;
  if (r >= mpad) return;
  const long o = (long)kk * mpad + r;
  const bool in = r < m && kk < k;
  if (AI) {
    double mm = 0.0, rr = 0.0;
    if (in) {
      const double2 v = *reinterpret_cast<const double2*>(A + 2 * (r + lda * (long)kk));
      vcphip_midrad_ru(v.x, v.y, mm, rr);
    }
    P0[o] = mm;
    P1[o] = rr;
  } else {
    P0[o] = in ? A[r + lda * (long)kk] : 0.0;
  }
}

// B through a 32 x 32 LDS tile: read along k (B's columns), write along n
template <bool BI>
__global__ __launch_bounds__(256) void k_planes_bt(int k, int n, int kpad, int npad, const double* __restrict__ B,
                                                   long ldb, double* __restrict__ P0, double* __restrict__ P1) {
  vcphip_mode_round_up();
  __shared__ double2 s[32][33];
  const int k0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int kk = k0 + tx, j = j0 + ty + 8 * i;
    double2 v = make_double2(0.0, 0.0);
    if (kk < k && j < n) {
      if (BI) v = *reinterpret_cast<const double2*>(B + 2 * (kk + ldb * (long)j));
      else v.x = B[kk + ldb * (long)j];
    }
    s[ty + 8 * i][tx] = v;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int kk = k0 + ty + 8 * i, j = j0 + tx;
    if (kk >= kpad) continue;   // the last block of 32 k may reach past kpad
    const double2 v = s[tx][ty + 8 * i];
    const long o = (long)kk * npad + j;
    if (BI) {
      double mm, rr;
      vcphip_midrad_ru(v.x, v.y, mm, rr);   // (0, 0) for the padding
      P0[o] = mm;
      P1[o] = rr;
    } else {
      P0[o] = v.x;
    }
  }
}

// Tiles of 16 TMR x 64 outputs: TMR = 4 (64 rows) for grids that fill the
// device, TMR = 2 (32 rows) for small ones, which would leave most CUs one
// workgroup (one wave per SIMD) or idle.  Interval x interval tiles hold four
// chains per output, point x point two, the mixed forms three; the waves per
// SIMD follow (64 rows: 2 / 4 / 3, 32 rows: 3 / 4 / 4).
template <bool AI, bool BI, int TMR>
__global__ __launch_bounds__(NT, TMR == 4 ? ((AI && BI) ? 2 : (AI || BI) ? 3 : 4) : ((AI && BI) ? 3 : 4))
void k_drgemm_planes(int m, int n, int mpad, int npad, int kpad, const double* __restrict__ A0,
                     const double* __restrict__ A1, const double* __restrict__ B0, const double* __restrict__ B1,
                     double* __restrict__ C, long ldc, double* __restrict__ C2, int accumulate) {
  constexpr bool II = AI && BI;
  constexpr int TBM = 16 * TMR;                  // tile rows
  constexpr int NA = AI ? 2 : 1, NB = BI ? 2 : 1, NP = NA + NB;
  constexpr int ASL = BK * TBM, BSL = BK * BN;   // doubles of one plane's k-slice
  constexpr int STAGE = NA * ASL + NB * BSL;
  static_assert(TMR == 4 || TMR == 2, "tile rows");
  static_assert(BN == 64 && BK == 16, "slice layout");
  vcphip_mode_round_up();
  // [stage][A mid (or A), A rad: k x TBM][B mid (or B), B rad: k x 64]
  __shared__ __attribute__((aligned(16))) double sL[2 * STAGE];
  const int ntm = mpad / TBM, ntn = npad / BN;
  const int tile = vcphip_xcd_remap(blockIdx.x, ntm * ntn);
  int tm, tn;
  vcphip_tile_coords(tile, ntm, ntn, 16, tm, tn);
  const int row0 = tm * TBM, col0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63;
  // wave w < NP loads plane w of every slice, one k-row of a plane slice
  // per 2^lh lanes (two doubles a lane): a load instruction moves 1 KB, i.e.
  // 64 / 2^lh k-rows, and a slice takes BK 2^lh / 64 of them
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool loader = w < NP;
  const bool wa = w < NA;
  const int lh = (wa && TBM == 32) ? 4 : 5;
  const int nq = BK >> (6 - lh), rpi = 64 >> lh;
  const double* plane = wa ? (w == 0 ? A0 : A1) : (w == NA ? B0 : B1);
  const long ld = wa ? mpad : npad;
  const double* src = plane + (long)(lane >> lh) * ld + (wa ? row0 : col0) + 2 * (lane & ((1 << lh) - 1));
  const int poff = !loader ? 0 : wa ? w * ASL : NA * ASL + (w - NA) * BSL;
  const uint32_t lds0 = vcphip_lds_addr(sL) + (uint32_t)poff * 8;
  auto issue = [&](int stage, int t) {
    if (!loader) return;
    const double* g = src + (long)t * BK * ld;
    const uint32_t d = lds0 + (uint32_t)stage * (STAGE * 8);
    for (int q = 0; q < nq; ++q) vcphip_lds_dma16(g + (long)(rpi * q) * ld, d + (uint32_t)q * 1024);
  };

  const int tx = tid % 16, ty = tid / 16;
  double up[TMR][TN], dn[TMR][TN], r1[TMR][TN], r2[TMR][TN];
#pragma unroll
  for (int i = 0; i < TMR; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      up[i][j] = 0.0;
      dn[i][j] = 0.0;
      r1[i][j] = 0.0;
      r2[i][j] = 0.0;
    }
  if (!AI && !BI && accumulate) {
    // udmatmul form: the chains continue from the values already in CU / CD
#pragma unroll
    for (int i = 0; i < TMR; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = row0 + ty * TMR + i, c = col0 + dr_col(tx, j);
        if (r < m && c < n) {
          up[i][j] = C[r + ldc * c];
          dn[i][j] = -C2[r + ldc * c];
        }
      }
  }

  const int nt = kpad / BK;
  issue(0, 0);
  for (int t = 0; t < nt; ++t) {
    // only slice t's loads are in flight: wait for them, then for every wave
    // (which also ends every wave's reads of the other stage, slice t - 1)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (t + 1 < nt) issue((t + 1) & 1, t + 1);
    const double* sAm = sL + (t & 1) * STAGE;
    const double* sAr = sAm + (AI ? ASL : 0);
    const double* sBm = sAm + NA * ASL;
    const double* sBr = sBm + (BI ? BSL : 0);
#pragma unroll 2
    for (int kk = 0; kk < BK; ++kk) {
      double am[TMR], ar[TMR], bm[TN], br[TN];
#pragma unroll
      for (int h = 0; h < TMR / 2; ++h) {
        const double2 x = *reinterpret_cast<const double2*>(&sAm[kk * TBM + ty * TMR + 2 * h]);
        am[2 * h] = x.x;
        am[2 * h + 1] = x.y;
      }
      {
        const double2 y0 = *reinterpret_cast<const double2*>(&sBm[kk * BN + 2 * tx]);
        const double2 y1 = *reinterpret_cast<const double2*>(&sBm[kk * BN + 2 * tx + 32]);
        bm[0] = y0.x; bm[1] = y0.y; bm[2] = y1.x; bm[3] = y1.y;
      }
      if (AI) {
#pragma unroll
        for (int h = 0; h < TMR / 2; ++h) {
          const double2 z = *reinterpret_cast<const double2*>(&sAr[kk * TBM + ty * TMR + 2 * h]);
          ar[2 * h] = z.x;
          ar[2 * h + 1] = z.y;
        }
      }
      if (BI) {
        const double2 u0 = *reinterpret_cast<const double2*>(&sBr[kk * BN + 2 * tx]);
        const double2 u1 = *reinterpret_cast<const double2*>(&sBr[kk * BN + 2 * tx + 32]);
        br[0] = u0.x; br[1] = u0.y; br[2] = u1.x; br[3] = u1.y;
      }
#pragma unroll
      for (int i = 0; i < TMR; ++i) {
        double as = 0.0;
        if (II) as = __builtin_fabs(am[i]) + ar[i];   // (|mA| + rA), RU (:277)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          up[i][j] = __builtin_fma(am[i], bm[j], up[i][j]);
          dn[i][j] = __builtin_fma(-am[i], bm[j], dn[i][j]);
          if (II) {
            r1[i][j] = __builtin_fma(as, br[j], r1[i][j]);                    // (|mA| + rA) rB (:277-279)
            r2[i][j] = __builtin_fma(ar[i], __builtin_fabs(bm[j]), r2[i][j]);  // rA |mB| (:281)
          } else if (AI) {
            r1[i][j] = __builtin_fma(ar[i], __builtin_fabs(bm[j]), r1[i][j]);  // rA*|B|  (:129)
          } else if (BI) {
            r1[i][j] = __builtin_fma(__builtin_fabs(am[i]), br[j], r1[i][j]);  // |B|*rA  (:185)
          }
        }
      }
    }
  }

#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int c = col0 + dr_col(tx, j);
    if (c >= n) continue;
#pragma unroll
    for (int i = 0; i < TMR; ++i) {
      const int r = row0 + ty * TMR + i;
      if (r >= m) continue;
      if (AI || BI) {
        const double tt = II ? r1[i][j] + r2[i][j] : r1[i][j];   // temp (:283, :129, :185)
        const double hi = up[i][j] + tt;                         // uC += temp, RU (:284)
        const double nlo = dn[i][j] + tt;                        // lC -= temp, RD = -(RU(-lC + t)) (:285-286)
        *reinterpret_cast<double2*>(C + 2 * (r + ldc * c)) = make_double2(-nlo, hi);
      } else {
        C[r + ldc * c] = up[i][j];
        C2[r + ldc * c] = -dn[i][j];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Batched verified residual C_p = R_p * IA_p - I (imats::linearsolve,
// vcp/imats.hpp:318-326) for many problems of uneven size in one persistent
// launch per tile shape (TMR = 4: 64 x 64, TMR = 2: 32 x 64): workgroups
// take the tiles of a list sorted by problem size (largest first) from a
// work queue, each tile being the PI form above.
// ---------------------------------------------------------------------------
// Work queue of the persistent launches: workgroup thread 0 takes the next
// tile index from the launch's counter (zeroed by k_queue_reset just before
// the launch, on the same stream), so tiles are taken in list order as
// workgroups come free.  (A fixed stride of the grid over the list was
// 15-20% slower on uneven sets: a workgroup that drew long tiles held up the
// end of the launch.)
__device__ __forceinline__ int next_tile(unsigned* __restrict__ queue, int& slot) {
  __syncthreads();   // every thread has read the previous index (and left the tile's LDS)
  if (threadIdx.x == 0) slot = (int)atomicAdd(queue, 1u);
  __syncthreads();
  return slot;
}

__global__ __launch_bounds__(64) void k_queue_reset(unsigned* __restrict__ queue) {
  if (threadIdx.x < 2) queue[threadIdx.x] = 0u;
}

template <int TMR>
__global__ __launch_bounds__(NT, 3) void k_resid_tiles(int ntiles, const VcphipResidTile* __restrict__ tiles,
                                                        const VcphipResidProblem* __restrict__ probs,
                                                        unsigned* __restrict__ queue) {
  vcphip_mode_round_up();
  __shared__ __attribute__((aligned(16))) DrLds<false, true> lds;
  __shared__ int slot;
  for (int t = next_tile(queue, slot); t < ntiles; t = next_tile(queue, slot)) {
    const VcphipResidTile tl = tiles[t];
    const VcphipResidProblem p = probs[tl.problem];
    dr_tile<false, true, true, TMR>(lds, p.n, p.n, p.n, p.R, p.ldr, p.A, p.lda, p.C, p.ldc, nullptr, 0, tl.row0,
                                    tl.col0);
  }
}

// Many interval products C_p = IA_p * IB_p of uneven (m, n, k) in one
// persistent launch per tile shape (pidblas::mulmm, vcp/pidblas.hpp:239-294,
// per problem): workgroups take the tiles of a list sorted by problem size
// from a work queue, each tile the interval x interval form of k_drgemm, so
// every C_p is bit-identical to the single product's.
template <int TMR>
__global__ __launch_bounds__(NT, 2) void k_igemm_tiles(int ntiles, const VcphipResidTile* __restrict__ tiles,
                                                        const VcphipIgemmProblem* __restrict__ probs,
                                                        unsigned* __restrict__ queue) {
  vcphip_mode_round_up();
  __shared__ __attribute__((aligned(16))) DrLds<true, true> lds;
  __shared__ int slot;
  for (int t = next_tile(queue, slot); t < ntiles; t = next_tile(queue, slot)) {
    const VcphipResidTile tl = tiles[t];
    const VcphipIgemmProblem p = probs[tl.problem];
    dr_tile<true, true, false, TMR>(lds, p.m, p.n, p.k, p.A, p.lda, p.B, p.ldb, p.C, p.ldc, nullptr, 0, tl.row0,
                                    tl.col0);
  }
}

// row sums of |C| (upward, in column order: mats::norminf, vcp/mats.hpp:1583-1591)
__global__ __launch_bounds__(256) void k_resid_rowsum(int nrows, const int* __restrict__ row_problem,
                                                      const VcphipResidProblem* __restrict__ probs,
                                                      double* __restrict__ rowsum) {
  vcphip_mode_round_up();
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= nrows) return;
  const int pi = row_problem[g];
  const VcphipResidProblem p = probs[pi];
  const int i = g - p.row_base;
  const double* row = p.C + 2 * (long)i;
  const long step = 2 * (long)p.ldc;
  // the upward sum runs in column order, as the reference's; a thread per
  // row is few waves per SIMD, so the loads of 32 columns are issued ahead of
  // their additions
  constexpr int U = 32;
  double s = 0.0;
  int j = 0;
  for (; j + U <= p.n; j += U) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const double2*>(row + step * (j + u));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double mag = fmax(__builtin_fabs(v[u].x), __builtin_fabs(v[u].y));
      s = (j + u == 0) ? mag : s + mag;
    }
  }
  for (; j < p.n; ++j) {
    const double2 v = *reinterpret_cast<const double2*>(row + step * j);
    const double mag = fmax(__builtin_fabs(v.x), __builtin_fabs(v.y));
    s = (j == 0) ? mag : s + mag;
  }
  rowsum[g] = s;
}

// ||C_p||_inf upper bound: the largest row sum (mats.hpp:1592-1598).  One
// workgroup per problem; the maximum is exact in any order.
__global__ __launch_bounds__(256) void k_resid_max(int count, const VcphipResidProblem* __restrict__ probs,
                                                   const double* __restrict__ rowsum, double* __restrict__ normub) {
  __shared__ double red[256];
  const int pi = blockIdx.x;
  const VcphipResidProblem p = probs[pi];
  double best = -1.0;   // row sums are >= 0
  for (int i = threadIdx.x; i < p.n; i += 256) {
    const double v = rowsum[p.row_base + i];
    best = v > best ? v : best;
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = red[threadIdx.x + h] > red[threadIdx.x] ? red[threadIdx.x + h] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) normub[pi] = p.n > 0 ? red[0] : 0.0;
}

// ---------------------------------------------------------------------------
// A^T A forms.  IA is k x n; tiles (tm <= tn) of the upper triangle only.
//   interval (pidblas::mulltmm, vcp/pidblas.hpp:296-358):
//     up, dn  = chains of mA_i . mA_j                (dsyrk, :308-315)
//     tij     = RU chain of |mA_i| . rA_j            (|mA|^T rA, :317-321)
//     tji     = RU chain of |mA_j| . rA_i
//     sij     = RU chain of rA_i . rA_j              (rA^T rA, :323)
//     T = RU(tij + RU(tji + sij)); upper = RU(up + T); lower = -RU(dn + T)
//   point (pidblas::vmulmm, :198-236): upper = RU chain, lower = RD chain.
// Element (i, j), i <= j, is written to (i, j) and (j, i).
// ---------------------------------------------------------------------------
template <bool INTV>
__global__ __launch_bounds__(NT, 2) void k_drsyrk(int k, int n, const double* __restrict__ A, long lda,
                                                  double* __restrict__ C, long ldc) {
  vcphip_mode_round_up();
  __shared__ __attribute__((aligned(16))) double sIm[BK][LA];
  __shared__ __attribute__((aligned(16))) double sIr[INTV ? BK : 1][LA];
  __shared__ __attribute__((aligned(16))) double sJm[BK][LB];
  __shared__ __attribute__((aligned(16))) double sJr[INTV ? BK : 1][LB];

  const int tid = threadIdx.x;
  const int nt = (n + BM - 1) / BM;
  // decode upper-triangle tile index b -> (tm, tn), tm <= tn, row-major over tm
  int b = blockIdx.x, tm = 0;
  while (b >= nt - tm) { b -= nt - tm; ++tm; }
  const int tn = tm + b;
  const int i0 = tm * BM, j0 = tn * BN;

  const int l_k = tid % BK, l_c = tid / BK;   // k-contiguous loads, 16 columns per pass
  double im[A_PER_T], ir[A_PER_T], jm[B_PER_T], jr[B_PER_T];
  auto load_tile = [&](int k0) {
#pragma unroll
    for (int q = 0; q < A_PER_T; ++q) {
      const int kk = k0 + l_k;
      const int ci = i0 + l_c + 16 * q, cj = j0 + l_c + 16 * q;
      if (INTV) {
        const double2 vi = (kk < k && ci < n) ? *reinterpret_cast<const double2*>(A + 2 * (kk + lda * ci)) : make_double2(0.0, 0.0);
        const double2 vj = (kk < k && cj < n) ? *reinterpret_cast<const double2*>(A + 2 * (kk + lda * cj)) : make_double2(0.0, 0.0);
        im[q] = vi.x; ir[q] = vi.y; jm[q] = vj.x; jr[q] = vj.y;
      } else {
        im[q] = (kk < k && ci < n) ? A[kk + lda * ci] : 0.0;
        jm[q] = (kk < k && cj < n) ? A[kk + lda * cj] : 0.0;
      }
    }
  };
  auto store_tile = [&]() {
#pragma unroll
    for (int q = 0; q < A_PER_T; ++q) {
      if (INTV) {
        double mm, rr;
        vcphip_midrad_ru(im[q], ir[q], mm, rr);
        sIm[l_k][l_c + 16 * q] = mm;
        sIr[l_k][l_c + 16 * q] = rr;
        vcphip_midrad_ru(jm[q], jr[q], mm, rr);
        sJm[l_k][l_c + 16 * q] = mm;
        sJr[l_k][l_c + 16 * q] = rr;
      } else {
        sIm[l_k][l_c + 16 * q] = im[q];
        sJm[l_k][l_c + 16 * q] = jm[q];
      }
    }
  };

  const int tx = tid % 16, ty = tid / 16;
  double up[TM][TN], dn[TM][TN], tij[TM][TN], tji[TM][TN], sij[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) { up[i][j] = 0.0; dn[i][j] = 0.0; tij[i][j] = 0.0; tji[i][j] = 0.0; sij[i][j] = 0.0; }

  const int ntiles = (k + BK - 1) / BK;
  load_tile(0);
  store_tile();
  __syncthreads();
  for (int t = 0; t < ntiles; ++t) {
    if (t + 1 < ntiles) load_tile((t + 1) * BK);
#pragma unroll(INTV ? 1 : 2)
    for (int kk = 0; kk < BK; ++kk) {
      double a[TM], ar[TM], bb[TN], br[TN];
      {
        const double2 x0 = *reinterpret_cast<const double2*>(&sIm[kk][ty * TM]);
        const double2 x1 = *reinterpret_cast<const double2*>(&sIm[kk][ty * TM + 2]);
        a[0] = x0.x; a[1] = x0.y; a[2] = x1.x; a[3] = x1.y;
      }
      if (INTV) {
        const double2 x0 = *reinterpret_cast<const double2*>(&sIr[kk][ty * TM]);
        const double2 x1 = *reinterpret_cast<const double2*>(&sIr[kk][ty * TM + 2]);
        ar[0] = x0.x; ar[1] = x0.y; ar[2] = x1.x; ar[3] = x1.y;
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) { bb[j] = sJm[kk][tx + 16 * j]; if (INTV) br[j] = sJr[kk][tx + 16 * j]; }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          up[i][j] = __builtin_fma(a[i], bb[j], up[i][j]);
          dn[i][j] = __builtin_fma(-a[i], bb[j], dn[i][j]);
          if (INTV) {
            tij[i][j] = __builtin_fma(__builtin_fabs(a[i]), br[j], tij[i][j]);
            tji[i][j] = __builtin_fma(__builtin_fabs(bb[j]), ar[i], tji[i][j]);
            sij[i][j] = __builtin_fma(ar[i], br[j], sij[i][j]);
          }
        }
    }
    __syncthreads();
    if (t + 1 < ntiles) {
      store_tile();
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = i0 + ty * TM + i, c = j0 + tx + 16 * j;
      if (r >= n || c >= n || r > c) continue;
      double lo, hi;
      if (INTV) {
        const double T = tij[i][j] + (tji[i][j] + sij[i][j]);
        hi = up[i][j] + T;
        lo = -(dn[i][j] + T);
      } else {
        hi = up[i][j];
        lo = -dn[i][j];
      }
      *reinterpret_cast<double2*>(C + 2 * (r + ldc * c)) = make_double2(lo, hi);
      *reinterpret_cast<double2*>(C + 2 * (c + ldc * r)) = make_double2(lo, hi);
    }
}

// The same A^T A forms from the planes of A (k x n) made by k_planes_bt:
// mid(A), rad(A) (or A) row-major kpad x npad, zero-padded.  Column i of A is
// then the contiguous run at i of every k-row, so one plane set feeds both
// operands: wave 0..3 load the slices of the row block (mid, rad) and of the
// column block (mid, rad) straight into LDS, two stages.  The chains and
// their order are k_drsyrk's.
template <bool INTV>
__global__ __launch_bounds__(NT, 2) void k_drsyrk_planes(int n, int npad, int kpad, const double* __restrict__ P0,
                                                         const double* __restrict__ P1, double* __restrict__ C,
                                                         long ldc) {
  constexpr int NP = INTV ? 4 : 2;
  vcphip_mode_round_up();
  // [stage][plane: I mid, (I rad), J mid, (J rad)][k][64]
  __shared__ __attribute__((aligned(16))) double sL[2][NP][BK][BM];
  const int tid = threadIdx.x, lane = tid & 63;
  const int nt = npad / BM;
  int b = blockIdx.x, tm = 0;
  while (b >= nt - tm) {
    b -= nt - tm;
    ++tm;
  }
  const int tn = tm + b;
  const int i0 = tm * BM, j0 = tn * BN;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool loader = w < NP;
  // plane of wave w: I mid, I rad, J mid, J rad (interval); I, J (point)
  const bool rad = INTV && (w & 1);
  const bool jside = INTV ? (w >= 2) : (w == 1);
  const double* src = (rad ? P1 : P0) + (long)(lane >> 5) * npad + (jside ? j0 : i0) + 2 * (lane & 31);
  const uint32_t lds0 = vcphip_lds_addr(&sL[0][0][0][0]) + (uint32_t)(loader ? w : 0) * (BK * BM * 8);
  auto issue = [&](int stage, int t) {
    if (!loader) return;
    const double* g = src + (long)t * BK * npad;
    const uint32_t d = lds0 + (uint32_t)stage * (NP * BK * BM * 8);
#pragma unroll
    for (int q = 0; q < BK / 2; ++q) vcphip_lds_dma16(g + (long)(2 * q) * npad, d + (uint32_t)q * (2 * BM * 8));
  };

  const int tx = tid % 16, ty = tid / 16;
  double up[TM][TN], dn[TM][TN], tij[TM][TN], tji[TM][TN], sij[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      up[i][j] = 0.0;
      dn[i][j] = 0.0;
      tij[i][j] = 0.0;
      tji[i][j] = 0.0;
      sij[i][j] = 0.0;
    }

  const int ns = kpad / BK;
  issue(0, 0);
  for (int t = 0; t < ns; ++t) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (t + 1 < ns) issue((t + 1) & 1, t + 1);
    const double(*sIm)[BM] = sL[t & 1][0];
    const double(*sIr)[BM] = sL[t & 1][INTV ? 1 : 0];
    const double(*sJm)[BM] = sL[t & 1][INTV ? 2 : 1];
    const double(*sJr)[BM] = sL[t & 1][INTV ? 3 : 1];
#pragma unroll(INTV ? 1 : 2)
    for (int kk = 0; kk < BK; ++kk) {
      double a[TM], ar[TM], bb[TN], br[TN];
      {
        const double2 x0 = *reinterpret_cast<const double2*>(&sIm[kk][ty * TM]);
        const double2 x1 = *reinterpret_cast<const double2*>(&sIm[kk][ty * TM + 2]);
        a[0] = x0.x; a[1] = x0.y; a[2] = x1.x; a[3] = x1.y;
        const double2 y0 = *reinterpret_cast<const double2*>(&sJm[kk][2 * tx]);
        const double2 y1 = *reinterpret_cast<const double2*>(&sJm[kk][2 * tx + 32]);
        bb[0] = y0.x; bb[1] = y0.y; bb[2] = y1.x; bb[3] = y1.y;
      }
      if (INTV) {
        const double2 x0 = *reinterpret_cast<const double2*>(&sIr[kk][ty * TM]);
        const double2 x1 = *reinterpret_cast<const double2*>(&sIr[kk][ty * TM + 2]);
        ar[0] = x0.x; ar[1] = x0.y; ar[2] = x1.x; ar[3] = x1.y;
        const double2 y0 = *reinterpret_cast<const double2*>(&sJr[kk][2 * tx]);
        const double2 y1 = *reinterpret_cast<const double2*>(&sJr[kk][2 * tx + 32]);
        br[0] = y0.x; br[1] = y0.y; br[2] = y1.x; br[3] = y1.y;
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          up[i][j] = __builtin_fma(a[i], bb[j], up[i][j]);
          dn[i][j] = __builtin_fma(-a[i], bb[j], dn[i][j]);
          if (INTV) {
            tij[i][j] = __builtin_fma(__builtin_fabs(a[i]), br[j], tij[i][j]);
            tji[i][j] = __builtin_fma(__builtin_fabs(bb[j]), ar[i], tji[i][j]);
            sij[i][j] = __builtin_fma(ar[i], br[j], sij[i][j]);
          }
        }
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = i0 + ty * TM + i, c = j0 + dr_col(tx, j);
      if (r >= n || c >= n || r > c) continue;
      double lo, hi;
      if (INTV) {
        const double T = tij[i][j] + (tji[i][j] + sij[i][j]);
        hi = up[i][j] + T;
        lo = -(dn[i][j] + T);
      } else {
        hi = up[i][j];
        lo = -dn[i][j];
      }
      *reinterpret_cast<double2*>(C + 2 * (r + ldc * c)) = make_double2(lo, hi);
      *reinterpret_cast<double2*>(C + 2 * (c + ldc * r)) = make_double2(lo, hi);
    }
}

// ---------------------------------------------------------------------------
// Inf-sup interval product (pidblas_fma::mulmm, vcp/pidblas_fma.hpp:62-71,
// with the interval FMA of vcp/ifma.hpp:44-55): per output element the
// k-ascending chains
//   up <- max over a in {inf A, sup A}, b in {inf B, sup B} of RU(a b + up)
//   dn <- min over the same of RD(a b + dn),        both from 0.
// RU is monotone, so the maximum of the four rounded FMAs is the rounded FMA
// of the largest product: the same result as the reference's sign-case
// selection (ifma.hpp:58-137), without its branches.  The RD chain is carried
// negated, d = -dn, as d <- max RU((-a) b + d).
//
// The maxima are v_max_f64, whose order is total here (-0 < +0, a NaN
// operand yields the other), so the maximum of the four does not depend on
// which end of an operand comes first: the operands are staged as (end of
// smaller magnitude, end of larger magnitude).  When every operand of a
// k-slice is finite and strictly on one side of zero, the extreme products
// are those of the smaller ends and of the larger ends, so two FMAs per chain
// give the same maximum: 4 FMAs and 2 maxima per (i, j, k) instead of 8 and 6.
// ---------------------------------------------------------------------------
// finite and strictly on one side of zero
__device__ __forceinline__ bool is_sign_definite(double lo, double hi) {
  return (lo > 0.0 || hi < 0.0) && __builtin_fabs(lo) < __builtin_inf() && __builtin_fabs(hi) < __builtin_inf();
}
// (end of smaller magnitude, end of larger magnitude) of a sign-definite
// interval; other intervals keep their order
__device__ __forceinline__ double2 ends_by_magnitude(double2 v) {
  return v.y < 0.0 ? make_double2(v.y, v.x) : v;
}

__global__ __launch_bounds__(NT) void k_isgemm(int m, int n, int k, const double* __restrict__ A, long lda,
                                               const double* __restrict__ B, long ldb, double* __restrict__ C,
                                               long ldc) {
  vcphip_mode_round_up();
  __shared__ __attribute__((aligned(16))) double sAi[BK][LA];
  __shared__ __attribute__((aligned(16))) double sAs[BK][LA];
  __shared__ __attribute__((aligned(16))) double sBi[BK][LB];
  __shared__ __attribute__((aligned(16))) double sBs[BK][LB];
  const int ntm = (m + BM - 1) / BM, ntn = (n + BN - 1) / BN;
  const int tiles = ntm * ntn;
  int tm, tn;
  vcphip_tile_coords(vcphip_xcd_remap(blockIdx.x, tiles), ntm, ntn, 8, tm, tn);
  const int row0 = tm * BM, col0 = tn * BN;
  const int tid = threadIdx.x;
  const int a_row = tid % BM, a_kc = tid / BM;
  const int b_k = tid % BK, b_col = tid / BK;
  double2 ra[A_PER_T], rb[B_PER_T];
  auto load_tile = [&](int k0) {
#pragma unroll
    for (int q = 0; q < A_PER_T; ++q) {
      const int kk = k0 + a_kc + 4 * q, r = row0 + a_row;
      ra[q] = (r < m && kk < k) ? *reinterpret_cast<const double2*>(A + 2 * (r + lda * kk)) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < B_PER_T; ++q) {
      const int kk = k0 + b_k, c = col0 + b_col + 16 * q;
      rb[q] = (kk < k && c < n) ? *reinterpret_cast<const double2*>(B + 2 * (kk + ldb * c)) : make_double2(0.0, 0.0);
    }
  };
  // stores the slice at k0 (ends by magnitude) and returns whether every
  // operand of this thread's share is finite and strictly sign-definite
  // (the zero padding past m, n and k takes either loop)
  auto store_tile = [&](int k0) {
    bool ok = true;
#pragma unroll
    for (int q = 0; q < A_PER_T; ++q) {
      const int kk = k0 + a_kc + 4 * q, r = row0 + a_row;
      if (r < m && kk < k) ok = ok && is_sign_definite(ra[q].x, ra[q].y);
      const double2 v = ends_by_magnitude(ra[q]);
      sAi[a_kc + 4 * q][a_row] = v.x;
      sAs[a_kc + 4 * q][a_row] = v.y;
    }
#pragma unroll
    for (int q = 0; q < B_PER_T; ++q) {
      const int kk = k0 + b_k, c = col0 + b_col + 16 * q;
      if (kk < k && c < n) ok = ok && is_sign_definite(rb[q].x, rb[q].y);
      const double2 v = ends_by_magnitude(rb[q]);
      sBi[b_k][b_col + 16 * q] = v.x;
      sBs[b_k][b_col + 16 * q] = v.y;
    }
    return ok;
  };
  const int tx = tid % 16, ty = tid / 16;
  double up[TM][TN], dn[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      up[i][j] = 0.0;
      dn[i][j] = -0.0;   // -(+0): the RD chain starts from +0
    }
  const int ntiles = (k + BK - 1) / BK;
  load_tile(0);
  bool two = __syncthreads_and(store_tile(0)) != 0;   // the slice takes the two-FMA loop
  for (int t = 0; t < ntiles; ++t) {
    if (t + 1 < ntiles) load_tile((t + 1) * BK);
    if (two) {
      // every operand of the slice is sign-definite: the smaller ends' and
      // the larger ends' products are the extreme ones (two FMAs per chain;
      // a separate loop, since one loop with the mixed products behind a
      // uniform branch measured 25% slower)
      for (int kk = 0; kk < BK; ++kk) {
        double as[TM], ag[TM], bs[TN], bg[TN];   // ends of smaller / larger magnitude
        {
          const double2 x0 = *reinterpret_cast<const double2*>(&sAi[kk][ty * TM]);
          const double2 x1 = *reinterpret_cast<const double2*>(&sAi[kk][ty * TM + 2]);
          const double2 x2 = *reinterpret_cast<const double2*>(&sAs[kk][ty * TM]);
          const double2 x3 = *reinterpret_cast<const double2*>(&sAs[kk][ty * TM + 2]);
          as[0] = x0.x; as[1] = x0.y; as[2] = x1.x; as[3] = x1.y;
          ag[0] = x2.x; ag[1] = x2.y; ag[2] = x3.x; ag[3] = x3.y;
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            bs[j] = sBi[kk][tx + 16 * j];
            bg[j] = sBs[kk][tx + 16 * j];
          }
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            const double u = up[i][j], d = dn[i][j];
            up[i][j] = __builtin_fmax(__builtin_fma(as[i], bs[j], u), __builtin_fma(ag[i], bg[j], u));
            dn[i][j] = __builtin_fmax(__builtin_fma(-as[i], bs[j], d), __builtin_fma(-ag[i], bg[j], d));
          }
      }
    } else {
#pragma unroll 2
      for (int kk = 0; kk < BK; ++kk) {
        double ai[TM], as[TM], bi[TN], bs[TN];
        {
          const double2 x0 = *reinterpret_cast<const double2*>(&sAi[kk][ty * TM]);
          const double2 x1 = *reinterpret_cast<const double2*>(&sAi[kk][ty * TM + 2]);
          const double2 x2 = *reinterpret_cast<const double2*>(&sAs[kk][ty * TM]);
          const double2 x3 = *reinterpret_cast<const double2*>(&sAs[kk][ty * TM + 2]);
          ai[0] = x0.x; ai[1] = x0.y; ai[2] = x1.x; ai[3] = x1.y;
          as[0] = x2.x; as[1] = x2.y; as[2] = x3.x; as[3] = x3.y;
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            bi[j] = sBi[kk][tx + 16 * j];
            bs[j] = sBs[kk][tx + 16 * j];
          }
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            const double u = up[i][j], d = dn[i][j];
            up[i][j] = __builtin_fmax(__builtin_fmax(__builtin_fma(as[i], bs[j], u), __builtin_fma(ai[i], bs[j], u)),
                                      __builtin_fmax(__builtin_fma(as[i], bi[j], u), __builtin_fma(ai[i], bi[j], u)));
            dn[i][j] = __builtin_fmax(__builtin_fmax(__builtin_fma(-as[i], bs[j], d), __builtin_fma(-ai[i], bs[j], d)),
                                      __builtin_fmax(__builtin_fma(-as[i], bi[j], d), __builtin_fma(-ai[i], bi[j], d)));
          }
      }
    }
    __syncthreads();
    if (t + 1 < ntiles) two = __syncthreads_and(store_tile((t + 1) * BK)) != 0;
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int c = col0 + tx + 16 * j;
    if (c >= n) continue;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int r = row0 + ty * TM + i;
      if (r < m) *reinterpret_cast<double2*>(C + 2 * (r + ldc * (long)c)) = make_double2(-dn[i][j], up[i][j]);
    }
  }
}

// p[0..n) = v (a kernel, so that it can precede graph replays without a
// memset node)
__global__ __launch_bounds__(256) void k_fill_f64(double* __restrict__ p, long n, double v) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

int vcphip_launch_isgemm(int m, int n, int k, const double* A, long lda, const double* B, long ldb, double* C,
                         long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  const int tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
  hipLaunchKernelGGL(k_isgemm, dim3(tiles), dim3(NT), 0, s, m, n, k, A, lda, B, ldb, C, ldc);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_drgemm(int kind, int m, int n, int k, const double* A, long lda, const double* B,
                         long ldb, double* C, long ldc, double* C2, int accumulate, hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  const int ntiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
  dim3 grid(ntiles), block(NT);
  switch (kind) {
    case VCPHIP_DR_II: hipLaunchKernelGGL((k_drgemm<true, true>), grid, block, 0, s, m, n, k, A, lda, B, ldb, C, ldc, C2, accumulate); break;
    case VCPHIP_DR_IP: hipLaunchKernelGGL((k_drgemm<true, false>), grid, block, 0, s, m, n, k, A, lda, B, ldb, C, ldc, C2, accumulate); break;
    case VCPHIP_DR_PI: hipLaunchKernelGGL((k_drgemm<false, true>), grid, block, 0, s, m, n, k, A, lda, B, ldb, C, ldc, C2, accumulate); break;
    case VCPHIP_DR_PP: hipLaunchKernelGGL((k_drgemm<false, false>), grid, block, 0, s, m, n, k, A, lda, B, ldb, C, ldc, C2, accumulate); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_planes(int kind, int m, int n, int k, const double* A, long lda, const double* B, long ldb,
                         double* A0, double* A1, double* B0, double* B1, hipStream_t s) {
  const int mpad = (m + BM - 1) / BM * BM, npad = (n + BN - 1) / BN * BN, kpad = (k + BK - 1) / BK * BK;
  if (m <= 0 || n <= 0 || kpad <= 0) return 0;
  const bool ai = kind == VCPHIP_DR_II || kind == VCPHIP_DR_IP, bi = kind == VCPHIP_DR_II || kind == VCPHIP_DR_PI;
  const dim3 ga(mpad / 256 + (mpad % 256 != 0), kpad), gb(kpad / 32 + (kpad % 32 != 0), npad / 32);
  if (ai) hipLaunchKernelGGL(k_planes_a<true>, ga, dim3(256), 0, s, m, k, mpad, A, lda, A0, A1);
  else hipLaunchKernelGGL(k_planes_a<false>, ga, dim3(256), 0, s, m, k, mpad, A, lda, A0, A1);
  if (bi) hipLaunchKernelGGL(k_planes_bt<true>, gb, dim3(256), 0, s, k, n, kpad, npad, B, ldb, B0, B1);
  else hipLaunchKernelGGL(k_planes_bt<false>, gb, dim3(256), 0, s, k, n, kpad, npad, B, ldb, B0, B1);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_drgemm_planes(int kind, int m, int n, int k, const double* A0, const double* A1, const double* B0,
                                const double* B1, double* C, long ldc, double* C2, int accumulate, int rows,
                                hipStream_t s) {
  const int mpad = (m + BM - 1) / BM * BM, npad = (n + BN - 1) / BN * BN, kpad = (k + BK - 1) / BK * BK;
  if (m <= 0 || n <= 0) return 0;
  if (kpad == 0 || (rows != 64 && rows != 32)) return -1;   // k = 0 takes the tile path
  const dim3 grid((mpad / rows) * (npad / BN)), block(NT);
#define VCPHIP_PLANES_CASE(KIND, AI, BI)                                                                          \
  case KIND:                                                                                                     \
    if (rows == 64)                                                                                              \
      hipLaunchKernelGGL((k_drgemm_planes<AI, BI, 4>), grid, block, 0, s, m, n, mpad, npad, kpad, A0, A1, B0, B1, \
                         C, ldc, C2, accumulate);                                                                \
    else                                                                                                         \
      hipLaunchKernelGGL((k_drgemm_planes<AI, BI, 2>), grid, block, 0, s, m, n, mpad, npad, kpad, A0, A1, B0, B1, \
                         C, ldc, C2, accumulate);                                                                \
    break;
  switch (kind) {
    VCPHIP_PLANES_CASE(VCPHIP_DR_II, true, true)
    VCPHIP_PLANES_CASE(VCPHIP_DR_IP, true, false)
    VCPHIP_PLANES_CASE(VCPHIP_DR_PI, false, true)
    VCPHIP_PLANES_CASE(VCPHIP_DR_PP, false, false)
    default: return -1;
  }
#undef VCPHIP_PLANES_CASE
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_drsyrk_planes(int interval, int k, int n, const double* A, long lda, double* P0, double* P1,
                                double* C, long ldc, hipStream_t s) {
  if (n <= 0) return 0;
  const int npad = (n + BN - 1) / BN * BN, kpad = (k + BK - 1) / BK * BK;
  if (kpad == 0) return -1;   // k = 0 takes the tile path
  const dim3 gb(kpad / 32 + (kpad % 32 != 0), npad / 32);
  const int nt = npad / BN;
  if (interval) {
    hipLaunchKernelGGL(k_planes_bt<true>, gb, dim3(256), 0, s, k, n, kpad, npad, A, lda, P0, P1);
    hipLaunchKernelGGL(k_drsyrk_planes<true>, dim3(nt * (nt + 1) / 2), dim3(NT), 0, s, n, npad, kpad, P0, P1, C, ldc);
  } else {
    hipLaunchKernelGGL(k_planes_bt<false>, gb, dim3(256), 0, s, k, n, kpad, npad, A, lda, P0, P1);
    hipLaunchKernelGGL(k_drsyrk_planes<false>, dim3(nt * (nt + 1) / 2), dim3(NT), 0, s, n, npad, kpad, P0, P1, C, ldc);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_drsyrk(int interval, int k, int n, const double* A, long lda, double* C, long ldc,
                         hipStream_t s) {
  if (n <= 0) return 0;
  const int nt = (n + BM - 1) / BM;
  dim3 grid(nt * (nt + 1) / 2), block(NT);
  if (interval) hipLaunchKernelGGL((k_drsyrk<true>), grid, block, 0, s, k, n, A, lda, C, ldc);
  else hipLaunchKernelGGL((k_drsyrk<false>), grid, block, 0, s, k, n, A, lda, C, ldc);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_igemm_tiles(int n64, int n32, const VcphipResidTile* tiles, const VcphipIgemmProblem* probs,
                              int grid64, int grid32, unsigned* queue, hipStream_t s) {
  if (n64 > 0 || n32 > 0) hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, queue);
  if (n64 > 0) hipLaunchKernelGGL(k_igemm_tiles<4>, dim3(grid64), dim3(NT), 0, s, n64, tiles, probs, queue);
  if (n32 > 0) hipLaunchKernelGGL(k_igemm_tiles<2>, dim3(grid32), dim3(NT), 0, s, n32, tiles + n64, probs, queue + 1);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_fill_f64(double* p, long n, double v, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_resid_tiles(int n64, int n32, const VcphipResidTile* tiles, const VcphipResidProblem* probs,
                              int grid64, int grid32, unsigned* queue, hipStream_t s) {
  if (n64 > 0 || n32 > 0) hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, queue);
  if (n64 > 0) hipLaunchKernelGGL(k_resid_tiles<4>, dim3(grid64), dim3(NT), 0, s, n64, tiles, probs, queue);
  if (n32 > 0) hipLaunchKernelGGL(k_resid_tiles<2>, dim3(grid32), dim3(NT), 0, s, n32, tiles + n64, probs, queue + 1);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int vcphip_launch_resid(int n64, int n32, const VcphipResidTile* tiles, const VcphipResidProblem* probs, int nrows,
                        const int* row_problem, double* rowsum, int count, double* normub, int grid64, int grid32,
                        unsigned* queue, hipStream_t s) {
  vcphip_launch_resid_tiles(n64, n32, tiles, probs, grid64, grid32, queue, s);
  if (nrows > 0) hipLaunchKernelGGL(k_resid_rowsum, dim3((nrows + 255) / 256), dim3(256), 0, s, nrows, row_problem, probs, rowsum);
  if (count > 0) hipLaunchKernelGGL(k_resid_max, dim3(count), dim3(256), 0, s, count, probs, rowsum, normub);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
