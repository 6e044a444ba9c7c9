// One step of the exact (restated-distance) computation shared by the exact
// kNN scan, range search and pairwise kernels.
//
// A workgroup of 256 lanes holds QW "query" rows in LDS (read as broadcasts)
// and gives each lane one "data" row, streamed from HBM in 16-byte loads; every
// loaded block of LANES coordinates is reused across the QW queries from
// registers.  When d exceeds the LDS chunk DC the queries are staged chunk by
// chunk (with workgroup barriers, so every lane must call exact_step).
#pragma once
#include "common.h"

namespace fa {

template <typename F, int QW>
struct ExactCfg {
  static constexpr int L = Lanes<F>::N;
  static constexpr int DC = (QW * 512 * (int)sizeof(F) <= 32768) ? 512 : ((32768 / (QW * (int)sizeof(F))) & ~7);
};

// Stage coordinates [c0, c1) of the QW query rows; rows >= nqb are zero.
template <typename F, int QW, int DC, typename QRow>
__device__ __forceinline__ void stage_queries(F (*qbuf)[DC], const QRow& qrow, int nqb, int c0, int c1) {
  const int w = c1 - c0;
  for (int e = threadIdx.x; e < QW * w; e += blockDim.x) {
    const int qq = e / w, c = e - qq * w;
    qbuf[qq][c] = qq < nqb ? qrow(qq)[c0 + c] : F(0);
  }
}

// Lanes and rounding of the restated distance.  ARITH_AVX: fuel's AVX+FMA
// kernels (src/distance/squared_euclidean.rs:13-67), LANES = 8 (f32) or 4 (f64)
// fused lanes, a left fold from -0.0 (Rust's Sum), a fused tail.  ARITH_FALLBACK:
// the kernel of hosts without AVX+FMA (:99-133), 8 lanes of fl(t * t) + acc for
// both dtypes, a left fold from +0, an unfused tail.
template <typename F, int ARITH>
struct Arith {
  static constexpr int L = ARITH == ARITH_FALLBACK ? 8 : Lanes<F>::N;
  __device__ static __forceinline__ F step(F t, F acc) {
    if constexpr (ARITH == ARITH_FALLBACK) return t * t + acc;  // two roundings (-ffp-contract=off)
    else return fma_rn(t, t, acc);
  }
  __device__ static __forceinline__ F fold_start() { return ARITH == ARITH_FALLBACK ? F(0) : F(-0.0); }
};

// s[qq] = restated squared distance between query qq and the lane's row xr.
// `single` means all d coordinates are already staged (d <= DC, a multiple of 8).
template <typename F, int QW, int DC, int ARITH = ARITH_AVX, typename QRow>
__device__ __forceinline__ void exact_step(F (*qbuf)[DC], const QRow& qrow, int nqb, bool single,
                                           const F* __restrict__ xr, int d, F (&s)[QW]) {
  using AR = Arith<F, ARITH>;
  constexpr int L = AR::L;
  static_assert(DC % L == 0, "lane groups do not straddle staged chunks");
  const int sd = d & ~(L - 1);
  F acc[QW][L];
#pragma unroll
  for (int qq = 0; qq < QW; ++qq)
#pragma unroll
    for (int j = 0; j < L; ++j) acc[qq][j] = F(0);

  for (int c0 = 0; c0 < d; c0 += DC) {
    const int c1 = c0 + DC < d ? c0 + DC : d;
    if (!single) {
      __syncthreads();
      stage_queries<F, QW, DC>(qbuf, qrow, nqb, c0, c1);
      __syncthreads();
    }
    const int iend = c1 < sd ? c1 : sd;
    for (int i = c0; i < iend; i += L) {
      F xv[L];
      if constexpr (sizeof(F) == 4) {
        const float4* p = reinterpret_cast<const float4*>(xr + i);
        const float4 u0 = p[0], u1 = p[1];
        xv[0] = u0.x; xv[1] = u0.y; xv[2] = u0.z; xv[3] = u0.w;
        xv[4] = u1.x; xv[5] = u1.y; xv[6] = u1.z; xv[7] = u1.w;
      } else {
        const double2* p = reinterpret_cast<const double2*>(xr + i);
#pragma unroll
        for (int h = 0; h < L / 2; ++h) {
          const double2 u = p[h];
          xv[2 * h] = u.x;
          xv[2 * h + 1] = u.y;
        }
      }
#pragma unroll
      for (int qq = 0; qq < QW; ++qq) {
#pragma unroll
        for (int j = 0; j < L; ++j) {
          const F t = qbuf[qq][i - c0 + j] - xv[j];
          acc[qq][j] = AR::step(t, acc[qq][j]);
        }
      }
    }
  }
  // Left-fold lane sum, then the tail over d mod L.  The tail coordinates are
  // in the last staged chunk.
  const int c0_last = d > 0 ? ((d - 1) / DC) * DC : 0;
#pragma unroll
  for (int qq = 0; qq < QW; ++qq) {
    F t = AR::fold_start();
#pragma unroll
    for (int j = 0; j < L; ++j) t = t + acc[qq][j];
    s[qq] = t;
  }
  for (int i = sd; i < d; ++i) {
    const F xv = xr[i];
#pragma unroll
    for (int qq = 0; qq < QW; ++qq) {
      const F t = qbuf[qq][i - c0_last] - xv;
      s[qq] = AR::step(t, s[qq]);
    }
  }
}

}  // namespace fa
