// fuel's math::pairwise_sqdist (src/math/mod.rs:562-579), the distance matrix
// of the k-means family (lloyd.rs:76, elkan.rs:83, hamerly.rs:89, selkan.rs:73,
// shamerly.rs:74): out[i][j] = sum over k of (p1[i][k] - p2[j][k])^2 in the
// reference's own rounding order, which differs from the search kernel:
//   d >= 4 (src/math/avx2.rs:888-963, :1036-1082): one sequential FMA chain
//          per output, acc = fma(diff, diff, acc) for k = 0, 1, ..., d-1 from +0;
//   d <  4 (src/math/scalar.rs:10-16, :133-163): the rounded squares added
//          left to right.
// Tiling: a workgroup computes a 64 x 64 block of outputs; k runs in chunks of
// KC staged through LDS (rows of both point sets, one padded row per point);
// each thread owns 4 x 4 outputs and walks k in order, so the per-output chain
// is exactly the reference's.
#include "common.h"
#include "kernels.h"

namespace fa {

template <typename F, bool FUSED>
__global__ __launch_bounds__(256) void kmeans_pairwise_kernel(KmeansPairwiseArgs<F> a) {
  constexpr int TM = 64, TN = 64, KC = 32, PAD = 1;
  __shared__ F sa[TM][KC + PAD];
  __shared__ F sb[TN][KC + PAD];
  const int tid = threadIdx.x;
  const int64_t i0 = int64_t(blockIdx.y) * TM;
  const int64_t j0 = int64_t(blockIdx.x) * TN;
  const int ti = (tid / 16) * 4;  // 4 rows of the block
  const int tj = (tid % 16) * 4;  // 4 columns of the block
  F acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = F(0);
  for (int k0 = 0; k0 < a.d; k0 += KC) {
    const int kc = (a.d - k0) < KC ? (a.d - k0) : KC;
    // stage 64 rows x kc of each point set (clamped rows are never written out)
    for (int e = tid; e < TM * KC; e += 256) {
      const int r = e / KC, c = e % KC;
      if (c < kc) {
        const int64_t ri = i0 + r < a.m ? i0 + r : a.m - 1;
        const int64_t rj = j0 + r < a.n ? j0 + r : a.n - 1;
        sa[r][c] = a.P1[ri * a.ld1 + k0 + c];
        sb[r][c] = a.P2[rj * a.ld2 + k0 + c];
      }
    }
    __syncthreads();
    for (int c = 0; c < kc; ++c) {
      F x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = sa[ti + u][c];
#pragma unroll
      for (int v = 0; v < 4; ++v) y[v] = sb[tj + v][c];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const F t = x[u] - y[v];
          if constexpr (FUSED)
            acc[u][v] = fma_rn(t, t, acc[u][v]);
          else
            acc[u][v] = acc[u][v] + t * t;  // no contraction: -ffp-contract=off
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t i = i0 + ti + u;
    if (i >= a.m) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t j = j0 + tj + v;
      if (j < a.n) a.out[i * a.ldo + j] = acc[u][v];
    }
  }
}

template <typename F>
hipError_t launch_kmeans_pairwise(const KmeansPairwiseArgs<F>& a, hipStream_t st) {
  if (a.m == 0 || a.n == 0) return hipSuccess;
  const dim3 grid((unsigned)((a.n + 63) / 64), (unsigned)((a.m + 63) / 64)), block(256);
  if (a.d >= 4)
    hipLaunchKernelGGL((kmeans_pairwise_kernel<F, true>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((kmeans_pairwise_kernel<F, false>), grid, block, 0, st, a);
  return hipGetLastError();
}

template hipError_t launch_kmeans_pairwise<float>(const KmeansPairwiseArgs<float>&, hipStream_t);
template hipError_t launch_kmeans_pairwise<double>(const KmeansPairwiseArgs<double>&, hipStream_t);

}  // namespace fa
