// Local outlier factor and kNN-outlier scores from a GPU self-kNN table
// (fuel src/outlier/lof.rs:14-81, src/outlier/knn.rs:14-32).  Each point's sums
// run sequentially in neighbour order in the data type, as Rust's `Sum` fold;
// divisions and sqrt are IEEE correctly rounded; nothing is contracted.
#include "kernels.h"

namespace fa {

template <typename F>
__global__ void kdist_kernel(LofArgs<F> a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  // k-distance = distance of the last neighbour (src/outlier/lof.rs:39-41)
  a.kdist[i] = a.k > 0 ? a.nb_dist[i * a.k + a.k - 1] : F(0);
}

template <typename F>
__global__ void lrd_kernel(LofArgs<F> a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (a.k == 0) {
    a.lrd[i] = F(INFINITY);
    return;
  }
  // lrd = |N| / sum max(kdist(o), d(p, o))  (src/outlier/lof.rs:48-62)
  F s = F(-0.0);
  for (int j = 0; j < a.k; ++j) {
    const int64_t o = a.nb_idx[i * a.k + j];
    s = s + fmax(a.kdist[o], a.nb_dist[i * a.k + j]);
  }
  a.lrd[i] = s > F(0) ? F(a.k) / s : F(INFINITY);
}

template <typename F>
__global__ void lof_score_kernel(LofArgs<F> a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const F l = a.lrd[i];
  // score = sum lrd(o) / (lrd(p) * |N|), 1 when |N| = 0 or lrd(p) is infinite
  // (src/outlier/lof.rs:64-78)
  if (a.k == 0 || isinf(l)) {
    a.out[i] = F(1);
    return;
  }
  F s = F(-0.0);
  for (int j = 0; j < a.k; ++j) s = s + a.lrd[a.nb_idx[i * a.k + j]];
  const F den = l * F(a.k);
  a.out[i] = s / den;
}

template <typename F>
__global__ void knn_outlier_kernel(LofArgs<F> a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  a.out[i] = a.k > 0 ? a.nb_dist[i * a.k + a.k - 1] : F(0);
}

template <typename F>
hipError_t launch_lof(const LofArgs<F>& a, int knn_outlier, hipStream_t st) {
  const dim3 grid((unsigned)((a.n + 255) / 256)), block(256);
  if (a.n == 0) return hipSuccess;
  if (knn_outlier) {
    hipLaunchKernelGGL((knn_outlier_kernel<F>), grid, block, 0, st, a);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((kdist_kernel<F>), grid, block, 0, st, a);
  hipLaunchKernelGGL((lrd_kernel<F>), grid, block, 0, st, a);
  hipLaunchKernelGGL((lof_score_kernel<F>), grid, block, 0, st, a);
  return hipGetLastError();
}

template hipError_t launch_lof<float>(const LofArgs<float>&, int, hipStream_t);
template hipError_t launch_lof<double>(const LofArgs<double>&, int, hipStream_t);

}  // namespace fa
