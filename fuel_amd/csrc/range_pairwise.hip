// Exact range search (two passes: count, fill; then a per-query sort) and
// pairwise distance matrices, both on the restated fuel distance.
//
// Range: fuel's linear_scan_range keeps every point with distance <= radius and
// sorts by (distance, index) (src/api/search.rs:154-166).  A per-query radius
// (the k-th distance of each query) gives KNNHeap's all-ties kNN result.  For the Euclidean
// metric the test sqrt(s) <= r is done as s <= S(r), S(r) the largest float whose
// correctly rounded sqrt is <= r, which is the same predicate.
// Pairwise: out[i][j] = restated(p_i, p_j); the restated distance is symmetric
// bit for bit, so computing both triangles equals fuel's mirrored upper triangle
// (src/api/square_distance_matrix.rs:12-42); condensed order i(i-1)/2+j, i>j
// (src/api/condensed_distance_matrix.rs:22-54).
#include "exact_core.h"
#include "kernels.h"

namespace fa {

template <typename F, int QW, int PASS, int ARITH>
__global__ __launch_bounds__(256) void range_kernel(RangeArgs<F> a) {
  constexpr int DC = ExactCfg<F, QW>::DC;
  __shared__ F qbuf[QW][DC];
  __shared__ int64_t run[QW];
  __shared__ int wcnt[QW][4];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int64_t qs0 = int64_t(blockIdx.x) * QW;
  if (qs0 >= a.nq) return;
  const int nqb = (int)((a.nq - qs0) < QW ? (a.nq - qs0) : QW);
  const int split = blockIdx.y;
  const int64_t r0 = int64_t(split) * a.rows_per_split;
  int64_t r1 = r0 + a.rows_per_split;
  if (r1 > a.n) r1 = a.n;
  const int d = a.d;
  if (tid < QW) run[tid] = (PASS == 1 && tid < nqb) ? a.offsets[(qs0 + tid) * a.nsplit + split] : 0;
  auto qrow = [&](int qq) -> const F* { return a.Q + (qs0 + qq) * a.ldq; };
  const bool single = d <= DC;
  if (single) stage_queries<F, QW, DC>(qbuf, qrow, nqb, 0, d);
  __syncthreads();
  F rad[QW], rs[QW], lo[QW];
#pragma unroll
  for (int qq = 0; qq < QW; ++qq) {
    rad[qq] = (a.radii && qq < nqb) ? a.radii[(qs0 + qq) * a.radii_stride] : a.radius;
    rs[qq] = a.metric == 0 ? sq_upper(rad[qq]) : rad[qq];
    lo[qq] = (a.skips && qq < nqb) ? a.skips[qs0 + qq] : F(-INFINITY);
  }
  for (int64_t rb = r0; rb < r1; rb += 256) {
    const int64_t row = rb + tid;
    const bool valid = row < r1;
    F s[QW];
    exact_step<F, QW, DC, ARITH>(qbuf, qrow, nqb, single, a.X + (valid ? row : r0) * a.ldx, d, s);
    uint64_t hits[QW];
#pragma unroll
    for (int qq = 0; qq < QW; ++qq) {
      bool hit = qq < nqb && valid && s[qq] <= rs[qq] && !(a.exclude_self && row == a.qbase + qs0 + qq);
      if (hit && a.metric == 0) hit = sqrt_rn(s[qq]) <= rad[qq];  // exact predicate
      if (hit && a.skips) hit = (a.metric == 0 ? sqrt_rn(s[qq]) : s[qq]) >= lo[qq];
      hits[qq] = __ballot(hit);
      if (__lane_id() == 0) wcnt[qq][wave] = __popcll(hits[qq]);
    }
    __syncthreads();
    if constexpr (PASS == 1) {
#pragma unroll
      for (int qq = 0; qq < QW; ++qq) {
        const uint64_t b = hits[qq];
        if (b & (1ull << __lane_id())) {
          int64_t pos = run[qq];
          for (int w = 0; w < wave; ++w) pos += wcnt[qq][w];
          pos += __popcll(b & ((1ull << __lane_id()) - 1));
          const F key = a.metric == 0 ? sqrt_rn(s[qq]) : s[qq];
          a.out_keys[pos] = ord_key(key);
          a.out_idx[pos] = row;
        }
      }
    }
    __syncthreads();
    if (tid < QW) run[tid] += wcnt[tid][0] + wcnt[tid][1] + wcnt[tid][2] + wcnt[tid][3];
    __syncthreads();
  }
  if constexpr (PASS == 0) {
    if (tid < nqb) a.counts[(qs0 + tid) * a.nsplit + split] = run[tid];
  }
}

template <typename F, int QW, int FORM, int ARITH>
__global__ __launch_bounds__(256) void pairwise_kernel(PairwiseArgs<F> a) {
  constexpr int DC = ExactCfg<F, QW>::DC;
  __shared__ F qbuf[QW][DC];
  const int tid = threadIdx.x;
  const int64_t qs0 = int64_t(blockIdx.x) * QW;
  if (qs0 >= a.n) return;
  const int nqb = (int)((a.n - qs0) < QW ? (a.n - qs0) : QW);
  const int64_t r0 = int64_t(blockIdx.y) * a.rows_per_split;
  int64_t r1 = r0 + a.rows_per_split;
  if (r1 > a.n) r1 = a.n;
  if (FORM == 1 && r1 > qs0 + nqb - 1) r1 = qs0 + nqb - 1;  // condensed: only j < i
  if (r0 >= r1) return;
  const int d = a.d;
  auto qrow = [&](int qq) -> const F* { return a.X + (qs0 + qq) * a.ldx; };
  const bool single = d <= DC;
  if (single) stage_queries<F, QW, DC>(qbuf, qrow, nqb, 0, d);
  __syncthreads();
  for (int64_t rb = r0; rb < r1; rb += 256) {
    const int64_t j = rb + tid;
    const bool valid = j < r1;
    F s[QW];
    exact_step<F, QW, DC, ARITH>(qbuf, qrow, nqb, single, a.X + (valid ? j : r0) * a.ldx, d, s);
#pragma unroll
    for (int qq = 0; qq < QW; ++qq) {
      const int64_t i = qs0 + qq;
      if (qq < nqb && valid) {
        const F v = a.metric == 0 ? sqrt_rn(s[qq]) : s[qq];
        if (FORM == 0) {
          a.out[i * a.n + j] = v;
        } else if (j < i) {
          a.out[i * (i - 1) / 2 + j] = v;
        }
      }
    }
  }
}

// Queries per workgroup: 16, and 8 for the f64 fallback form, whose 8 double
// lanes per query would not stay in registers for 16 (the grid follows).
template <typename F>
hipError_t launch_range(const RangeArgs<F>& a, int pass, int nsplit, hipStream_t st) {
  constexpr int QW = 16;
  constexpr int QW8 = 8;
  if (a.arith == ARITH_FALLBACK) {
    if constexpr (sizeof(F) == 8) {
      dim3 grid((unsigned)((a.nq + QW8 - 1) / QW8), (unsigned)nsplit);
      if (pass == 0)
        hipLaunchKernelGGL((range_kernel<F, QW8, 0, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((range_kernel<F, QW8, 1, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
    } else {
      dim3 grid((unsigned)((a.nq + QW - 1) / QW), (unsigned)nsplit);
      if (pass == 0)
        hipLaunchKernelGGL((range_kernel<F, QW, 0, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((range_kernel<F, QW, 1, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
    }
    return hipGetLastError();
  }
  dim3 grid((unsigned)((a.nq + QW - 1) / QW), (unsigned)nsplit);
  if (pass == 0)
    hipLaunchKernelGGL((range_kernel<F, QW, 0, ARITH_AVX>), grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((range_kernel<F, QW, 1, ARITH_AVX>), grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

template <typename F>
hipError_t launch_pairwise(const PairwiseArgs<F>& a, int form, int nsplit, hipStream_t st) {
  constexpr int QW = 16;
  constexpr int QW8 = 8;
  if (a.arith == ARITH_FALLBACK) {
    if constexpr (sizeof(F) == 8) {
      dim3 grid((unsigned)((a.n + QW8 - 1) / QW8), (unsigned)nsplit);
      if (form == 0)
        hipLaunchKernelGGL((pairwise_kernel<F, QW8, 0, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((pairwise_kernel<F, QW8, 1, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
    } else {
      dim3 grid((unsigned)((a.n + QW - 1) / QW), (unsigned)nsplit);
      if (form == 0)
        hipLaunchKernelGGL((pairwise_kernel<F, QW, 0, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((pairwise_kernel<F, QW, 1, ARITH_FALLBACK>), grid, dim3(256), 0, st, a);
    }
    return hipGetLastError();
  }
  dim3 grid((unsigned)((a.n + QW - 1) / QW), (unsigned)nsplit);
  if (form == 0)
    hipLaunchKernelGGL((pairwise_kernel<F, QW, 0, ARITH_AVX>), grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((pairwise_kernel<F, QW, 1, ARITH_AVX>), grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// Convert sorted order-preserving keys back to distances.
template <typename F>
__global__ void unord_kernel(const typename OrdKey<F>::T* __restrict__ keys, int64_t nnz, F* __restrict__ dist) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nnz) dist[i] = unord_key<F>(keys[i]);
}

// Per-query CSR offsets from the (query, split) exclusive scan.
__global__ void gather_offsets_kernel(const int64_t* __restrict__ qs_offsets, int64_t m, int nsplit,
                                      int64_t total, int64_t* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) out[i] = qs_offsets[i * nsplit];
  if (i == m) out[m] = total;
}

template <typename F>
hipError_t launch_unord(const typename OrdKey<F>::T* keys, int64_t nnz, F* dist, hipStream_t st) {
  if (nnz == 0) return hipSuccess;
  hipLaunchKernelGGL((unord_kernel<F>), dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, keys, nnz, dist);
  return hipGetLastError();
}

hipError_t launch_gather_offsets(const int64_t* qs_offsets, int64_t m, int nsplit, int64_t total, int64_t* out,
                                 hipStream_t st) {
  hipLaunchKernelGGL(gather_offsets_kernel, dim3((unsigned)((m + 1 + 255) / 256)), dim3(256), 0, st, qs_offsets,
                     m, nsplit, total, out);
  return hipGetLastError();
}

template hipError_t launch_range<float>(const RangeArgs<float>&, int, int, hipStream_t);
template hipError_t launch_range<double>(const RangeArgs<double>&, int, int, hipStream_t);
template hipError_t launch_pairwise<float>(const PairwiseArgs<float>&, int, int, hipStream_t);
template hipError_t launch_pairwise<double>(const PairwiseArgs<double>&, int, int, hipStream_t);
template hipError_t launch_unord<float>(const uint32_t*, int64_t, float*, hipStream_t);
template hipError_t launch_unord<double>(const uint64_t*, int64_t, double*, hipStream_t);

}  // namespace fa
