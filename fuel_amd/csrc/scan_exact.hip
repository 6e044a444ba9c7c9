// Exact linear scan: every (query, point) pair gets the restated fuel distance
// (bit-identical to the AVX2 kernel), and a fused per-query selection keeps the
// KP lexicographically smallest (distance, index) pairs per DB split without
// ever writing the distance matrix.  This is the parity anchor and the fallback
// for queries whose mixed-precision candidate set could not be certified.
//
// Work decomposition: one workgroup = 4 waves = 256 lanes, QW queries per
// workgroup (exact_core.h), one DB row per lane per step; grid.y splits the DB
// rows so that small query batches still fill the GPU.
// Selection: a per-query threshold in LDS filters candidates (in squared space
// for the Euclidean metric, so the correctly rounded sqrt runs only for
// survivors); survivors are appended to the query's slot in HBM and merged into
// the sorted KP-list by a wave-cooperative bitonic sort when the buffer fills.
#include "exact_core.h"
#include "kernels.h"

namespace fa {

template <typename F, int QW, int V, int ARITH>
__global__ __launch_bounds__(256) void scan_exact_kernel(ScanExactArgs<F> a) {
  using C = Comp<F>;
  using T = typename C::T;
  constexpr int NS = 64 * V;
  constexpr int DC = ExactCfg<F, QW>::DC;
  __shared__ F qbuf[QW][DC];
  __shared__ T thr[QW];
  __shared__ F thr_s[QW];
  __shared__ int bcnt[QW];
  __shared__ int lcnt[QW];
  __shared__ T lo_s[QW];

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
  const int KP = a.KP;
  const int CAP = NS - KP;
  T* slotbase = a.slots + (size_t(split) * a.nq + qs0) * NS;

  if (tid < QW) {
    thr[tid] = C::MAX;
    thr_s[tid] = F(INFINITY);
    bcnt[tid] = 0;
    lcnt[tid] = 0;
    lo_s[tid] = (a.lo && tid < nqb) ? a.lo[qs0 + tid] : T(0);
  }
  const bool has_lo = a.lo != nullptr;
  auto qrow = [&](int qq) -> const F* {
    const int64_t r = a.qrows ? a.qrows[qs0 + qq] : qs0 + qq;
    return a.Q + r * a.ldq;
  };
  const bool single = d <= DC;
  if (single) stage_queries<F, QW, DC>(qbuf, qrow, nqb, 0, d);
  __syncthreads();

  for (int64_t rb = r0; rb < r1; rb += 256) {
    const int64_t row = rb + tid;
    const bool valid = row < r1;
    F s[QW];
    exact_step<F, QW, DC, ARITH>(qbuf, qrow, nqb, single, a.X + (valid ? row : r0) * a.ldx, d, s);

#pragma unroll
    for (int qq = 0; qq < QW; ++qq) {
      if (qq < nqb && valid && s[qq] <= thr_s[qq]) {
        const F key = a.metric == 0 ? sqrt_rn(s[qq]) : s[qq];
        const T c = C::make(key, (uint32_t)row);
        if (c < thr[qq] && (!has_lo || c > lo_s[qq])) {
          const int pos = atomicAdd(&bcnt[qq], 1);
          slotbase[size_t(qq) * NS + KP + pos] = c;
        }
      }
    }
    __syncthreads();
    const bool last = rb + 256 >= r1;
    for (int qq = wave; qq < nqb; qq += 4) {
      const int b = bcnt[qq];
      if (b > 0 && (last || b > CAP - 256)) {
        T t;
        const int l = merge_slot<T, V>(slotbase + size_t(qq) * NS, KP, lcnt[qq], b, t);
        if (__lane_id() == 0) {
          lcnt[qq] = l;
          bcnt[qq] = 0;
          if (l == KP) {
            thr[qq] = t;
            const F kf = C::key(t);
            thr_s[qq] = a.metric == 0 ? sq_upper(kf) : kf;
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid < nqb) a.lcnts[size_t(split) * a.nq + qs0 + tid] = lcnt[tid];
}

// Merge the per-split lists of one query (one wave per query slot) into the
// lexicographic top-KK, then apply fuel's exclusion / truncation / padding
// semantics and write the caller's (index, distance) row.
template <typename F, int V>
__global__ __launch_bounds__(64) void finalize_exact_kernel(FinalizeArgs<F> a) {
  using C = Comp<F>;
  using T = typename C::T;
  constexpr int NS = 64 * V;
  const int64_t qs = blockIdx.x;
  if (qs >= a.nq) return;
  const int lane = __lane_id();
  const int KK = a.KK;
  T x[V];
#pragma unroll
  for (int v = 0; v < V; ++v) x[v] = C::MAX;
  int fill = 0;
  for (int s = 0; s < a.nsplit; ++s) {
    const T* slot = a.slots + (size_t(s) * a.nq + qs) * a.slot_stride;
    int cnt = a.lcnts[size_t(s) * a.nq + qs];
    if (cnt > KK) cnt = KK;
    int off = 0;
    while (off < cnt) {
      const int room = NS - fill;
      const int take = (cnt - off) < room ? (cnt - off) : room;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int e = v * 64 + lane;
        if (e >= fill && e < fill + take) x[v] = slot[off + e - fill];
      }
      fill += take;
      off += take;
      if (fill == NS) {
        wave_bitonic_sort<T, V>(x);
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (v * 64 + lane >= KK) x[v] = C::MAX;
        fill = KK;
      }
    }
  }
  wave_bitonic_sort<T, V>(x);
  const int count = fill < KK ? fill : KK;
  if (a.raw_out) {
    T* row = a.raw_out + qs * a.raw_stride + a.raw_off;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int e = v * 64 + lane;
      if (e < KK) row[e] = e < count ? x[v] : C::MAX;
    }
    const T last = count == KK ? wave_get<T, V>(x, KK - 1) : C::MAX;
    if (lane == 0) {
      a.lo_out[qs] = last;
      a.cnt_out[qs] += count;
    }
    return;
  }
  const int64_t qr = a.qrows ? a.qrows[qs] : qs;
  write_knn_row<F, V>(x, count, a.qbase + qr, a.k, a.mode, a.out_idx + qr * a.k, a.out_dist + qr * a.k);
}

// One thread per query slot: the concatenated rounds are sorted, so the row is
// written in one pass (self position, the k-th key for EXCL_PY_LINEAR, then
// every kept entry up to k, then -1 / +inf padding).
template <typename F>
__global__ __launch_bounds__(64) void write_long_rows_kernel(LongRowArgs<F> a) {
  using C = Comp<F>;
  const int64_t qs = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (qs >= a.nq) return;
  const typename C::T* row = a.raw + qs * a.stride;
  const int cnt = a.cnt[qs];
  const int64_t qr = a.qrows ? a.qrows[qs] : qs;
  const int64_t self = a.qbase + qr;
  int64_t* oidx = a.out_idx + qr * a.k;
  F* odist = a.out_dist + qr * a.k;
  int sp = -1;
  if (a.mode != EXCL_NONE)
    for (int e = 0; e < cnt; ++e)
      if ((int64_t)C::idx(row[e]) == self) {
        sp = e;
        break;
      }
  const F kth = (a.mode == EXCL_PY_LINEAR && a.k >= 1 && cnt >= a.k) ? C::key(row[a.k - 1]) : F(INFINITY);
  int j = 0;
  for (int e = 0; e < cnt && j < a.k; ++e) {
    if (e == sp) continue;
    const F key = C::key(row[e]);
    if (a.mode == EXCL_PY_LINEAR && key > kth) break;
    oidx[j] = (int64_t)C::idx(row[e]);
    odist[j] = key;
    ++j;
  }
  for (; j < a.k; ++j) {
    oidx[j] = -1;
    odist[j] = F(INFINITY);
  }
}

template <typename F>
hipError_t launch_write_long_rows(const LongRowArgs<F>& a, hipStream_t st) {
  if (a.nq == 0) return hipSuccess;
  hipLaunchKernelGGL((write_long_rows_kernel<F>), dim3((unsigned)((a.nq + 63) / 64)), dim3(64), 0, st, a);
  return hipGetLastError();
}
template hipError_t launch_write_long_rows<float>(const LongRowArgs<float>&, hipStream_t);
template hipError_t launch_write_long_rows<double>(const LongRowArgs<double>&, hipStream_t);

template <typename F>
hipError_t launch_scan_exact(const ScanExactArgs<F>& a, int qw, int nsplit, hipStream_t st) {
  dim3 block(256);
  // the f64 fallback form keeps 8 double lanes per query: at most 8 queries per
  // workgroup keep its accumulators in registers
  if (a.arith == ARITH_FALLBACK && sizeof(F) == 8 && qw > 8) qw = 8;
  dim3 grid((unsigned)((a.nq + qw - 1) / qw), (unsigned)nsplit);
  if (a.arith == ARITH_FALLBACK) {
    if (qw >= 16)
      hipLaunchKernelGGL((scan_exact_kernel<F, 16, 8, ARITH_FALLBACK>), grid, block, 0, st, a);
    else if (qw == 8)
      hipLaunchKernelGGL((scan_exact_kernel<F, 8, 8, ARITH_FALLBACK>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((scan_exact_kernel<F, 4, 8, ARITH_FALLBACK>), grid, block, 0, st, a);
  } else if (qw == 16) {
    hipLaunchKernelGGL((scan_exact_kernel<F, 16, 8, ARITH_AVX>), grid, block, 0, st, a);
  } else if (qw == 8) {
    hipLaunchKernelGGL((scan_exact_kernel<F, 8, 8, ARITH_AVX>), grid, block, 0, st, a);
  } else {
    hipLaunchKernelGGL((scan_exact_kernel<F, 4, 8, ARITH_AVX>), grid, block, 0, st, a);
  }
  return hipGetLastError();
}

template <typename F>
hipError_t launch_finalize_exact(const FinalizeArgs<F>& a, hipStream_t st) {
  hipLaunchKernelGGL((finalize_exact_kernel<F, 8>), dim3((unsigned)a.nq), dim3(64), 0, st, a);
  return hipGetLastError();
}

template hipError_t launch_scan_exact<float>(const ScanExactArgs<float>&, int, int, hipStream_t);
template hipError_t launch_scan_exact<double>(const ScanExactArgs<double>&, int, int, hipStream_t);
template hipError_t launch_finalize_exact<float>(const FinalizeArgs<float>&, hipStream_t);
template hipError_t launch_finalize_exact<double>(const FinalizeArgs<double>&, hipStream_t);

}  // namespace fa
