// Exact rerank and certification of the MFMA candidates (one wave per query).
//
//  1. Union of the per-split candidate lists, cut to the KP smallest approximate
//     keys (composite order).  Every row outside the kept set has an
//     approximate key >= T, the KP-th kept key (rows of a full split list are
//     >= that split's KP-th key >= T; rows cut here are >= T by the sort).
//  2. Exact rerank: the restated fuel distance of every kept candidate, sorted
//     by (distance, index) -> the top KK.
//  3. Certification: a proven lower bound on the restated distance of every
//     excluded row.  With s~ = fl(hx - g) the approximate key and E_s a bound on
//     |s~ - (||x||^2/2 - q.x)|, every excluded x has
//         D(x) = ||q - x||^2 >= ||q||^2 + 2 T - 2 E_s = L,
//     and its restated f32 value r(x) >= L (1 - gamma_r) (or, when every
//     coordinate is an integer of magnitude <= 256, r(x) >= min(L, 2^24 (1 -
//     gamma_r)), since then r(x) = D(x) whenever D(x) < 2^24).  If the largest
//     float B below that bound already orders after the KK-th exact key
//     (sqrt(B) > key for 'euclidean', B > key for 'sqeuclidean'), no excluded
//     row can enter the answer, ties included, and the top KK is fuel's exact
//     answer.  Otherwise the query is queued for the exact fallback scan.
//  E_s collects, for key~ = -(MFMA of x~ . q~ with C input -fl32(||x||^2/2)),
//  low 4 bits replaced by the row tag: operand rounding |q~.x~ - q.x| <= |q|
//  |x~-x| + |q~-q| |x| + |q~-q| |x~-x| (zero for the f32 operands); the
//  rounding of fl32(||x||^2/2); MFMA accumulation <= 4 (d+2) u (|q~| |x~| + H),
//  H = max fl32(||x||^2/2) (zero when every partial sum is a multiple of 1/2
//  below 2^23, which holds for integer data with small norms); the tag, <= 15
//  ulps of the result, <= 2^-19 (|q~| |x~| + H) + 15 * 2^-149.  u = 2^-24.  All of it is
//  evaluated in f64 with directed slack; the index-wide maxima make the bound
//  uniform over the excluded rows.
#include "common.h"
#include "kernels.h"

namespace fa {

// Restated f32 distance of two 32-byte-aligned rows (fuel's AVX2 order).
__device__ __forceinline__ float restated_aligned(const float* __restrict__ a, const float* __restrict__ b, int d) {
  const int sd = d & ~7;
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < sd; i += 8) {
    const float4 a0 = *reinterpret_cast<const float4*>(a + i), a1 = *reinterpret_cast<const float4*>(a + i + 4);
    const float4 b0 = *reinterpret_cast<const float4*>(b + i), b1 = *reinterpret_cast<const float4*>(b + i + 4);
    float t;
    t = a0.x - b0.x; acc[0] = fma_rn(t, t, acc[0]);
    t = a0.y - b0.y; acc[1] = fma_rn(t, t, acc[1]);
    t = a0.z - b0.z; acc[2] = fma_rn(t, t, acc[2]);
    t = a0.w - b0.w; acc[3] = fma_rn(t, t, acc[3]);
    t = a1.x - b1.x; acc[4] = fma_rn(t, t, acc[4]);
    t = a1.y - b1.y; acc[5] = fma_rn(t, t, acc[5]);
    t = a1.z - b1.z; acc[6] = fma_rn(t, t, acc[6]);
    t = a1.w - b1.w; acc[7] = fma_rn(t, t, acc[7]);
  }
  float s = -0.0f;
#pragma unroll
  for (int j = 0; j < 8; ++j) s = s + acc[j];
  for (int i = sd; i < d; ++i) {
    const float t = a[i] - b[i];
    s = fma_rn(t, t, s);
  }
  return s;
}

__device__ __forceinline__ bool certify(const RerankArgs& a, int64_t qs, float T, float kk) {
  const double u = 0x1p-24;
  const int d = a.d;
  const double sl = (d + 8) * 0x1p-52;
  const double X2 = __longlong_as_double((long long)a.consts[0]);  // upper bounds already
  const double ex = sqrt(__longlong_as_double((long long)a.consts[1])) * (1 + 0x1p-50);
  const double xt = sqrt(__longlong_as_double((long long)a.consts[2])) * (1 + 0x1p-50);
  const double xn = sqrt(X2) * (1 + 0x1p-50);
  // beyond these norms keys may reach the scan's finite threshold sentinel
  if (!(X2 < MFMA_NORM2_LIMIT) || !(a.qn2[qs] < MFMA_NORM2_LIMIT)) return false;
  const double qn2_lo = a.qn2[qs] * (1 - sl);
  const double qn = sqrt(a.qn2[qs] * (1 + sl)) * (1 + 0x1p-50);
  const double eq = sqrt(a.eq2[qs] * (1 + sl)) * (1 + 0x1p-50);
  const double qt = sqrt(a.qt2[qs] * (1 + sl)) * (1 + 0x1p-50);
  const bool int_regime = a.flags[0] != 0 && a.qint[qs] != 0;
  const double dot_mag = qt * xt * (1 + 0x1p-50);
  const double e_hx = __longlong_as_double((long long)a.consts[3]);
  const double H = __longlong_as_double((long long)a.consts[4]);
  const bool mfma_exact = int_regime && e_hx == 0.0 && dot_mag + H < 0x1p23;
  const double gacc = mfma_exact ? 0.0 : 4.0 * (d + 2) * u;
  const double e_round = qn * ex + eq * xn + eq * ex;
  const double e_tag = 0x1p-19 * (dot_mag + H) * (1 + 0x1p-10) + 0x1p-145;
  double e_s = e_round + e_hx + gacc * (dot_mag + H) + e_tag + (mfma_exact ? 0.0 : 2.0 * (d + 2) * 0x1p-126);
  e_s *= (1 + 0x1p-40);
  double L = 2.0 * (double)T + qn2_lo;
  L -= (fabs(2.0 * (double)T) + qn2_lo) * 0x1p-50;
  L -= 2.0 * e_s;
  const int n_r = d / 8 + 8 + (d % 8) + 2;
  const double gr = n_r * u * (1 + 0x1p-10);
  double Lr;
  if (int_regime)
    Lr = fmin(L, 0x1p24 * (1 - gr));
  else
    Lr = L * (1 - gr) - (2.0 * d + 16) * 0x1p-149;
  if (!(Lr > 0.0)) return false;
  float B = (float)Lr;
  if ((double)B > Lr) B = nextafterf(B, -INFINITY);
  return a.metric == 0 ? (sqrt_rn(B) > kk) : (B > kk);
}

template <int V>
__global__ __launch_bounds__(64) void rerank_kernel(RerankArgs a) {
  using C = Comp<float>;
  using T = uint64_t;
  constexpr int NS = 64 * V;
  const int64_t qs = blockIdx.x;
  if (qs >= a.nq) return;
  const int lane = __lane_id();
  const int KP = a.KP;

  // 1. union of split lists -> KP smallest approximate keys
  T x[V];
#pragma unroll
  for (int v = 0; v < V; ++v) x[v] = C::MAX;
  int fill = 0;
  int64_t total = 0;
  for (int s = 0; s < a.nsplit; ++s) {
    const T* sl = a.slots + (size_t(s) * a.nq_pad + qs) * a.slot_ns;
    const int cnt = a.lcnts[size_t(s) * a.nq_pad + qs];
    total += cnt;
    int off = 0;
    while (off < cnt) {
      const int room = NS - fill;
      const int take = (cnt - off) < room ? (cnt - off) : room;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int e = v * 64 + lane;
        if (e >= fill && e < fill + take) x[v] = sl[off + e - fill];
      }
      fill += take;
      off += take;
      if (fill == NS) {
        wave_bitonic_sort<T, V>(x);
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (v * 64 + lane >= KP) x[v] = C::MAX;
        fill = KP;
      }
    }
  }
  wave_bitonic_sort<T, V>(x);
  const int ncand = total < KP ? (int)total : KP;
  const float T_C = total >= KP ? C::key(wave_get<T, V>(x, KP - 1)) : INFINITY;

  // 2. exact rerank
  const float* qrow = a.Q32 + qs * a.ld32;
  T y[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    y[v] = C::MAX;
    if (e < ncand) {
      const uint32_t idx = C::idx(x[v]);
      const float s = restated_aligned(qrow, a.X + size_t(idx) * a.ldx, a.d);
      const float key = a.metric == 0 ? sqrt_rn(s) : s;
      y[v] = C::make(key, idx);
    }
  }
  wave_bitonic_sort<T, V>(y);

  // 3. certification
  bool cert = true;
  if (total >= KP) {
    const float kk = C::key(wave_get<T, V>(y, a.KK - 1));
    cert = certify(a, qs, T_C, kk);
  }
  const int64_t orow = a.rows ? a.rows[qs] : qs;
  if (cert) {
    const int cnt = ncand < a.KK ? ncand : a.KK;
    write_knn_row<float, V>(y, cnt, a.qbase + orow, a.k, a.mode, a.out_idx + orow * a.k, a.out_dist + orow * a.k);
  } else if (lane == 0) {
    const int pos = atomicAdd(a.fail_count, 1);
    a.fail_list[pos] = orow;
  }
}

hipError_t launch_rerank(const RerankArgs& a, int V, hipStream_t st) {
  if (a.nq == 0) return hipSuccess;
  if (V == 4)
    hipLaunchKernelGGL((rerank_kernel<4>), dim3((unsigned)a.nq), dim3(64), 0, st, a);
  else if (V == 8)
    hipLaunchKernelGGL((rerank_kernel<8>), dim3((unsigned)a.nq), dim3(64), 0, st, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace fa
