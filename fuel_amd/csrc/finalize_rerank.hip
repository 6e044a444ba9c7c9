// Exact rerank and certification of the MFMA candidates (one wave per query).
//
//  1. Union of the per-split candidate lists, cut to the KP smallest approximate
//     keys (composite order).  Every row outside the kept set has an
//     approximate key >= T, the KP-th kept key (rows of a full split list are
//     >= that split's KP-th key >= T; rows cut here are >= T by the sort).
//  2. Exact rerank: the restated fuel distance of every kept candidate, sorted
//     by (distance, index) -> the top KK.
//  3. Certification: a proven lower bound on the restated distance of every
//     excluded row.  T below is the KP-th kept key when this kernel cut the
//     union; otherwise the scan's final shared threshold, below which every
//     row a split dropped cannot lie.  With key~ the approximate key and E_s
//     a bound on |key~ - (||x||^2/2 - q.x)|, every excluded x has
//         D(x) = ||q - x||^2 >= ||q||^2 + 2 T - 2 E_s = L,
//     and its restated f32 value r(x) >= L (1 - gamma_r) (or, when every
//     coordinate is an integer of magnitude <= 256, r(x) >= min(L, 2^24 (1 -
//     gamma_r)), since then r(x) = D(x) whenever D(x) < 2^24).  If the largest
//     float B below that bound already orders after the KK-th exact key
//     (sqrt(B) > key for 'euclidean', B > key for 'sqeuclidean'), no excluded
//     row can enter the answer, ties included, and the top KK is fuel's exact
//     answer.  Otherwise the query is queued for the next level of the chain,
//     with the threshold theta for which the certificate would hold given its
//     current KK-th exact key: a second scan that starts at theta keeps every
//     row that can still enter the answer, and certifies with T = theta.
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

// Restated f32 distance of two 32-byte-aligned rows in fuel's order: 8 lanes
// interleaved over the coordinates, a left fold, then the tail; ARITH_AVX fuses
// every lane and tail update (src/distance/squared_euclidean.rs:13-39),
// ARITH_FALLBACK rounds the square first (:99-133).
template <int ARITH>
__device__ __forceinline__ float restated_aligned(const float* __restrict__ a, const float* __restrict__ b, int d) {
  auto step = [](float t, float acc) {
    if constexpr (ARITH == ARITH_FALLBACK) return t * t + acc;  // two roundings (-ffp-contract=off)
    else return fma_rn(t, t, acc);
  };
  const int sd = d & ~7;
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < sd; i += 8) {
    const float4 a0 = *reinterpret_cast<const float4*>(a + i), a1 = *reinterpret_cast<const float4*>(a + i + 4);
    const float4 b0 = *reinterpret_cast<const float4*>(b + i), b1 = *reinterpret_cast<const float4*>(b + i + 4);
    acc[0] = step(a0.x - b0.x, acc[0]);
    acc[1] = step(a0.y - b0.y, acc[1]);
    acc[2] = step(a0.z - b0.z, acc[2]);
    acc[3] = step(a0.w - b0.w, acc[3]);
    acc[4] = step(a1.x - b1.x, acc[4]);
    acc[5] = step(a1.y - b1.y, acc[5]);
    acc[6] = step(a1.z - b1.z, acc[6]);
    acc[7] = step(a1.w - b1.w, acc[7]);
  }
  float s = ARITH == ARITH_FALLBACK ? 0.0f : -0.0f;
#pragma unroll
  for (int j = 0; j < 8; ++j) s = s + acc[j];
  for (int i = sd; i < d; ++i) s = step(a[i] - b[i], s);
  return s;
}

// Per-query terms of the bound (shared by the certificate and the threshold of
// the certified second pass).
struct BoundTerms {
  double e_s;     // bound on |key~ - (||x||^2/2 - q.x)| over every row
  double qn2_lo;  // lower bound of ||q||^2
  double gr;      // relative error of the restated f32 distance
  int d;
  bool int_regime;
  bool usable;    // norms inside the scan's sentinel range
};

__device__ __forceinline__ BoundTerms bound_terms(const RerankArgs& a, int64_t qs) {
  const double u = 0x1p-24;
  BoundTerms b;
  const int d = a.d;
  b.d = d;
  const double sl = (d + 8) * 0x1p-52;
  const double X2 = __longlong_as_double((long long)a.consts[0]);  // upper bounds already
  const double ex = sqrt(__longlong_as_double((long long)a.consts[1])) * (1 + 0x1p-50);
  const double xt = sqrt(__longlong_as_double((long long)a.consts[2])) * (1 + 0x1p-50);
  const double xn = sqrt(X2) * (1 + 0x1p-50);
  // beyond these norms keys may reach the scan's finite threshold sentinel
  b.usable = X2 < MFMA_NORM2_LIMIT && a.qn2[qs] < MFMA_NORM2_LIMIT;
  b.qn2_lo = a.qn2[qs] * (1 - sl);
  const double qn = sqrt(a.qn2[qs] * (1 + sl)) * (1 + 0x1p-50);
  const double eq = sqrt(a.eq2[qs] * (1 + sl)) * (1 + 0x1p-50);
  const double qt = sqrt(a.qt2[qs] * (1 + sl)) * (1 + 0x1p-50);
  b.int_regime = a.flags[0] != 0 && a.qint[qs] != 0;
  const double dot_mag = qt * xt * (1 + 0x1p-50);
  const double e_hx = __longlong_as_double((long long)a.consts[3]);
  const double H = __longlong_as_double((long long)a.consts[4]);
  const bool mfma_exact = b.int_regime && e_hx == 0.0 && dot_mag + H < 0x1p23;
  const double gacc = mfma_exact ? 0.0 : 4.0 * (d + 2) * u;
  const double e_round = qn * ex + eq * xn + eq * ex;
  const double e_tag = 0x1p-19 * (dot_mag + H) * (1 + 0x1p-10) + 0x1p-145;
  double e_s = e_round + e_hx + gacc * (dot_mag + H) + e_tag + (mfma_exact ? 0.0 : 2.0 * (d + 2) * 0x1p-126);
  b.e_s = e_s * (1 + 0x1p-40);
  // relative error of the restated sum of non-negative terms: a lane term sees
  // the subtraction (2u), its square's rounding (fused or not: u), its lane's
  // remaining updates (d/8 - 1), the 7 fold additions and the d%8 tail
  // updates; a tail term at most 2u + u + d%8 u.  Both rounding forms fit.
  const int n_r = d / 8 + 8 + (d % 8) + 2;
  b.gr = n_r * u * (1 + 0x1p-10);
  return b;
}

// Every excluded row has key~ >= T.  True iff no excluded row can order before
// the KK-th exact key kk (ties included).
__device__ __forceinline__ bool certify(const RerankArgs& a, const BoundTerms& b, float T, float kk) {
  if (!b.usable) return false;
  double L = 2.0 * (double)T + b.qn2_lo;
  L -= (fabs(2.0 * (double)T) + b.qn2_lo) * 0x1p-50;
  L -= 2.0 * b.e_s;
  double Lr;
  if (b.int_regime)
    Lr = fmin(L, 0x1p24 * (1 - b.gr));
  else
    Lr = L * (1 - b.gr) - (2.0 * b.d + 16) * 0x1p-149;
  if (!(Lr > 0.0)) return false;
  float B = (float)Lr;
  if ((double)B > Lr) B = nextafterf(B, -INFINITY);
  return a.metric == 0 ? (sqrt_rn(B) > kk) : (B > kk);
}

// The smallest float T (up to a small margin) for which certify(T, kk) holds:
// a second scan that keeps every row with key~ < T finds every row that can
// still order before kk, and its answer is then certified with T itself.
// +inf when no such T exists (the integer regime's cap).
__device__ __forceinline__ float certified_threshold(const RerankArgs& a, const BoundTerms& b, float kk) {
  if (!b.usable || !(kk < INFINITY)) return INFINITY;
  // target for the restated-distance bound B: strictly above kk (squared form)
  double target;
  if (a.metric == 0) {
    const double kn = (double)nextafterf(kk, INFINITY);
    target = kn * kn;
  } else {
    target = (double)kk;
  }
  target = target * (1 + 0x1p-20) + 0x1p-140;
  double needL;
  if (b.int_regime) {
    if (!(target < 0x1p24 * (1 - b.gr))) return INFINITY;
    needL = target;
  } else {
    needL = (target + (2.0 * b.d + 16) * 0x1p-149) / (1 - b.gr);
  }
  const double twoT = needL - b.qn2_lo + 2.0 * b.e_s;
  double T = 0.5 * twoT;
  T += fabs(T) * 0x1p-40 + (fabs(twoT) + b.qn2_lo) * 0x1p-48;
  float Tf = (float)T;
  if ((double)Tf < T) Tf = nextafterf(Tf, INFINITY);
  for (int i = 0; i < 8 && !certify(a, b, Tf, kk); ++i) Tf = nextafterf(Tf, INFINITY);
  return certify(a, b, Tf, kk) ? Tf : INFINITY;
}

// Residual operand: an excluded row's stored key s >= T is, up to the row tag
// (<= 15 ulps of the MFMA result), a lower bound of D(x) / 2, so
// D(x) >= 2 (T - 2^-19 |T| (1 + 2^-10) - 2^-145) = L (s - c|s| increases with s),
// and its restated f32 value r(x) >= L (1 - gamma_r) - (2 d + 16) 2^-149.
__device__ __forceinline__ double resid_lower(double T) {
  return 2.0 * (T - 0x1p-19 * fabs(T) * (1 + 0x1p-10) - 0x1p-145);
}

__device__ __forceinline__ bool certify_resid(const RerankArgs& a, const BoundTerms& b, float T, float kk) {
  if (!b.usable) return false;
  const double L = resid_lower((double)T);
  const double Lr = L * (1 - b.gr) - (2.0 * b.d + 16) * 0x1p-149;
  if (!(Lr > 0.0)) return false;
  float B = (float)Lr;
  if ((double)B > Lr) B = nextafterf(B, -INFINITY);
  return a.metric == 0 ? (sqrt_rn(B) > kk) : (B > kk);
}

__device__ __forceinline__ float certified_threshold_resid(const RerankArgs& a, const BoundTerms& b, float kk) {
  if (!b.usable || !(kk < INFINITY)) return INFINITY;
  double target;
  if (a.metric == 0) {
    const double kn = (double)nextafterf(kk, INFINITY);
    target = kn * kn;
  } else {
    target = (double)kk;
  }
  target = target * (1 + 0x1p-20) + 0x1p-140;
  const double needL = (target + (2.0 * b.d + 16) * 0x1p-149) / (1 - b.gr);
  // 2 (T - c T - 2^-145) >= needL for T >= 0
  double T = (0.5 * needL + 0x1p-145) / (1 - 0x1p-19 * (1 + 0x1p-10));
  T += fabs(T) * 0x1p-40 + 0x1p-140;
  float Tf = (float)T;
  if ((double)Tf < T) Tf = nextafterf(Tf, INFINITY);
  for (int i = 0; i < 8 && !certify_resid(a, b, Tf, kk); ++i) Tf = nextafterf(Tf, INFINITY);
  return certify_resid(a, b, Tf, kk) ? Tf : INFINITY;
}

// Ascending sort of the first `count` of 64*V entries whose tail is C::MAX:
// only the smallest register prefix that holds them is sorted (the rest is
// already in place).
template <typename T, int V>
__device__ __forceinline__ void sort_prefix(T (&x)[V], int count) {
  if (count <= 64 || V == 1) {
    wave_bitonic_sort<T, 1>(reinterpret_cast<T(&)[1]>(x[0]));
  } else if (V == 2 || count <= 128) {
    wave_bitonic_sort<T, (V < 2 ? V : 2)>(reinterpret_cast<T(&)[(V < 2 ? V : 2)]>(x[0]));
  } else if (V == 4 || count <= 256) {
    wave_bitonic_sort<T, (V < 4 ? V : 4)>(reinterpret_cast<T(&)[(V < 4 ? V : 4)]>(x[0]));
  } else {
    wave_bitonic_sort<T, V>(x);
  }
}

// The n-th smallest (from 1) high word of the entries in registers [0, nr) of
// the wave: the smallest t with #(high word <= t) >= n, by a 32-step search.
template <int V>
__device__ __forceinline__ uint32_t wave_kth_hi(const uint64_t (&y)[V], int nr, int n) {
  uint32_t p = 0;
  for (int b = 31; b >= 0; --b) {
    const uint32_t probe = p | ((1u << b) - 1u);
    int c = 0;
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (v < nr) c += __popcll(__ballot(uint32_t(y[v] >> 32) <= probe));
    if (c < n) p |= (1u << b);
  }
  return p;
}

template <int V, int ARITH>
__device__ __forceinline__ void rerank_body(const RerankArgs& a, const int64_t qs) {
  using C = Comp<float>;
  using T = uint64_t;
  constexpr int NS = 64 * V;
  if (qs >= a.nq) return;
  const int lane = __lane_id();
  const int KP = a.KP;

  // 1. union of split lists -> KP smallest approximate keys.  Entries are
  //    compacted into an LDS buffer; whenever it holds NS, it is sorted and cut
  //    back to its KP smallest, and `cut` (the KP-th of those) drops every later
  //    entry at or above it: such an entry cannot be among the union's KP
  //    smallest (composites are distinct: each row belongs to one split).
  __shared__ T buf[NS + 64];
  T x[V];
  T cut = C::MAX;
  int fill = 0;
  int64_t total = 0;
  auto absorb = [&](const T val) {
    const bool keep = val < cut;
    const uint64_t mask = __ballot(keep);
    if (keep) buf[fill + __popcll(mask & ((uint64_t(1) << lane) - 1))] = val;
    fill += __popcll(mask);
    if (fill >= NS) {
      __syncthreads();
#pragma unroll
      for (int v = 0; v < V; ++v) x[v] = buf[v * 64 + lane];
      const T over = lane < fill - NS ? buf[NS + lane] : C::MAX;
      wave_bitonic_sort<T, V>(x);
      cut = wave_get<T, V>(x, KP - 1);
      __syncthreads();
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (v * 64 + lane < KP) buf[v * 64 + lane] = x[v];
      if (lane < fill - NS) buf[KP + lane] = over;
      fill = KP + (fill - NS);
      __syncthreads();
    }
  };
  // The counts of 64 splits come in one load (lane s: split s0 + s); only the
  // non-empty lists are visited, PF at a time with their first 64 entries
  // loaded together (small batches run hundreds of splits, and a query that
  // starts from a tight threshold leaves most of their lists empty).
  constexpr int PF = 4;
  int64_t my_total = 0;
  for (int s0 = 0; s0 < a.nsplit; s0 += 64) {
    const int c = s0 + lane < a.nsplit ? a.lcnts[size_t(s0 + lane) * a.nq_pad + qs] : 0;
    my_total += c;
    uint64_t mask = __ballot(c > 0);
    while (mask) {
      int cnt[PF];
      const T* sl[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        cnt[u] = 0;
        sl[u] = a.slots;
        if (mask) {
          const int b = __builtin_ctzll(mask);
          mask &= mask - 1;
          cnt[u] = __builtin_amdgcn_readlane(c, b);
          sl[u] = a.slots + (size_t(s0 + b) * a.nq_pad + qs) * a.slot_ns;
        }
      }
      T v0[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) v0[u] = lane < cnt[u] ? sl[u][lane] : C::MAX;
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        absorb(v0[u]);
        for (int off = 64; off < cnt[u]; off += 64) absorb(off + lane < cnt[u] ? sl[u][off + lane] : C::MAX);
      }
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) my_total += __shfl_xor(my_total, m);
  total = my_total;
  __syncthreads();
#pragma unroll
  for (int v = 0; v < V; ++v) x[v] = v * 64 + lane < fill ? buf[v * 64 + lane] : C::MAX;
  sort_prefix<T, V>(x, fill);
  const int ncand = total < KP ? (int)total : KP;
  // Bound on the excluded rows: with a cut here, the KP-th kept key; otherwise
  // the scan's final shared threshold (every row a split dropped lies at or
  // above it).  At or above KEY_LIMIT no real row was excluded.
  const float T_C = total > KP ? C::key(wave_get<T, V>(x, KP - 1)) : unord_f32(a.gthr[qs]);

  // 2. exact rerank, 64 candidates at a time in the order of their approximate
  //    keys.  Once the rounds done hold KK exact keys, kk' (the KK-th of them)
  //    bounds the final KK-th exact key from above; a later candidate whose own
  //    key~ = s already certifies against kk' (the bound of step 3 with T = s:
  //    its restated distance orders strictly after kk') cannot enter the top KK
  //    and is not gathered.  The top KK, kk and everything after are unchanged.
  const float* qrow = a.Q32 + qs * a.ld32;
  const BoundTerms bt = bound_terms(a, qs);
  T y[V];
  float kkp = INFINITY;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    y[v] = C::MAX;
    if (v * 64 >= ncand) continue;
    bool eval = e < ncand;
    if (eval && kkp < INFINITY) {
      const float s = C::key(x[v]);
      if (s < MFMA_KEY_LIMIT && (a.resid ? certify_resid(a, bt, s, kkp) : certify(a, bt, s, kkp))) eval = false;
    }
    if (eval) {
      uint32_t idx = C::idx(x[v]);
      if (a.resid) idx = a.rperm[idx];  // padded position -> row (padding rows never survive)
      if (idx != RPERM_NONE) {
        const float s = restated_aligned<ARITH>(qrow, a.X + size_t(idx) * a.ldx, a.d);
        const float key = a.metric == 0 ? sqrt_rn(s) : s;
        y[v] = C::make(key, idx);
      }
    }
    if (!(kkp < INFINITY) && (v + 1) * 64 >= a.KK && (v + 1) * 64 < ncand) {
      const uint32_t h = wave_kth_hi<V>(y, v + 1, a.KK);  // every entry of rounds 0..v was evaluated
      if (h != 0xFFFFFFFFu) kkp = unord_f32(h);
    }
  }
  sort_prefix<T, V>(y, ncand);

  // 3. certification
  bool cert = true;
  float kk = INFINITY;
  if (T_C < MFMA_KEY_LIMIT) {
    kk = ncand >= a.KK ? C::key(wave_get<T, V>(y, a.KK - 1)) : INFINITY;
    cert = a.resid ? certify_resid(a, bt, T_C, kk) : certify(a, bt, T_C, kk);
  }
  const int64_t orow = a.rows ? a.rows[qs] : qs;
  if (cert) {
    const int cnt = ncand < a.KK ? ncand : a.KK;
    write_knn_row<float, V>(y, cnt, a.qbase + orow, a.k, a.mode, a.out_idx + orow * a.k, a.out_dist + orow * a.k);
  } else {
    const float theta = !a.fail_theta ? INFINITY
                        : (a.resid ? certified_threshold_resid(a, bt, kk) : certified_threshold(a, bt, kk));
    if (lane == 0) {
      const int pos = atomicAdd(a.fail_count, 1);
      a.fail_list[pos] = orow;
      if (a.fail_theta) a.fail_theta[pos] = theta;
    }
  }
}

template <int V, int ARITH>
__global__ __launch_bounds__(64) void rerank_kernel(RerankArgs a) {
  rerank_body<V, ARITH>(a, blockIdx.x);
}

// Grouped form: one wave per query of every problem of the table.
template <int V, int ARITH>
__global__ __launch_bounds__(64) void rerank_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const RerankArgs a = static_cast<const RerankArgs*>(g.tab)[p];
  rerank_body<V, ARITH>(a, int64_t(blockIdx.x) - g.start[p]);
}

hipError_t launch_rerank_group(const GroupTab& g, int V, int arith, hipStream_t st) {
  if (g.np <= 0 || g.nblocks <= 0) return hipSuccess;
  const bool fb = arith == ARITH_FALLBACK;
  const dim3 grid((unsigned)g.nblocks);
  if (V == 4 && !fb)
    hipLaunchKernelGGL((rerank_group_kernel<4, ARITH_AVX>), grid, dim3(64), 0, st, g);
  else if (V == 8 && !fb)
    hipLaunchKernelGGL((rerank_group_kernel<8, ARITH_AVX>), grid, dim3(64), 0, st, g);
  else if (V == 4)
    hipLaunchKernelGGL((rerank_group_kernel<4, ARITH_FALLBACK>), grid, dim3(64), 0, st, g);
  else if (V == 8)
    hipLaunchKernelGGL((rerank_group_kernel<8, ARITH_FALLBACK>), grid, dim3(64), 0, st, g);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rerank(const RerankArgs& a, int V, hipStream_t st) {
  if (a.nq == 0) return hipSuccess;
  const bool fb = a.arith == ARITH_FALLBACK;
  if (V == 4 && !fb)
    hipLaunchKernelGGL((rerank_kernel<4, ARITH_AVX>), dim3((unsigned)a.nq), dim3(64), 0, st, a);
  else if (V == 8 && !fb)
    hipLaunchKernelGGL((rerank_kernel<8, ARITH_AVX>), dim3((unsigned)a.nq), dim3(64), 0, st, a);
  else if (V == 4)
    hipLaunchKernelGGL((rerank_kernel<4, ARITH_FALLBACK>), dim3((unsigned)a.nq), dim3(64), 0, st, a);
  else if (V == 8)
    hipLaunchKernelGGL((rerank_kernel<8, ARITH_FALLBACK>), dim3((unsigned)a.nq), dim3(64), 0, st, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace fa
