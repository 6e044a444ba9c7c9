// Device-side building blocks shared by the scan, finalize, range and pairwise
// kernels: the restated fuel distance, order-preserving composite keys and the
// wave-cooperative bitonic merge that maintains per-query candidate lists.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace fa {

constexpr int WAVE = 64;

// ---------------------------------------------------------------------------
// Restated squared distance (fuel src/distance/squared_euclidean.rs:13-39 for f32,
// :41-67 for f64).  LANES accumulators interleaved over the coordinates
// (8 for f32, 4 for f64), each updated with a fused multiply-add of the rounded
// difference; a left-fold lane sum; then a fused multiply-add tail over
// d mod LANES.  The order of every rounding matches the reference, so the
// result is bit-identical to the AVX2 kernel.  Compiled with -ffp-contract=off.
template <typename F>
struct Lanes;
template <>
struct Lanes<float> {
  static constexpr int N = 8;
};
template <>
struct Lanes<double> {
  static constexpr int N = 4;
};

// Rounding forms of the restated distance (the index option
// FUEL_AMD_FLAG_FALLBACK_ARITH selects the second).
constexpr int ARITH_AVX = 0;
constexpr int ARITH_FALLBACK = 1;

__device__ __forceinline__ float fma_rn(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_rn(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_rn(double x) { return __builtin_sqrt(x); }

// ---------------------------------------------------------------------------
// Composite keys.  A candidate is (key, index) ordered lexicographically, the
// total order of fuel's DistPair (src/api/search.rs:23-30).  For f32 keys the
// pair packs into one u64: the key mapped to an order-preserving u32 in the high
// half, the index in the low half.  For f64 keys the pair is a u64 key plus the
// index, packed into an unsigned __int128.
__device__ __host__ __forceinline__ uint32_t ord_f32(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __host__ __forceinline__ float unord_f32(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __builtin_bit_cast(float, u);
}
__device__ __host__ __forceinline__ uint64_t ord_f64(double f) {
  uint64_t u = __builtin_bit_cast(uint64_t, f);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __host__ __forceinline__ double unord_f64(uint64_t o) {
  uint64_t u = (o & 0x8000000000000000ull) ? (o & 0x7fffffffffffffffull) : ~o;
  return __builtin_bit_cast(double, u);
}

template <typename F>
struct Comp;
template <>
struct Comp<float> {
  using T = uint64_t;
  static constexpr T MAX = ~uint64_t(0);
  __device__ __host__ static __forceinline__ T make(float key, uint32_t idx) {
    return (uint64_t(ord_f32(key)) << 32) | idx;
  }
  __device__ __host__ static __forceinline__ float key(T c) { return unord_f32(uint32_t(c >> 32)); }
  __device__ __host__ static __forceinline__ uint32_t idx(T c) { return uint32_t(c); }
};
template <>
struct Comp<double> {
  using T = unsigned __int128;
  static constexpr T MAX = ~T(0);
  __device__ __host__ static __forceinline__ T make(double key, uint32_t idx) {
    return (T(ord_f64(key)) << 64) | T(idx);
  }
  __device__ __host__ static __forceinline__ double key(T c) { return unord_f64(uint64_t(c >> 64)); }
  __device__ __host__ static __forceinline__ uint32_t idx(T c) { return uint32_t(c); }
};

// ---------------------------------------------------------------------------
// Cross-lane exchange (ds_bpermute; 64-bit values move as two dwords).
__device__ __forceinline__ uint32_t shfl_xor_u32(uint32_t v, int mask) {
  int lane = __lane_id();
  return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ mask) << 2, (int)v);
}
__device__ __forceinline__ uint64_t shfl_xor(uint64_t v, int mask) {
  uint32_t lo = shfl_xor_u32(uint32_t(v), mask);
  uint32_t hi = shfl_xor_u32(uint32_t(v >> 32), mask);
  return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ unsigned __int128 shfl_xor(unsigned __int128 v, int mask) {
  uint64_t lo = shfl_xor(uint64_t(v), mask);
  uint64_t hi = shfl_xor(uint64_t(v >> 64), mask);
  return ((unsigned __int128)hi << 64) | lo;
}

// Bitonic sort (ascending) of 64*V values held V per lane; element e = v*64 + lane.
template <typename T, int V>
__device__ __forceinline__ void wave_bitonic_sort(T (&x)[V]) {
  const int lane = __lane_id();
  constexpr int NS = 64 * V;
#pragma unroll
  for (int k = 2; k <= NS; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const int vj = j >> 6;
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int pv = v ^ vj;
          if (pv > v) {
            const int e = v * 64 + lane;
            const bool asc = (e & k) == 0;
            T a = x[v], b = x[pv];
            const bool sw = asc ? (a > b) : (a < b);
            x[v] = sw ? b : a;
            x[pv] = sw ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int e = v * 64 + lane;
          const bool asc = (e & k) == 0;
          const bool lower = (lane & j) == 0;
          T p = shfl_xor(x[v], j);
          T mn = x[v] < p ? x[v] : p;
          T mx = x[v] < p ? p : x[v];
          x[v] = (lower == asc) ? mn : mx;
        }
      }
    }
  }
}

// A per-(split, query) slot in global memory of NS composites:
//   [0, KP)   the current best KP candidates, sorted ascending (lcnt valid)
//   [KP, NS)  the append buffer (bcnt valid)
// merge_slot sorts both parts together (one wave, all lanes active), keeps the
// KP smallest and returns the new list length; the new threshold is slot[KP-1]
// when the list is full.
template <typename T, int V>
__device__ __forceinline__ T wave_get(const T (&x)[V], int e);

template <typename T, int V>
__device__ __forceinline__ int merge_slot(T* __restrict__ slot, int KP, int lcnt, int bcnt, T& kth) {
  const int lane = __lane_id();
  constexpr T TMAX = ~T(0);
  T x[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    const bool valid = (e < lcnt) || (e >= KP && e < KP + bcnt);
    x[v] = valid ? slot[e] : TMAX;
  }
  wave_bitonic_sort<T, V>(x);
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    if (e < KP) slot[e] = x[v];
  }
  __threadfence_block();
  kth = wave_get<T, V>(x, KP - 1);
  int total = lcnt + bcnt;
  return total < KP ? total : KP;
}

// Two-segment variant: the append buffer is split between the two halves of a
// wave that share a query ([KP, KP+csub) and [KP+csub, KP+2*csub)).
template <typename T, int V>
__device__ __forceinline__ int merge_slot2(T* __restrict__ slot, int KP, int lcnt, int csub, int c0, int c1, T& kth) {
  const int lane = __lane_id();
  constexpr T TMAX = ~T(0);
  T x[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    const bool valid = (e < lcnt) || (e >= KP && e < KP + c0) || (e >= KP + csub && e < KP + csub + c1);
    x[v] = valid ? slot[e] : TMAX;
  }
  wave_bitonic_sort<T, V>(x);
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    if (e < KP) slot[e] = x[v];
  }
  __threadfence_block();
  kth = wave_get<T, V>(x, KP - 1);
  const int total = lcnt + c0 + c1;
  return total < KP ? total : KP;
}

template <typename T, int V>
__device__ __forceinline__ T wave_get(const T (&x)[V], int e) {
  // Broadcast element e (uniform) to all lanes.
  const int v = e >> 6;
  const int src = e & 63;
  T out = 0;
#pragma unroll
  for (int i = 0; i < V; ++i)
    if (i == v) {
      if constexpr (sizeof(T) == 8) {
        uint32_t lo = __builtin_amdgcn_readlane((int)uint32_t(x[i]), src);
        uint32_t hi = __builtin_amdgcn_readlane((int)uint32_t(uint64_t(x[i]) >> 32), src);
        out = (T(hi) << 32) | lo;
      } else {
        unsigned __int128 c = x[i];
        uint32_t w0 = __builtin_amdgcn_readlane((int)uint32_t(c), src);
        uint32_t w1 = __builtin_amdgcn_readlane((int)uint32_t(c >> 32), src);
        uint32_t w2 = __builtin_amdgcn_readlane((int)uint32_t(c >> 64), src);
        uint32_t w3 = __builtin_amdgcn_readlane((int)uint32_t(c >> 96), src);
        out = ((unsigned __int128)w3 << 96) | ((unsigned __int128)w2 << 64) |
              ((unsigned __int128)w1 << 32) | w0;
      }
    }
  return out;
}

// A float S with sqrt_rn(s) <= t  ==>  s <= S (t >= 0).  Lets scans filter in
// squared space before paying for a correctly rounded sqrt; the exact test on
// the sqrt'd key follows for every survivor, so S only has to be an upper bound.
// sqrt_rn(s) <= t implies s < (t + ulp(t))^2 <= t^2 (1 + 2^-22 + 2^-46), and
// S = t^2 (1 + 2^-20) rounded up covers it.  Negative or NaN t admits nothing.
__device__ __forceinline__ float sq_upper(float t) {
  if (!(t >= 0.0f)) return -INFINITY;
  if (!(t < INFINITY)) return INFINITY;
  const double s = (double)t * (double)t * (1.0 + 0x1p-20);
  float f = (float)s;
  if ((double)f < s) f = nextafterf(f, INFINITY);
  return f;
}
__device__ __forceinline__ double sq_upper(double t) {
  if (!(t >= 0.0)) return -INFINITY;
  if (!(t < INFINITY)) return INFINITY;
  return nextafter(t * t * (1.0 + 0x1p-48), INFINITY);
}

}  // namespace fa
