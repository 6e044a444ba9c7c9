// Candidate generation on the matrix cores: key(q, x) = ||x||^2/2 - q.x, the
// query-independent part of ||q - x||^2 / 2, with the inner products on MFMA
// (bf16 x bf16 -> f32, or f32 x f32 -> f32), tiled through LDS, and a fused
// per-query selection that keeps the KP smallest keys of every DB split.  The
// distance matrix is never written.  Exactness is recovered afterwards by the
// rerank kernel (exact restated distances + a proven error bound).
//
// Layout (v_mfma_f32_32x32x16_bf16 / v_mfma_f32_32x32x2f32):
//   A operand = 32 DB rows of the tile, staged in LDS (row stride padded by 16
//   bytes so the 32 rows of a fragment read hit distinct banks);
//   B operand = 32 queries, held in registers for the whole scan;
//   C = 32x32 f32: lane l holds query (l & 31) against 16 DB rows
//       8*(r/4) + 4*(l/32) + (r%4), r = 0..15.
// Every lane-half h = l/32 supplies, per K step, the 16 bytes at byte offset
// 32*s + 16*h of its row (8 bf16, or 4 f32 feeding 4 chained 32x32x2 MFMAs);
// A and B use the same assignment, so the sums run over every coordinate once.
//
// Workgroup: 4 waves, each owning A*32 queries; the tile of R*32 DB rows is
// shared through LDS by the 4 waves.  blockIdx.x = split * nqb + query block.
//
// Selection: lanes compare key < thr(query) (strict; an excluded row has
// key >= the KP-th kept key); survivors are appended to the query's slot in
// HBM, one segment per lane-half, and merged by a wave-cooperative bitonic
// sort before a segment can overflow.
#include "common.h"
#include "kernels.h"

namespace fa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

template <int KIND>
struct Mfma;

template <>
struct Mfma<0> {  // bf16: one 32x32x16 MFMA per 16-byte operand
  __device__ static __forceinline__ void step(floatx16& acc, const uint4& a, const uint4& b) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                  0, 0, 0);
  }
};

template <>
struct Mfma<1> {  // f32: four chained 32x32x2 MFMAs per 16-byte operand
  __device__ static __forceinline__ void step(floatx16& acc, const uint4& a, const uint4& b) {
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.x), __uint_as_float(b.x), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.y), __uint_as_float(b.y), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.z), __uint_as_float(b.z), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.w), __uint_as_float(b.w), acc, 0, 0, 0);
  }
};

// Merge every query of group g whose lanes flag `need` (wave-uniform loop; the
// merge of query j resets the counters held by lanes j and j + 32).
#define MERGE_WHERE(g, need)                                                              \
  do {                                                                                    \
    uint64_t mask_ = __ballot(need);                                                      \
    while (mask_) {                                                                       \
      const int j_ = __builtin_ctzll(mask_) & 31;                                         \
      const int c0_ = __builtin_amdgcn_readlane(cnt[g], j_);                              \
      const int c1_ = __builtin_amdgcn_readlane(cnt[g], j_ + 32);                         \
      const int lc_ = __builtin_amdgcn_readlane(lcnt[g], j_);                             \
      uint64_t* sl_ = a.slots + (size_t(split) * a.nq_pad + q0 + (g) * 32 + j_) * NS;     \
      uint64_t kth_;                                                                      \
      const int nl_ = merge_slot2<uint64_t, V>(sl_, KP, lc_, csub, c0_, c1_, kth_);       \
      const float nthr_ = nl_ == KP ? C::key(kth_) : INFINITY;                            \
      if (col == j_) {                                                                    \
        cnt[g] = 0;                                                                       \
        lcnt[g] = nl_;                                                                    \
        thr[g] = nthr_;                                                                   \
      }                                                                                   \
      mask_ &= ~((1ull << j_) | (1ull << (j_ + 32)));                                     \
    }                                                                                     \
  } while (0)

template <int KIND, int KS, int A, int R, int V>
__global__ __launch_bounds__(256) void scan_mfma_kernel(ScanMfmaArgs a) {
  using C = Comp<float>;
  constexpr int NS = 64 * V;
  constexpr int QPW = 32 * A;
  constexpr int QPB = 4 * QPW;
  constexpr int BN = 32 * R;
  constexpr int ROWB = KS * 32 + 16;  // LDS row stride in bytes
  constexpr int CHUNKS = KS * 2;      // 16-byte chunks per row
  __shared__ __attribute__((aligned(16))) char tile[BN * ROWB];
  __shared__ __attribute__((aligned(16))) float tile_hx[BN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int half = lane >> 5;
  const int col = lane & 31;
  const int wave = tid >> 6;
  const int qb = (int)(blockIdx.x % (unsigned)a.nqb);
  const int split = (int)(blockIdx.x / (unsigned)a.nqb);
  const int64_t q0 = int64_t(qb) * QPB + wave * QPW;
  const int64_t r0 = int64_t(split) * a.rows_per_split;
  int64_t r1 = r0 + a.rows_per_split;
  if (r1 > a.n_pad) r1 = a.n_pad;
  const int KP = a.KP;
  const int csub = a.Csub;
  const size_t kp_bytes = size_t(KS) * 32;

  // query fragments, resident for the whole scan
  uint4 qf[A][KS];
#pragma unroll
  for (int g = 0; g < A; ++g) {
    const char* qrow = reinterpret_cast<const char*>(a.Qop) + size_t(q0 + g * 32 + col) * kp_bytes + half * 16;
#pragma unroll
    for (int s = 0; s < KS; ++s) qf[g][s] = *reinterpret_cast<const uint4*>(qrow + s * 32);
  }
  float thr[A];
  int cnt[A], lcnt[A];
  uint64_t* slot[A];
#pragma unroll
  for (int g = 0; g < A; ++g) {
    thr[g] = INFINITY;
    cnt[g] = 0;
    lcnt[g] = 0;
    slot[g] = a.slots + (size_t(split) * a.nq_pad + q0 + g * 32 + col) * NS;
  }

  for (int64_t t = r0; t < r1; t += BN) {
    __syncthreads();
    for (int c = tid; c < BN * CHUNKS; c += 256) {
      const int row = c / CHUNKS, part = c - row * CHUNKS;
      const uint4 v = *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(a.Xop) +
                                                      size_t(t + row) * kp_bytes + part * 16);
      *reinterpret_cast<uint4*>(tile + row * ROWB + part * 16) = v;
    }
    if (tid < BN) tile_hx[tid] = a.hx[t + tid];
    __syncthreads();

    floatx16 acc[A][R];
#pragma unroll
    for (int g = 0; g < A; ++g)
#pragma unroll
      for (int rb = 0; rb < R; ++rb) acc[g][rb] = floatx16{};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      uint4 af[R];
#pragma unroll
      for (int rb = 0; rb < R; ++rb)
        af[rb] = *reinterpret_cast<const uint4*>(tile + (rb * 32 + col) * ROWB + s * 32 + half * 16);
#pragma unroll
      for (int g = 0; g < A; ++g)
#pragma unroll
        for (int rb = 0; rb < R; ++rb) Mfma<KIND>::step(acc[g][rb], af[rb], qf[g][s]);
    }

    // key = hx - q.x ; row of register r: 8*(r/4) + 4*half + (r%4)
    float hv[R][16];
#pragma unroll
    for (int rb = 0; rb < R; ++rb)
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4) {
        const float4 h4 = *reinterpret_cast<const float4*>(tile_hx + rb * 32 + q4 * 8 + half * 4);
        hv[rb][q4 * 4 + 0] = h4.x;
        hv[rb][q4 * 4 + 1] = h4.y;
        hv[rb][q4 * 4 + 2] = h4.z;
        hv[rb][q4 * 4 + 3] = h4.w;
      }
    bool any = false;
#pragma unroll
    for (int g = 0; g < A; ++g)
#pragma unroll
      for (int rb = 0; rb < R; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) any |= (hv[rb][r] - acc[g][rb][r]) < thr[g];
    if (__ballot(any)) {
#pragma unroll
      for (int g = 0; g < A; ++g) {
#pragma unroll
        for (int rb = 0; rb < R; ++rb)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const float key = hv[rb][r] - acc[g][rb][r];
            if (key < thr[g]) {
              const uint32_t row = (uint32_t)(t + rb * 32 + (r >> 2) * 8 + half * 4 + (r & 3));
              slot[g][KP + half * csub + cnt[g]] = C::make(key, row);
              ++cnt[g];
            }
          }
        MERGE_WHERE(g, cnt[g] > csub - 16 * R);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < A; ++g) MERGE_WHERE(g, cnt[g] > 0);
  if (half == 0) {
#pragma unroll
    for (int g = 0; g < A; ++g) a.lcnts[size_t(split) * a.nq_pad + q0 + g * 32 + col] = lcnt[g];
  }
}

// K step buckets: bf16 covers 16 coordinates per step, f32 covers 8.  Up to 16
// steps a wave holds 2 query groups and the tile has 64 rows; beyond, 1 query
// group and 32-row tiles keep the query fragments at <= 224 registers and the
// tile within 64 KB of LDS.
static const int kBf16Steps[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 32, 40, 49, 56};
static const int kF32Steps[] = {1, 2, 3, 4, 6, 8, 12, 16, 20, 24, 32};

int mfma_ksteps_bucket(int kind, int d) {
  const int per = kind == 0 ? 16 : 8;
  int need = (d + per - 1) / per;
  if (need < 1) need = 1;
  const int* tab = kind == 0 ? kBf16Steps : kF32Steps;
  const int nt = kind == 0 ? (int)(sizeof(kBf16Steps) / sizeof(int)) : (int)(sizeof(kF32Steps) / sizeof(int));
  for (int i = 0; i < nt; ++i)
    if (tab[i] >= need) return tab[i];
  return -1;
}

int mfma_queries_per_block(int kind, int ksteps) { return 4 * 32 * (ksteps <= 16 ? 2 : 1); }
int mfma_rows_per_tile(int kind, int ksteps) { return 32 * (ksteps <= 16 ? 2 : 1); }

template <int KIND, int KS>
static hipError_t launch_ks(const ScanMfmaArgs& a, int V, int nsplit, hipStream_t st) {
  constexpr int A = KS <= 16 ? 2 : 1;
  constexpr int R = KS <= 16 ? 2 : 1;
  const dim3 grid((unsigned)(a.nqb * nsplit)), block(256);
  if (V == 4)
    hipLaunchKernelGGL((scan_mfma_kernel<KIND, KS, A, R, 4>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((scan_mfma_kernel<KIND, KS, A, R, 8>), grid, block, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_scan_mfma(const ScanMfmaArgs& a, int kind, int ks, int V, int nsplit, hipStream_t st) {
  if (V != 4 && V != 8) return hipErrorInvalidValue;
  if (kind == 0) {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_ks<0, K>(a, V, nsplit, st);
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8) KS_CASE(10)
      KS_CASE(12) KS_CASE(14) KS_CASE(16) KS_CASE(20) KS_CASE(24) KS_CASE(32) KS_CASE(40) KS_CASE(49)
      KS_CASE(56)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_ks<1, K>(a, V, nsplit, st);
    KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(6) KS_CASE(8) KS_CASE(12) KS_CASE(16) KS_CASE(20)
    KS_CASE(24) KS_CASE(32)
#undef KS_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace fa
