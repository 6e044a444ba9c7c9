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
      if (nl_ == KP && lane == j_) atomicMin(&a.gthr[q0 + (g) * 32 + j_], (uint32_t)(kth_ >> 32)); \
      if (col == j_) {                                                                    \
        cnt[g] = 0;                                                                       \
        lcnt[g] = nl_;                                                                    \
        thr[g] = nthr_;                                                                   \
      }                                                                                   \
      mask_ &= ~((1ull << j_) | (1ull << (j_ + 32)));                                     \
    }                                                                                     \
  } while (0)

// Asynchronous global -> LDS copy of 16 (or 4) bytes per lane into
// lds_base + 16*lane (4*lane).  Issued from inline assembly so that the compiler
// does not serialise later LDS reads of *other* buffers behind it; completion is
// awaited explicitly (s_waitcnt) before the barrier that publishes the buffer.
// M0 (the LDS base operand, latched at issue) is saved and restored.
__device__ __forceinline__ void lds_dma16(uint32_t lds_base, const void* src) {
  uint32_t tmp;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %1\n\t"
      "global_load_lds_dwordx4 %2, off\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(tmp)
      : "s"(lds_base), "v"(src)
      : "memory");
}
__device__ __forceinline__ void lds_dma4(uint32_t lds_base, const void* src) {
  uint32_t tmp;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %1\n\t"
      "global_load_lds_dword %2, off\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(tmp)
      : "s"(lds_base), "v"(src)
      : "memory");
}

template <int KIND, int KS, int A, int R, int V>
__device__ __forceinline__ void scan_mfma_body(const ScanMfmaArgs& a) {
  using C = Comp<float>;
  // bf16 rows carry -||x||^2/2 split into three bf16 columns (queries carry 1s
  // there), so the MFMA result is already acc = q.x - ||x||^2/2 = -key.
  constexpr bool FOLD = KIND == 0;
  constexpr int NS = 64 * V;
  constexpr int QPW = 32 * A;
  constexpr int QPB = 4 * QPW;
  constexpr int BN = 32 * R;
  constexpr int CH = KS * 2;                  // 16-byte chunks of a row
  constexpr int ROWC = CH + 1;                // + one pad chunk (bank spread)
  constexpr int ROWB = ROWC * 16;             // LDS row stride in bytes
  constexpr int TILE_CHUNKS = BN * ROWC;
  constexpr int WAVE_LOADS = (TILE_CHUNKS + 255) / 256;  // 64-chunk loads per wave per tile
  constexpr int TILE_BYTES = WAVE_LOADS * 256 * 16;       // rounded up to whole wave loads
  constexpr int NBUF = (2 * TILE_BYTES <= 65536) ? 2 : 1;
  __shared__ __attribute__((aligned(16))) char tile[NBUF][TILE_BYTES];
  // 64 floats per buffer: the hx DMA writes one dword per lane of a wave
  __shared__ __attribute__((aligned(16))) float tile_hx[FOLD ? 1 : NBUF][FOLD ? 4 : 64];

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
  const char* xop = reinterpret_cast<const char*>(a.Xop);

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
#pragma unroll
  for (int g = 0; g < A; ++g) {
    thr[g] = INFINITY;
    cnt[g] = 0;
    lcnt[g] = 0;
  }

  // Tile staging: lane i of DMA j of wave w fills LDS chunk p = (j*4 + w)*64 + i
  // = row*ROWC + c; the pad chunk (c == CH) and the round-up tail re-read a
  // valid chunk.  The per-lane source offsets do not depend on the tile.
  uint32_t soff[WAVE_LOADS];
#pragma unroll
  for (int j = 0; j < WAVE_LOADS; ++j) {
    const int p = (j * 4 + wave) * 64 + lane;
    int row = p / ROWC, c = p - row * ROWC;
    if (c >= CH) c = CH - 1;
    if (row >= BN) {
      row = BN - 1;
      c = CH - 1;
    }
    soff[j] = (uint32_t)(row * (KS * 32) + c * 16);
  }
  const uint32_t lds_tile = (uint32_t)(uintptr_t)&tile[0][0];
  const uint32_t lds_hx = (uint32_t)(uintptr_t)&tile_hx[0][0];
  auto stage = [&](int b, int64_t t) {
    const char* base = xop + size_t(t) * kp_bytes;
#pragma unroll
    for (int j = 0; j < WAVE_LOADS; ++j)
      lds_dma16(__builtin_amdgcn_readfirstlane(lds_tile + b * TILE_BYTES + (j * 4 + wave) * 1024), base + soff[j]);
    if (!FOLD && wave == 0) {
      const int r = lane < BN ? lane : BN - 1;
      lds_dma4(__builtin_amdgcn_readfirstlane(lds_hx + (NBUF == 2 ? b : 0) * 64 * 4), a.hx + t + r);
    }
  };

  int64_t it = 0;
  if (r0 < r1) stage(0, r0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (NBUF == 2 && r0 + BN < r1) stage(1, r0 + BN);
  int buf = 0;
  for (int64_t t = r0; t < r1; t += BN, ++it) {
    floatx16 acc[A][R];
#pragma unroll
    for (int g = 0; g < A; ++g)
#pragma unroll
      for (int rb = 0; rb < R; ++rb) acc[g][rb] = floatx16{};
    const char* tb = tile[NBUF == 2 ? buf : 0];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      uint4 af[R];
#pragma unroll
      for (int rb = 0; rb < R; ++rb)
        af[rb] = *reinterpret_cast<const uint4*>(tb + (rb * 32 + col) * ROWB + s * 32 + half * 16);
#pragma unroll
      for (int g = 0; g < A; ++g)
#pragma unroll
        for (int rb = 0; rb < R; ++rb) Mfma<KIND>::step(acc[g][rb], af[rb], qf[g][s]);
    }

    if constexpr (!FOLD) {
      // acc = q.x - hx = -key.  Register r of a 32x32 tile holds row
      // 8*(r/4) + 4*half + (r%4).
#pragma unroll
      for (int rb = 0; rb < R; ++rb)
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
          const float4 h4 =
              *reinterpret_cast<const float4*>(&tile_hx[FOLD ? 0 : (NBUF == 2 ? buf : 0)][rb * 32 + q4 * 8 + half * 4]);
          const float hv[4] = {h4.x, h4.y, h4.z, h4.w};
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int g = 0; g < A; ++g) acc[g][rb][q4 * 4 + u] = acc[g][rb][q4 * 4 + u] - hv[u];
        }
    }
    // Selection: one max-reduction per (query group, row block); only blocks
    // with a survivor in some lane take the append path.  The threshold is the
    // smaller of this split's KP-th key and the one other splits of the query
    // published: any split's KP-th key bounds the KP-th key of the union of
    // all splits' lists, so rows above it can never be needed.
#pragma unroll
    for (int g = 0; g < A; ++g) {
      const float nt = -thr[g];
      bool appended = false;
#pragma unroll
      for (int rb = 0; rb < R; ++rb) {
        float mx = acc[g][rb][0];
#pragma unroll
        for (int r = 1; r < 16; ++r) mx = fmaxf(mx, acc[g][rb][r]);
        if (__ballot(mx > nt)) {
          appended = true;
          uint64_t* sl = a.slots + (size_t(split) * a.nq_pad + q0 + g * 32 + col) * NS + KP + half * csub;
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            if (acc[g][rb][r] > nt) {
              const uint32_t row = (uint32_t)(t + rb * 32 + (r >> 2) * 8 + half * 4 + (r & 3));
              sl[cnt[g]] = C::make(-acc[g][rb][r], row);
              ++cnt[g];
            }
          }
        }
      }
      if (appended) MERGE_WHERE(g, cnt[g] > csub - 16 * R);
    }
    // Every 8 tiles, adopt the thresholds the query's other splits published
    // (read here, where no tile DMA is in flight, so the load does not hold up
    // the prefetch).
    auto refresh = [&]() {
      if ((it & 7) == 7) {
#pragma unroll
        for (int g = 0; g < A; ++g)
          thr[g] = fminf(thr[g], unord_f32(__hip_atomic_load(&a.gthr[q0 + g * 32 + col], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)));
      }
    };
    if constexpr (NBUF == 2) {
      // the next tile's DMA (issued one iteration ago) has landed for every
      // wave; this tile's buffer is free again and takes the tile after next
      __builtin_amdgcn_s_waitcnt(0);
      refresh();
      __syncthreads();
      if (t + 2 * BN < r1) stage(buf, t + 2 * BN);
      buf ^= 1;
    } else {
      refresh();
      __syncthreads();
      if (t + BN < r1) stage(0, t + BN);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
  }
#pragma unroll
  for (int g = 0; g < A; ++g) MERGE_WHERE(g, cnt[g] > 0);
  if (half == 0) {
#pragma unroll
    for (int g = 0; g < A; ++g) a.lcnts[size_t(split) * a.nq_pad + q0 + g * 32 + col] = lcnt[g];
  }
}

// Two entry points with different register budgets: shapes whose working set
// fits 256 registers run two waves per SIMD (one wave's MFMAs overlap the other's
// selection epilogue); wider shapes keep one wave per SIMD and no spills.
template <int KIND, int KS, int A, int R, int V>
__global__ __launch_bounds__(256, 2) void scan_mfma_kernel(ScanMfmaArgs a) {
  scan_mfma_body<KIND, KS, A, R, V>(a);
}
template <int KIND, int KS, int A, int R, int V>
__global__ __launch_bounds__(256) void scan_mfma_kernel_wide(ScanMfmaArgs a) {
  scan_mfma_body<KIND, KS, A, R, V>(a);
}

// K step buckets: bf16 covers 16 columns per step (d coordinates + the three
// norm columns), f32 covers 8.  Up to 16 steps a wave holds 2 query groups and
// the tile has 64 rows; beyond, 1 query group and 32-row tiles keep the query
// fragments at <= 224 registers and the two tile buffers within 64 KB of LDS.
static const int kBf16Steps[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 16, 17, 20, 24, 32, 40, 50, 56};
static const int kF32Steps[] = {1, 2, 3, 4, 6, 8, 12, 16, 20, 24, 32};

int mfma_ksteps_bucket(int kind, int d) {
  const int per = kind == 0 ? 16 : 8;
  const int k = kind == 0 ? d + 3 : d;  // bf16 rows carry three extra columns
  int need = (k + per - 1) / per;
  if (need < 1) need = 1;
  const int* tab = kind == 0 ? kBf16Steps : kF32Steps;
  const int nt = kind == 0 ? (int)(sizeof(kBf16Steps) / sizeof(int)) : (int)(sizeof(kF32Steps) / sizeof(int));
  for (int i = 0; i < nt; ++i)
    if (tab[i] >= need) return tab[i];
  return -1;
}

int mfma_queries_per_block(int kind, int ksteps) { return 4 * 32 * (ksteps <= 16 ? 2 : 1); }
int mfma_rows_per_tile(int kind, int ksteps) { return 32 * (ksteps <= 16 ? 2 : 1); }
int mfma_workgroups_per_cu(int kind, int ksteps) { return (kind == 0 ? ksteps <= 12 : ksteps <= 8) ? 2 : 1; }

template <int KIND, int KS>
static hipError_t launch_ks(const ScanMfmaArgs& a, int V, int nsplit, hipStream_t st) {
  constexpr int A = KS <= 16 ? 2 : 1;
  constexpr int R = KS <= 16 ? 2 : 1;
  constexpr bool two_waves = KIND == 0 ? KS <= 12 : KS <= 8;
  const dim3 grid((unsigned)(a.nqb * nsplit)), block(256);
  if constexpr (two_waves) {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_kernel<KIND, KS, A, R, 4>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((scan_mfma_kernel<KIND, KS, A, R, 8>), grid, block, 0, st, a);
  } else {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_kernel_wide<KIND, KS, A, R, 4>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((scan_mfma_kernel_wide<KIND, KS, A, R, 8>), grid, block, 0, st, a);
  }
  return hipGetLastError();
}

hipError_t launch_scan_mfma(const ScanMfmaArgs& a, int kind, int ks, int V, int nsplit, hipStream_t st) {
  if (V != 4 && V != 8) return hipErrorInvalidValue;
  if (kind == 0) {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_ks<0, K>(a, V, nsplit, st);
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8) KS_CASE(9)
      KS_CASE(10) KS_CASE(12) KS_CASE(14) KS_CASE(16) KS_CASE(17) KS_CASE(20) KS_CASE(24) KS_CASE(32)
      KS_CASE(40) KS_CASE(50) KS_CASE(56)
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
