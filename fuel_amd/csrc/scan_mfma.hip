// Candidate generation on the matrix cores: key(q, x) = ||x||^2/2 - q.x, the
// part of ||q - x||^2 / 2 that varies with x, with the inner products on MFMA
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
//   The C input of the first MFMA of a tile is hxn = -fl32(||x||^2/2) of those
//   rows (staged beside the tile), so the result is acc = q.x - ||x||^2/2 = -key
//   with no extra pass over the accumulators.
// Every lane-half h = l/32 supplies, per K step, the 16 bytes at byte offset
// 32*s + 16*h of its row (8 bf16, or 4 f32 feeding 4 chained 32x32x2 MFMAs);
// A and B use the same assignment, so the sums run over every coordinate once.
//
// Workgroup: 4 waves (1 for small residual batches), each owning A*32 queries;
// the tile of R*32 DB rows is shared through LDS by the waves, with one barrier
// per tile, and taken in chunks of RC 32-row blocks whose accumulators are the
// only live ones; two or three tile buffers, the global->LDS copy of tile t+2
// issued as soon as its buffer is free.
//
// Selection: a row block is skipped when no lane's maximum beats its query's
// threshold (strict: an excluded row has key >= the KP-th kept key).  Otherwise
// the 4-bit row position is written into the low mantissa bits (a perturbation
// of at most 15 ulps, covered by the rerank's error bound), so one max tree
// yields the best value and its row; survivors are appended to the query's slot
// in HBM, one segment per lane-half, peeling one per lane per step.  Before a
// segment can overflow, the query's kept list and both segments are merged by a
// radix select (no sort): a 32-step ballot/popcount search finds the KP-th key,
// the keys below it (and enough equal ones) become the new kept list, and the
// KP-th key is the new threshold.  Thresholds are shared by the splits of a
// query through a per-query atomicMin; every tile's copy brings the values the
// other splits published.
//
// E4M3 operand (KIND 2, residual only): rows hold the E4M3 codes of r 2^-e_x
// with a per-row E8M0 scale 2^e_x staged beside the tile; the query side is
// q - c_j split into hi 2^e_q + lo 2^(e_q - 4), both E4M3.  Two 32-byte steps
// form one v_mfma_scale_f32_32x32x64_f8f6f4 operand (the lane half supplies
// its 16 bytes of each), issued once against hi and once against lo with the
// row's and the query's scales: lanes l and l + 32 carry the same scale, so
// the product is sa sb sum a b whatever K block a byte belongs to.
#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "fp8.h"
#include "kernels.h"

namespace fa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));
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

__device__ __forceinline__ uint16_t bf16_bits_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ int lane_prefix(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0));
}

// Merge of one query's slot (one wave): kept list slot[0, lcnt) and the two
// append segments slot[KP, KP + c0), slot[KP + csub, KP + csub + c1).  When the
// total reaches KP, t = the KP-th smallest key (high words, order-preserving
// bits) is found by a 32-step search for the smallest t with #(key <= t) >= KP;
// slot[0, KP) then receives every key < t and as many keys == t as fit.  The
// list is left unsorted.  Returns the new kept count; t_out = t (all-ones when
// fewer than KP keys exist).
template <int V>
__device__ __forceinline__ int merge_select(uint64_t* __restrict__ slot, int KP, int lcnt, int csub, int c0, int c1,
                                            uint32_t& t_out) {
  const int lane = __lane_id();
  uint64_t x[V];
  uint64_t vm[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    const bool valid = (e < lcnt) || (e >= KP && e < KP + c0) || (e >= KP + csub && e < KP + csub + c1);
    x[v] = valid ? slot[e] : ~0ull;
    vm[v] = __ballot(valid);
  }
  const int total = lcnt + c0 + c1;
  uint32_t t = 0xFFFFFFFFu;
  if (total >= KP) {
    uint32_t p = 0;
    for (int b = 31; b >= 0; --b) {
      const uint32_t probe = p | ((1u << b) - 1u);
      int c = 0;
#pragma unroll
      for (int v = 0; v < V; ++v) c += __popcll(__ballot(uint32_t(x[v] >> 32) <= probe) & vm[v]);
      if (c < KP) p |= (1u << b);
    }
    t = p;
  }
  // all loads above have been consumed; the stores below may overwrite slot
  int base = 0;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const bool f = ((vm[v] >> lane) & 1) && (total < KP || uint32_t(x[v] >> 32) < t);
    const uint64_t m = __ballot(f);
    if (f) slot[base + lane_prefix(m)] = x[v];
    base += __popcll(m);
  }
  if (total >= KP) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const bool f = ((vm[v] >> lane) & 1) && uint32_t(x[v] >> 32) == t;
      const uint64_t m = __ballot(f);
      const int pos = base + lane_prefix(m);
      if (f && pos < KP) slot[pos] = x[v];
      base += __popcll(m);
    }
  }
  t_out = t;
  return total < KP ? total : KP;
}

// Merge every query of group g whose lanes flag `need` (wave-uniform loop; the
// merge of query j resets the counters held by lanes j and j + 32).  The slot
// stores and loads of one wave are ordered by the hardware (wavefront scope),
// so the appends of this tile are visible to the merge's loads.
#define MERGE_WHERE(g, need)                                                              \
  do {                                                                                    \
    uint64_t mask_ = __ballot(need);                                                      \
    while (mask_) {                                                                       \
      const int j_ = __builtin_ctzll(mask_) & 31;                                         \
      const int c0_ = __builtin_amdgcn_readlane(cnt[g], j_);                              \
      const int c1_ = __builtin_amdgcn_readlane(cnt[g], j_ + 32);                         \
      const int lc_ = __builtin_amdgcn_readlane(lcnt[g], j_);                             \
      uint64_t* sl_ = a.slots + scan_slot_index(split, a.nq_pad, q0 + (g) * 32 + j_, NS, 0); \
      uint32_t kt_;                                                                       \
      const int nl_ = merge_select<V>(sl_, KP, lc_, csub, c0_, c1_, kt_);                 \
      const float nthr_ = nl_ == KP ? unord_f32(kt_) : MFMA_KEY_LIMIT;                    \
      if (nl_ == KP && lane == j_) atomicMin(&a.gthr[q0 + (g) * 32 + j_], kt_);           \
      if (col == j_) {                                                                    \
        cnt[g] = 0;                                                                       \
        lcnt[g] = nl_;                                                                    \
        thr[g] = nthr_;                                                                   \
      }                                                                                   \
      mask_ &= ~((1ull << j_) | (1ull << (j_ + 32)));                                     \
    }                                                                                     \
  } while (0)

// Tile geometry of a shape (KS K steps, R 32-row blocks per tile, W waves):
// each row takes 2 KS 16-byte chunks plus one pad chunk (bank spread); a tile
// is copied by whole 64-lane loads of every wave, so its LDS size is rounded
// up to them.  The scan body and the launch planning both use these.
constexpr int scan_wave_loads(int ks, int R, int W) { return (32 * R * (2 * ks + 1) + 64 * W - 1) / (64 * W); }
constexpr int scan_tile_bytes(int ks, int R, int W) { return scan_wave_loads(ks, R, W) * 64 * W * 16; }
// C inputs of a tile: one float per row, at least 64 per buffer, copied 64 at a
// time by the first (32 R + 63) / 64 waves.
constexpr int scan_hx_words(int R) { return 32 * R < 64 ? 64 : 32 * R; }
constexpr int scan_hx_waves(int R) { return (32 * R + 63) / 64; }
// LDS of one workgroup: nb tile buffers, their C inputs and the published
// thresholds (64 words per wave and buffer), plus the persistent form's work
// item.
constexpr int scan_lds_bytes(int ks, int R, int W, bool batch, int nb = 2) {
  return nb * scan_tile_bytes(ks, R, W) + nb * scan_hx_words(R) * 4 + nb * W * 64 * 4 + (batch ? 16 : 0);
}
// per buffer, 64 words of the E4M3 operand's row scales and (residual operand)
// 64 copies of the tile's list
constexpr int scan_lds_extra(int kind, int nb = 2, bool resid = false) {
  return nb * 64 * 4 * ((kind == 2 ? 1 : 0) + (resid ? 1 : 0));
}
constexpr int kLdsPerCu = 160 * 1024;  // gfx950
constexpr int kLdsGranule = 512;       // allocation granularity per workgroup

// Shape of a K step count: query groups per wave (A), 32-row blocks per tile
// (R), two waves per SIMD or one.  Plain bf16 shapes of 14 and 16 steps take
// 32-row tiles so that two waves per SIMD fit (their query fragments take 128
// registers); beyond 16 steps one query group per wave.
// The residual variants keep 64-row tiles and one wave per SIMD from 14 steps
// (their list-change rebuild of the fragments leaves no room for two).
constexpr int scan_groups(int ks) { return ks <= 16 ? 2 : 1; }
// (E4M3, kind 2: 64-row tiles; two waves per SIMD up to 4 32-byte steps, whose
// hi and lo fragments then fit without spills.)
constexpr bool scan_two_waves(int kind, int ks, bool resid) {
  return kind == 0 ? ks <= (resid ? 12 : 16) : (kind == 1 ? ks <= 8 : ks <= 4);
}
// Shapes whose fragments fit 128 registers run four workgroups (four waves per
// SIMD): at few K steps the selection, not the MFMA, bounds the scan, and more
// waves hide its latency.  They keep the accumulators of one 32-row block at a
// time, so that the 128 registers hold them without spills, and their tiles
// hold 4 such blocks up to 3 K steps, else 2 (whose LDS then still leaves four
// workgroups resident; residual: 2, one list per 64-row block), so that the
// barrier and the copy issue of a tile are shared by several blocks' MFMAs and
// selection.
constexpr bool scan_four_wg(int kind, int ks) { return kind == 0 ? ks <= 4 : (kind == 1 && ks <= 2); }
// Plain bf16 shapes of 5-8 K steps take 128-row tiles in two chunks of two
// blocks, those of 14-16 steps 64-row tiles in two chunks of one (one barrier
// per tile; two tile buffers then still leave two workgroups resident).
constexpr bool scan_wide_tile(int kind, int ks, bool resid) {
  return kind == 0 && !resid && ((ks >= 5 && ks <= 8) || (ks >= 14 && ks <= 16));
}
constexpr int scan_row_blocks(int kind, int ks, bool resid) {
  return scan_four_wg(kind, ks)            ? (resid || ks > 3 ? 2 : 4)
         : scan_wide_tile(kind, ks, resid) ? (ks <= 8 ? 4 : 2)
                                           : ((kind == 0 && !resid ? ks <= 12 : ks <= 16) ? 2 : 1);
}
// Row blocks whose accumulators are live at once (R for the other shapes).
constexpr int scan_row_chunk(int kind, int ks, bool resid) {
  return scan_four_wg(kind, ks)              ? 1
         : scan_wide_tile(kind, ks, resid) ? (ks <= 8 ? 2 : 1)
                                             : scan_row_blocks(kind, ks, resid);
}

// Tile buffers of a shape: three (the copy of tile t + 2 leaves when tile t
// starts, so two copies overlap the reads of one) when the larger LDS share
// leaves as many workgroups resident as two buffers do (the resident count by
// registers is 4 for one-wave workgroups, whose 270-390 registers leave one
// wave per SIMD, and the launch bound otherwise); else two (the copy of tile
// t + 2 leaves when tile t's buffer is free).  Fewer workgroups with deeper
// copies measured slower (one-wave bf16 residual scan of 10M x 96: 0.53 ms with
// four workgroups per CU and two buffers, 0.66 ms with three and three).
constexpr int scan_wgs_by_lds(int lds) { return kLdsPerCu / ((lds + kLdsGranule - 1) / kLdsGranule * kLdsGranule); }
constexpr int scan_buffers(int kind, int ks, int R, int W, bool resid) {
  const int per = W == 1 ? 4 : (scan_four_wg(kind, ks) ? 4 : (scan_two_waves(kind, ks, resid) ? 2 : 1));
  const int n2 = std::min(per, scan_wgs_by_lds(scan_lds_bytes(ks, R, W, true, 2) + scan_lds_extra(kind, 2, resid)));
  const int n3 = std::min(per, scan_wgs_by_lds(scan_lds_bytes(ks, R, W, true, 3) + scan_lds_extra(kind, 3, resid)));
  return n3 >= n2 && n3 > 0 ? 3 : 2;
}
// s_waitcnt vmcnt(n) with the export and LDS counters left alone (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt has 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}


// Tile staging: lane i of copy j of wave w fills LDS chunk p = (j W + w) 64 + i
// = row (2 KS + 1) + c of the tile; the pad chunk (c == 2 KS) and the round-up
// tail re-read a valid chunk.  Returns the lane's source byte offset from the
// tile's first operand row (rows of KS * 32 bytes).
__host__ __device__ __forceinline__ uint32_t scan_stage_src(int ks, int R, int W, int j, int wave, int lane) {
  const int CH = 2 * ks, ROWC = CH + 1, BN = 32 * R;
  const int p = (j * W + wave) * 64 + lane;
  int row = p / ROWC, c = p - row * ROWC;
  if (c >= CH) c = CH - 1;
  if (row >= BN) {
    row = BN - 1;
    c = CH - 1;
  }
  return (uint32_t)(row * (ks * 32) + c * 16);
}
// First query slot of a wave, and the row range [r0, r1) of a split.
__host__ __device__ __forceinline__ int64_t scan_wave_q0(int qb, int wave, int A, int W) {
  return int64_t(qb) * (W * 32 * A) + wave * (32 * A);
}
__host__ __device__ __forceinline__ void scan_split_rows(int split, int64_t rps, int64_t n_pad, int64_t& r0,
                                                         int64_t& r1) {
  r0 = int64_t(split) * rps;
  r1 = r0 + rps;
  if (r1 > n_pad) r1 = n_pad;
}
// Element index of entry e of query slot q's list in split `split`.
__host__ __device__ __forceinline__ size_t scan_slot_index(int split, int64_t nq_pad, int64_t q, int NS, int e) {
  return (size_t(split) * nq_pad + q) * NS + e;
}

template <int KIND, int KS, int A, int R, int V, bool RESID = false, int W = 4>
__device__ __forceinline__ void scan_mfma_body(const ScanMfmaArgs& a, const int qb, const int split) {
  static_assert(KIND != 2 || (RESID && KS % 2 == 0), "the E4M3 operand is residual, in whole 64-byte steps");
  using C = Comp<float>;
  constexpr int NS = 64 * V;
  constexpr int QPW = 32 * A;
  constexpr int QPB = W * QPW;  // W waves per workgroup (4; 1 for small residual batches)
  constexpr int BN = 32 * R;
  constexpr int CH = KS * 2;                  // 16-byte chunks of a row
  constexpr int ROWC = CH + 1;                // + one pad chunk (bank spread)
  constexpr int ROWB = ROWC * 16;             // LDS row stride in bytes
  constexpr int WAVE_LOADS = scan_wave_loads(KS, R, W);  // 64-chunk loads per wave per tile
  constexpr int TILE_BYTES = scan_tile_bytes(KS, R, W);  // rounded up to whole wave loads
  static_assert(WAVE_LOADS * 64 * W >= BN * ROWC, "tile copy covers the tile");
  constexpr int NB = scan_buffers(KIND, KS, R, W, RESID);
  constexpr int RC = scan_row_chunk(KIND, KS, RESID);
  static_assert(R % RC == 0, "whole chunks of row blocks");
  // vector memory loads one stage() issues in every wave: the tile chunks and
  // the thresholds (the first HXW waves add the C inputs, wave 0 the E4M3 row
  // scales and list id, so this is the smallest count of any wave)
  constexpr int STAGE_OPS = WAVE_LOADS + 1;
  constexpr int HXN = scan_hx_words(R), HXW = scan_hx_waves(R);
  static_assert(HXW <= W, "every C input has a wave that copies it");
  static_assert(!RESID || BN <= 64, "a residual tile lies in one 64-row block, hence in one list");
  __shared__ __attribute__((aligned(16))) char tile[NB][TILE_BYTES];
  __shared__ __attribute__((aligned(16))) float tile_hx[NB][HXN];  // one dword per lane of the first HXW waves
  __shared__ __attribute__((aligned(16))) uint32_t thr_lds[NB][W][64];  // published thresholds, per wave
  __shared__ __attribute__((aligned(16))) uint32_t tile_sc[NB][KIND == 2 ? 64 : 1];  // E8M0 row scales (E4M3)
  // residual operand: the tile's list, copied with the tile, so that reading it
  // waits on the LDS counter only (a vector load here would wait for the copy
  // of the next tile too)
  __shared__ __attribute__((aligned(16))) uint32_t tile_li[NB][RESID ? 64 : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int half = lane >> 5;
  const int col = lane & 31;
  const int wave = tid >> 6;
  const int64_t q0 = scan_wave_q0(qb, wave, A, W);
  int64_t r0, r1;
  scan_split_rows(split, a.rows_per_split, a.n_pad, r0, r1);
  const int KP = a.KP;
  const int csub = a.Csub;
  const size_t kp_bytes = size_t(KS) * 32;
  const char* xop = reinterpret_cast<const char*>(a.Xop);

  // query fragments, resident for the whole scan (residual operand: rebuilt at
  // every change of list, below)
  uint4 qf[A][KS];
  uint4 qlo[KIND == 2 ? A : 1][KIND == 2 ? KS : 1];  // E4M3: the lo part of q - c_j
  int qsh[A], qsl[A];                                 // E4M3: E8M0 codes of the hi and lo scales
  float qt[A];
#pragma unroll
  for (int g = 0; g < A; ++g) {
    qt[g] = 0.0f;
    qsh[g] = qsl[g] = 127;
  }
  int cur_list = -1;
  if constexpr (!RESID) {
#pragma unroll
    for (int g = 0; g < A; ++g) {
      const char* qrow = reinterpret_cast<const char*>(a.Qop) + size_t(q0 + g * 32 + col) * kp_bytes + half * 16;
#pragma unroll
      for (int s = 0; s < KS; ++s) qf[g][s] = *reinterpret_cast<const uint4*>(qrow + s * 32);
    }
  }
  float thr[A];
  int cnt[A], lcnt[A];
#pragma unroll
  for (int g = 0; g < A; ++g) {
    // a caller may start a query at a proven threshold (the certified second pass)
    thr[g] = fminf(MFMA_KEY_LIMIT, unord_f32(a.gthr[q0 + g * 32 + col]));
    cnt[g] = 0;
    lcnt[g] = 0;
  }

  // Tile staging (scan_stage_src): the per-lane source offsets do not depend on
  // the tile.
  uint32_t soff[WAVE_LOADS];
#pragma unroll
  for (int j = 0; j < WAVE_LOADS; ++j) soff[j] = scan_stage_src(KS, R, W, j, wave, lane);
  const uint32_t lds_tile = (uint32_t)(uintptr_t)&tile[0][0];
  const uint32_t lds_hx = (uint32_t)(uintptr_t)&tile_hx[0][0];
  const uint32_t lds_thr = (uint32_t)(uintptr_t)&thr_lds[0][0][0];
  const uint32_t lds_sc = (uint32_t)(uintptr_t)&tile_sc[0][0];
  const uint32_t lds_li = (uint32_t)(uintptr_t)&tile_li[0][0];
  const uint32_t* gthr_src = a.gthr + q0 + (lane < QPW ? lane : QPW - 1);
  auto stage = [&](int b, int64_t t) {
    const char* base = xop + size_t(t) * kp_bytes;
#pragma unroll
    for (int j = 0; j < WAVE_LOADS; ++j)
      lds_dma16(__builtin_amdgcn_readfirstlane(lds_tile + b * TILE_BYTES + (j * W + wave) * 1024), base + soff[j]);
    if (wave < HXW) {
      const int r = wave * 64 + lane < BN ? wave * 64 + lane : BN - 1;
      lds_dma4(__builtin_amdgcn_readfirstlane(lds_hx + (b * HXN + wave * 64) * 4), a.hx + t + r);
    }
    if (wave == 0) {
      if constexpr (KIND == 2) {
        // the BN scale bytes of the tile, a word per lane (lanes past BN/4 repeat the last)
        const int w = lane < BN / 4 ? lane : BN / 4 - 1;
        lds_dma4(__builtin_amdgcn_readfirstlane(lds_sc + b * 64 * 4), reinterpret_cast<const uint32_t*>(a.xsc + t) + w);
      }
      if constexpr (RESID)  // every lane copies the same word: the list of the tile's 64-row block
        lds_dma4(__builtin_amdgcn_readfirstlane(lds_li + b * 64 * 4),
                 reinterpret_cast<const uint32_t*>(a.tile_list) + (t >> 6));
    }
    // the thresholds other splits published for this wave's queries ride along
    lds_dma4(__builtin_amdgcn_readfirstlane(lds_thr + (b * W + wave) * 256), gthr_src);
  };

  // Loads of one wave complete in the order they were issued, so when at most
  // STAGE_OPS loads are outstanding, every stage before the last one has
  // landed, whatever stores (appends, merges) are still pending: those only add
  // to the count.  (Wave 0 may wait for a few chunks of the last stage too.)
  auto wait_all_but_last_stage = [&]() { wait_vmcnt<STAGE_OPS>(); };
  if (r0 < r1) stage(0, r0);
  if constexpr (NB == 3) {
    if (r0 + BN < r1) {
      stage(1, r0 + BN);
      wait_all_but_last_stage();
    } else {
      __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();
  } else {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (r0 + BN < r1) stage(1, r0 + BN);
  }
  int buf = 0;
  for (int64_t t = r0; t < r1; t += BN) {
    // three buffers: the one of tile t - BN is free (every wave passed the last
    // barrier), and takes tile t + 2 BN now
    if constexpr (NB == 3) {
      if (t + 2 * BN < r1) stage(buf == 0 ? 2 : buf - 1, t + 2 * BN);
    }
    const char* tb = tile[buf];
    const float* hb = tile_hx[buf];
    if constexpr (RESID) {
      const int jl = __builtin_amdgcn_readfirstlane((int)tile_li[buf][0]);
      if (jl != cur_list) {
        // B operand of this list: bf16(fl32(q - c_j)) of the lane's 8 coordinates
        // per K step (16 s + 8 half ...), zero beyond d; padding query slots
        // repeat the last query (their results are never read)
        cur_list = jl;
        const float* cj = a.cents + size_t(jl) * a.ldc;
        if constexpr (KIND == 2) {
          // E4M3 fragments of this list, prepared per (list, query slot) by
          // prep_qfrag_fp8_kernel: hi codes, then lo codes, in the operand's byte
          // layout, and the E8M0 code of the hi scale (lo: 2^-4 of it)
#pragma unroll
          for (int g = 0; g < A; ++g) {
            const int64_t qs = q0 + g * 32 + col;
            const size_t fi = size_t(jl) * a.nq_pad + qs;
            const char* fh = reinterpret_cast<const char*>(a.qfrag) + fi * kp_bytes + half * 16;
            const char* fl = fh + size_t(a.L) * a.nq_pad * kp_bytes;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
              qf[g][s] = *reinterpret_cast<const uint4*>(fh + s * 32);
              qlo[g][s] = *reinterpret_cast<const uint4*>(fl + s * 32);
            }
            qsh[g] = a.qsc[fi];
            qsl[g] = qsh[g] - 4;
            qt[g] = a.qtab[fi];
          }
        } else {
#pragma unroll
        for (int g = 0; g < A; ++g) {
          const int64_t qs = q0 + g * 32 + col;
          const float* qv = a.Q32 + (qs < a.m ? qs : a.m - 1) * a.ld32;
#pragma unroll
          for (int s = 0; s < KS; ++s) {
            const int k0 = s * 16 + half * 8;
            float v[8];
            if (KS >= 14 && k0 + 8 <= a.d && (a.ldc & 3) == 0) {
              // rows of Q32 are 32-byte aligned and k0 is a multiple of 8; centroid
              // rows are 16-byte aligned when their pitch is a multiple of 4
              const float4 q0 = *reinterpret_cast<const float4*>(qv + k0);
              const float4 q1 = *reinterpret_cast<const float4*>(qv + k0 + 4);
              const float4 c0 = *reinterpret_cast<const float4*>(cj + k0);
              const float4 c1 = *reinterpret_cast<const float4*>(cj + k0 + 4);
              v[0] = q0.x - c0.x; v[1] = q0.y - c0.y; v[2] = q0.z - c0.z; v[3] = q0.w - c0.w;
              v[4] = q1.x - c1.x; v[5] = q1.y - c1.y; v[6] = q1.z - c1.z; v[7] = q1.w - c1.w;
            } else {
#pragma unroll
              for (int e = 0; e < 8; ++e) v[e] = k0 + e < a.d ? qv[k0 + e] - cj[k0 + e] : 0.0f;
            }
            uint32_t wv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
              wv[e] = uint32_t(bf16_bits_rne(v[2 * e])) | (uint32_t(bf16_bits_rne(v[2 * e + 1])) << 16);
            qf[g][s] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
          }
          qt[g] = a.qtab[size_t(jl) * a.nq_pad + qs];
        }
        }
      }
    }
    // Row blocks c0 .. c0 + RC - 1 of the tile at a time (RC = R but for the
    // shapes of few K steps, whose tiles hold more row blocks than they keep
    // accumulators for).
    // (unrolled: same-box A/B, 0.7 % on the 8-step scan, 2 % on the batched
    // d = 256 problems, whose 64-row tiles were slower than 32-row ones when
    // rolled, and 1 % on the whole batched pass for the four-workgroup shapes)
#pragma unroll(R / RC)
    for (int c0 = 0; c0 < R; c0 += RC) {
    // C input: hxn of the 16 rows this lane holds, shared by the query groups
    floatx16 acc[A][RC];
#pragma unroll
    for (int rc = 0; rc < RC; ++rc) {
      const int rb = c0 + rc;
      floatx16 h;
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4) {
        const float4 h4 = *reinterpret_cast<const float4*>(hb + rb * 32 + q4 * 8 + half * 4);
        h[q4 * 4 + 0] = h4.x;
        h[q4 * 4 + 1] = h4.y;
        h[q4 * 4 + 2] = h4.z;
        h[q4 * 4 + 3] = h4.w;
      }
      if constexpr (RESID) {
#pragma unroll
        for (int g = 0; g < A; ++g)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[g][rc][r] = h[r] + qt[g];
      } else {
#pragma unroll
        for (int g = 0; g < A; ++g) acc[g][rc] = h;
      }
    }
    if constexpr (KIND == 2) {
      const uint8_t* scb = reinterpret_cast<const uint8_t*>(&tile_sc[buf][0]);
      int sa[RC];
#pragma unroll
      for (int rc = 0; rc < RC; ++rc) sa[rc] = scb[(c0 + rc) * 32 + col];
#pragma unroll
      for (int p = 0; p < KS / 2; ++p) {
        intx8 a8[RC];
#pragma unroll
        for (int rc = 0; rc < RC; ++rc) {
          const char* rp = tb + ((c0 + rc) * 32 + col) * ROWB + p * 64 + half * 16;
          const uint4 x0 = *reinterpret_cast<const uint4*>(rp);
          const uint4 x1 = *reinterpret_cast<const uint4*>(rp + 32);
          a8[rc] = intx8{(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
        }
#pragma unroll
        for (int g = 0; g < A; ++g) {
          const uint4 h0 = qf[g][2 * p], h1 = qf[g][2 * p + 1];
          const intx8 bh{(int)h0.x, (int)h0.y, (int)h0.z, (int)h0.w, (int)h1.x, (int)h1.y, (int)h1.z, (int)h1.w};
#pragma unroll
          for (int rc = 0; rc < RC; ++rc)
            acc[g][rc] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8[rc], bh, acc[g][rc], 0, 0, 0, sa[rc], 0,
                                                                         qsh[g]);
        }
#pragma unroll
        for (int g = 0; g < A; ++g) {
          const uint4 l0 = qlo[g][2 * p], l1 = qlo[g][2 * p + 1];
          const intx8 bl{(int)l0.x, (int)l0.y, (int)l0.z, (int)l0.w, (int)l1.x, (int)l1.y, (int)l1.z, (int)l1.w};
#pragma unroll
          for (int rc = 0; rc < RC; ++rc)
            acc[g][rc] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8[rc], bl, acc[g][rc], 0, 0, 0, sa[rc], 0,
                                                                         qsl[g]);
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        uint4 af[RC];
#pragma unroll
        for (int rc = 0; rc < RC; ++rc)
          af[rc] = *reinterpret_cast<const uint4*>(tb + ((c0 + rc) * 32 + col) * ROWB + s * 32 + half * 16);
#pragma unroll
        for (int g = 0; g < A; ++g)
#pragma unroll
          for (int rc = 0; rc < RC; ++rc) Mfma<KIND>::step(acc[g][rc], af[rc], qf[g][s]);
      }
    }

#pragma unroll
    for (int g = 0; g < A; ++g) {
      const float nt = -thr[g];
      bool appended = false;
#pragma unroll
      for (int rc = 0; rc < RC; ++rc) {
        const int rb = c0 + rc;
        float mx = acc[g][rc][0];
#pragma unroll
        for (int r = 1; r < 16; ++r) mx = fmaxf(mx, acc[g][rc][r]);
        if (__ballot(mx > nt)) {
          appended = true;
          uint64_t* sl = a.slots + scan_slot_index(split, a.nq_pad, q0 + g * 32 + col, NS, KP + half * csub);
          float w[16];
#pragma unroll
          for (int r = 0; r < 16; ++r)
            w[r] = __uint_as_float((__float_as_uint(acc[g][rc][r]) & ~15u) | uint32_t(r));
          float mt = w[0];
#pragma unroll
          for (int r = 1; r < 16; ++r) mt = fmaxf(mt, w[r]);
          // at most 16 rounds: each removes one survivor of every active lane
          while (__ballot(mt > nt)) {
            if (mt > nt) {
              const uint32_t r = __float_as_uint(mt) & 15u;
              const uint32_t row = (uint32_t)(t + rb * 32 + (r >> 2) * 8 + half * 4 + (r & 3));
              sl[cnt[g]] = C::make(-mt, row);
              ++cnt[g];
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) w[r] = w[r] < mt ? w[r] : -INFINITY;
            mt = w[0];
#pragma unroll
            for (int r = 1; r < 16; ++r) mt = fmaxf(mt, w[r]);
          }
        }
      }
      // a merge leaves room for the next chunk's appends (16 RC per lane)
      if (appended) MERGE_WHERE(g, cnt[g] > csub - 16 * RC);
    }
    }
    // The next tile (and the thresholds that came with it) have landed for
    // every wave; with two buffers this tile's one is free again and takes the
    // tile after next.
    const int nbuf = buf + 1 == NB ? 0 : buf + 1;
    if (NB == 3 && t + 2 * BN < r1)
      wait_all_but_last_stage();  // the stage of tile t + 2 BN may still be in flight
    else
      __builtin_amdgcn_s_waitcnt(0);
    if (t + BN < r1) {
#pragma unroll
      for (int g = 0; g < A; ++g) thr[g] = fminf(thr[g], unord_f32(thr_lds[nbuf][wave][g * 32 + col]));
    }
    __syncthreads();
    if constexpr (NB == 2) {
      if (t + 2 * BN < r1) stage(buf, t + 2 * BN);
    }
    buf = nbuf;
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
// Split-major order: all query blocks of split 0 start first, so that the
// later splits of a query begin with the thresholds the earlier ones reached.
template <int KIND, int KS, int A, int R, int V, bool RESID>
__global__ __launch_bounds__(256, 2) void scan_mfma_kernel(ScanMfmaArgs a) {
  scan_mfma_body<KIND, KS, A, R, V, RESID>(a, (int)(blockIdx.x % (unsigned)a.nqb),
                                           (int)(blockIdx.x / (unsigned)a.nqb));
}
template <int KIND, int KS, int A, int R, int V, bool RESID>
__global__ __launch_bounds__(256) void scan_mfma_kernel_wide(ScanMfmaArgs a) {
  scan_mfma_body<KIND, KS, A, R, V, RESID>(a, (int)(blockIdx.x % (unsigned)a.nqb),
                                           (int)(blockIdx.x / (unsigned)a.nqb));
}
// Shapes of at most 4 bf16 K steps: tiles are small (a few KB) and a tile's work
// is short next to the latency of its copy, so four workgroups share a CU.
template <int KIND, int KS, int A, int R, int V, bool RESID>
__global__ __launch_bounds__(256, 4) void scan_mfma_kernel_4wg(ScanMfmaArgs a) {
  scan_mfma_body<KIND, KS, A, R, V, RESID>(a, (int)(blockIdx.x % (unsigned)a.nqb),
                                           (int)(blockIdx.x / (unsigned)a.nqb));
}
// One wave per workgroup (64 query slots): small batches (serving) waste
// fewer padded query slots, and more workgroups share a CU, each on its own split.
template <int KIND, int KS, int A, int R, int V, bool RESID>
__global__ __launch_bounds__(64, 8) void scan_mfma_kernel_1w(ScanMfmaArgs a) {
  scan_mfma_body<KIND, KS, A, R, V, RESID, 1>(a, (int)(blockIdx.x % (unsigned)a.nqb),
                                              (int)(blockIdx.x / (unsigned)a.nqb));
}

// Persistent workgroups over the work items of many problems.  One lane takes
// the next item of a range from that range's counter and publishes it through
// LDS; the barriers around the exchange also separate one item's LDS tiles from
// the next one's (the body ends with every tile read and every copy into LDS
// completed).  A workgroup drains its XCD's range first, then the others in
// turn, and leaves once every range's counter has passed its end.
template <int KIND, int KS, int A, int R, int V, bool RESID>
__device__ __forceinline__ void scan_mfma_batch_body(const ScanBatchArgs& b) {
  __shared__ int64_t item_s;
  const int home = (int)(blockIdx.x % (unsigned)SCAN_XCDS);
  for (int c = 0; c < SCAN_XCDS; ++c) {
   const int x = (home + c) % SCAN_XCDS;
   const int64_t lo = b.range_start[x], hi = b.range_start[x + 1];
   for (;;) {
    if (threadIdx.x == 0) item_s = lo + (int64_t)atomicAdd(b.counter + x * SCAN_COUNTER_STRIDE, 1u);
    __syncthreads();
    const int64_t item = item_s;
    __syncthreads();
    if (item >= hi) break;
    int lo = 0, hi = b.nprob - 1;  // last problem whose first item is <= item
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (b.item_start[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const int p = __builtin_amdgcn_readfirstlane(lo);
    const ScanMfmaArgs a = b.probs[p];
    const int64_t local = item - b.item_start[p];
    scan_mfma_body<KIND, KS, A, R, V, RESID>(a, (int)(local % a.nqb), (int)(local / a.nqb));
   }
  }
}
template <int KIND, int KS, int A, int R, int V, bool RESID>
__global__ __launch_bounds__(256, 2) void scan_mfma_batch_kernel(ScanBatchArgs b) {
  scan_mfma_batch_body<KIND, KS, A, R, V, RESID>(b);
}
template <int KIND, int KS, int A, int R, int V, bool RESID>
__global__ __launch_bounds__(256) void scan_mfma_batch_kernel_wide(ScanBatchArgs b) {
  scan_mfma_batch_body<KIND, KS, A, R, V, RESID>(b);
}
template <int KIND, int KS, int A, int R, int V, bool RESID>
__global__ __launch_bounds__(256, 4) void scan_mfma_batch_kernel_4wg(ScanBatchArgs b) {
  scan_mfma_batch_body<KIND, KS, A, R, V, RESID>(b);
}

// K step buckets: bf16 covers 16 coordinates per step, f32 covers 8.  Up to 16
// steps a wave holds 2 query groups and the tile has 64 rows; beyond, 1 query
// group and 32-row tiles keep the query fragments at <= 256 registers and the
// two tile buffers within one CU's LDS.
static const int kBf16Steps[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 56, 64};
static const int kF32Steps[] = {1, 2, 3, 4, 6, 8, 12, 16, 20, 24, 32};
// E4M3 steps are 32 coordinates (32 bytes), taken two at a time by the MFMA.
static const int kFp8Steps[] = {2, 4, 6, 8};

int mfma_ksteps_bucket(int kind, int d) {
  const int per = kind == 0 ? 16 : (kind == 1 ? 8 : 32);
  int need = (d + per - 1) / per;
  if (need < 1) need = 1;
  const int* tab = kind == 0 ? kBf16Steps : (kind == 1 ? kF32Steps : kFp8Steps);
  const int nt = kind == 0   ? (int)(sizeof(kBf16Steps) / sizeof(int))
                 : kind == 1 ? (int)(sizeof(kF32Steps) / sizeof(int))
                             : (int)(sizeof(kFp8Steps) / sizeof(int));
  for (int i = 0; i < nt; ++i)
    if (tab[i] >= need) return tab[i];
  return -1;
}

int mfma_queries_per_block(int kind, int ksteps, int waves) { return waves * 32 * scan_groups(ksteps); }
int mfma_rows_per_tile(int kind, int ksteps, bool resid) { return 32 * scan_row_blocks(kind, ksteps, resid); }
// Resident workgroups per CU: the waves per SIMD the launch bounds allow (one
// wave of a one-wave workgroup per SIMD and bound), capped by what the
// workgroup's LDS leaves room for.
int mfma_workgroups_per_cu(int kind, int ksteps, int waves, bool resid) {
  const int per = scan_four_wg(kind, ksteps) ? 4 : (scan_two_waves(kind, ksteps, resid) ? 2 : 1);
  const int by_waves = waves == 1 ? 4 * std::min(per, 2) : per;
  const int R = scan_row_blocks(kind, ksteps, resid), W = waves == 1 ? 1 : 4;
  const int nb = scan_buffers(kind, ksteps, R, W, resid);
  const int by_lds = scan_wgs_by_lds(scan_lds_bytes(ksteps, R, W, false, nb) + scan_lds_extra(kind, nb, resid));
  return std::max(1, std::min(by_waves, by_lds));
}
bool mfma_one_wave_supported(int kind, int ksteps, int V, bool resid) {
  return resid && ksteps <= 12 && ((kind == 0 && V == 4) || (kind == 2 && (V == 4 || V == 8)));
}

// ---- seeded thresholds (SeedArgs in kernels.h) ----------------------------------
// The tile pass of scan_mfma_body without its selection: per (query group, row
// block) a max tree over the lane's 16 accumulators, folded into the lane's
// group register; the group registers rotate by R per tile, so tile i's row
// block rb lands in group (i R + rb) mod SEED_GROUPS.
template <int KIND, int KS, int A, int R>
__device__ __forceinline__ void seed_mfma_body(const SeedArgs& a, const int qb, const int part) {
  constexpr int W = 4;
  constexpr int G = SEED_GROUPS;
  static_assert(G % R == 0, "groups rotate by whole tiles");
  constexpr int BN = 32 * R;
  constexpr int ROWB = (2 * KS + 1) * 16;
  constexpr int WAVE_LOADS = scan_wave_loads(KS, R, W);
  constexpr int TILE_BYTES = scan_tile_bytes(KS, R, W);
  constexpr int HXN = scan_hx_words(R), HXW = scan_hx_waves(R);
  constexpr int RC = scan_row_chunk(KIND, KS, false);
  static_assert(HXW <= W, "every C input has a wave that copies it");
  __shared__ __attribute__((aligned(16))) char tile[2][TILE_BYTES];
  __shared__ __attribute__((aligned(16))) float tile_hx[2][HXN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int half = lane >> 5;
  const int col = lane & 31;
  const int wave = tid >> 6;
  const int64_t q0 = scan_wave_q0(qb, wave, A, W);
  const size_t kp_bytes = size_t(KS) * 32;
  const char* xop = reinterpret_cast<const char*>(a.Xop);
  uint4 qf[A][KS];
#pragma unroll
  for (int g = 0; g < A; ++g) {
    const char* qrow = reinterpret_cast<const char*>(a.Qop) + size_t(q0 + g * 32 + col) * kp_bytes + half * 16;
#pragma unroll
    for (int s = 0; s < KS; ++s) qf[g][s] = *reinterpret_cast<const uint4*>(qrow + s * 32);
  }
  float best[A][G];  // largest accumulator = smallest key, per group
#pragma unroll
  for (int g = 0; g < A; ++g)
#pragma unroll
    for (int j = 0; j < G; ++j) best[g][j] = -INFINITY;
  uint32_t soff[WAVE_LOADS];
#pragma unroll
  for (int j = 0; j < WAVE_LOADS; ++j) soff[j] = scan_stage_src(KS, R, W, j, wave, lane);
  const uint32_t lds_tile = (uint32_t)(uintptr_t)&tile[0][0];
  const uint32_t lds_hx = (uint32_t)(uintptr_t)&tile_hx[0][0];
  auto stage = [&](int b, int64_t t) {
    const char* base = xop + size_t(t) * kp_bytes;
#pragma unroll
    for (int j = 0; j < WAVE_LOADS; ++j)
      lds_dma16(__builtin_amdgcn_readfirstlane(lds_tile + b * TILE_BYTES + (j * W + wave) * 1024), base + soff[j]);
    if (wave < HXW) {
      const int r = wave * 64 + lane < BN ? wave * 64 + lane : BN - 1;
      lds_dma4(__builtin_amdgcn_readfirstlane(lds_hx + (b * HXN + wave * 64) * 4), a.hx + t + r);
    }
  };
  const int64_t step = a.parts;
  int64_t i = part;
  if (i < a.ntiles) stage(0, i * a.stride);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (i + step < a.ntiles) stage(1, (i + step) * a.stride);
  int buf = 0;
  for (; i < a.ntiles; i += step) {
    const char* tb = tile[buf];
    const float* hb = tile_hx[buf];
    // chunks of RC row blocks, as the scan takes them; block c0 + rc of the
    // tile lands in group register rc, and the registers rotate by RC per chunk
    // (so by R per tile)
#pragma unroll 1
    for (int c0 = 0; c0 < R; c0 += RC) {
      floatx16 acc[A][RC];
#pragma unroll
      for (int rc = 0; rc < RC; ++rc) {
        floatx16 h;
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
          const float4 h4 = *reinterpret_cast<const float4*>(hb + (c0 + rc) * 32 + q4 * 8 + half * 4);
          h[q4 * 4 + 0] = h4.x;
          h[q4 * 4 + 1] = h4.y;
          h[q4 * 4 + 2] = h4.z;
          h[q4 * 4 + 3] = h4.w;
        }
#pragma unroll
        for (int g = 0; g < A; ++g) acc[g][rc] = h;
      }
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        uint4 af[RC];
#pragma unroll
        for (int rc = 0; rc < RC; ++rc)
          af[rc] = *reinterpret_cast<const uint4*>(tb + ((c0 + rc) * 32 + col) * ROWB + s * 32 + half * 16);
#pragma unroll
        for (int g = 0; g < A; ++g)
#pragma unroll
          for (int rc = 0; rc < RC; ++rc) Mfma<KIND>::step(acc[g][rc], af[rc], qf[g][s]);
      }
#pragma unroll
      for (int g = 0; g < A; ++g)
#pragma unroll
        for (int rc = 0; rc < RC; ++rc) {
          float mx = acc[g][rc][0];
#pragma unroll
          for (int r = 1; r < 16; ++r) mx = fmaxf(mx, acc[g][rc][r]);
          best[g][rc] = fmaxf(best[g][rc], mx);
        }
#pragma unroll
      for (int g = 0; g < A; ++g) {
        float rot[G];
#pragma unroll
        for (int j = 0; j < G; ++j) rot[j] = best[g][(j + RC) % G];
#pragma unroll
        for (int j = 0; j < G; ++j) best[g][j] = rot[j];
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (i + 2 * step < a.ntiles) stage(buf, (i + 2 * step) * a.stride);
    buf ^= 1;
  }
#pragma unroll
  for (int g = 0; g < A; ++g) {
    float* out = a.mins + ((size_t(q0 + g * 32 + col) * a.parts + part) * 2 + half) * G;
#pragma unroll
    for (int j = 0; j < G; ++j) out[j] = -best[g][j];
  }
}

template <int KIND, int KS, int A, int R>
__global__ __launch_bounds__(256, 2) void seed_mfma_kernel(SeedArgs a) {
  seed_mfma_body<KIND, KS, A, R>(a, (int)(blockIdx.x % (unsigned)a.nqb), (int)(blockIdx.x / (unsigned)a.nqb));
}

// One wave per real query: the KP-th smallest of its parts * 2 * SEED_GROUPS
// group minima (a 32-step ballot search on the order-preserving bits), one ulp
// up so that the row attaining it still passes the scan's strict test, then
// the shared threshold drops to it.
// VMAX registers of 64 hold up to 64 VMAX minima.
template <int VMAX>
__device__ __forceinline__ void seed_select_body(const SeedArgs& a, const int64_t q) {
  if (q >= a.m) return;
  const int lane = __lane_id();
  const int total = a.parts * 2 * SEED_GROUPS;
  const float* v = a.mins + size_t(q) * total;
  uint32_t x[VMAX];
#pragma unroll
  for (int r = 0; r < VMAX; ++r) {
    const int e = r * 64 + lane;
    x[r] = e < total ? ord_f32(v[e]) : 0xFFFFFFFFu;
  }
  if (total < a.KP) return;
  uint32_t p = 0;
  for (int b = 31; b >= 0; --b) {
    const uint32_t probe = p | ((1u << b) - 1u);
    int c = 0;
#pragma unroll
    for (int r = 0; r < VMAX; ++r) c += __popcll(__ballot(x[r] <= probe));
    if (c < a.KP) p |= (1u << b);
  }
  const float t = unord_f32(p);
  if (!(t < INFINITY)) return;  // fewer than KP non-empty groups
  if (lane == 0) atomicMin(&a.gthr[q], ord_f32(nextafterf(t, INFINITY)));
}

__global__ __launch_bounds__(64) void seed_select_kernel(SeedArgs a) { seed_select_body<16>(a, blockIdx.x); }

// Grouped form: one wave per real query of every problem of the table.
template <int VMAX>
__global__ __launch_bounds__(64) void seed_select_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const SeedArgs a = static_cast<const SeedArgs*>(g.tab)[p];
  seed_select_body<VMAX>(a, int64_t(blockIdx.x) - g.start[p]);
}

hipError_t launch_seed_select_group(const GroupTab& g, int vmax, hipStream_t st) {
  if (g.np <= 0 || g.nblocks <= 0) return hipSuccess;
  const dim3 grid((unsigned)g.nblocks);
  if (vmax <= 4)
    hipLaunchKernelGGL(seed_select_group_kernel<4>, grid, dim3(64), 0, st, g);
  else if (vmax <= 8)
    hipLaunchKernelGGL(seed_select_group_kernel<8>, grid, dim3(64), 0, st, g);
  else if (vmax <= 16)
    hipLaunchKernelGGL(seed_select_group_kernel<16>, grid, dim3(64), 0, st, g);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// Grouped form: workgroup b of problem p is query block b % nqb of part b / nqb.
template <int KIND, int KS, int A, int R>
__global__ __launch_bounds__(256, 2) void seed_mfma_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const SeedArgs a = static_cast<const SeedArgs*>(g.tab)[p];
  const int64_t lb = int64_t(blockIdx.x) - g.start[p];
  seed_mfma_body<KIND, KS, A, R>(a, (int)(lb % a.nqb), (int)(lb / a.nqb));
}

// One problem (s) or a grouped launch (g).
template <int KIND, int KS>
static hipError_t launch_seed_ks(const SeedArgs* s, const GroupTab* g, hipStream_t st) {
  constexpr int A = scan_groups(KS);
  constexpr int R = scan_row_blocks(KIND, KS, false);
  if (g)
    hipLaunchKernelGGL((seed_mfma_group_kernel<KIND, KS, A, R>), dim3((unsigned)g->nblocks), dim3(256), 0, st, *g);
  else
    hipLaunchKernelGGL((seed_mfma_kernel<KIND, KS, A, R>), dim3((unsigned)(s->nqb * s->parts)), dim3(256), 0, st,
                       *s);
  return hipGetLastError();
}

static hipError_t launch_seed_pass(const SeedArgs* s, const GroupTab* g, int kind, int ks, hipStream_t st) {
  hipError_t e = hipErrorInvalidValue;
  if (kind == 0) {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    e = launch_seed_ks<0, K>(s, g, st); \
    break;
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8) KS_CASE(9)
      KS_CASE(10) KS_CASE(12) KS_CASE(14) KS_CASE(16) KS_CASE(20) KS_CASE(24) KS_CASE(28) KS_CASE(32)
      KS_CASE(40) KS_CASE(48) KS_CASE(56) KS_CASE(64)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    e = launch_seed_ks<1, K>(s, g, st); \
    break;
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(6) KS_CASE(8) KS_CASE(12) KS_CASE(16) KS_CASE(20)
      KS_CASE(24) KS_CASE(32)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  return e;
}

hipError_t launch_seed_mfma(const SeedArgs& s, int kind, int ks, hipStream_t st, bool select) {
  if (s.ntiles <= 0 || s.m <= 0) return hipSuccess;
  if (s.parts * 2 * SEED_GROUPS > 1024 || s.parts * 2 * SEED_GROUPS < s.KP) return hipErrorInvalidValue;
  const hipError_t e = launch_seed_pass(&s, nullptr, kind, ks, st);
  if (e != hipSuccess || !select) return e;
  hipLaunchKernelGGL(seed_select_kernel, dim3((unsigned)s.m), dim3(64), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_seed_mfma_group(const GroupTab& g, int kind, int ks, hipStream_t st) {
  if (g.np <= 0 || g.nblocks <= 0) return hipSuccess;
  return launch_seed_pass(nullptr, &g, kind, ks, st);
}

template <int KIND, int KS, bool RESID>
static hipError_t launch_ks(const ScanMfmaArgs& a, int V, int nsplit, hipStream_t st) {
  constexpr int A = scan_groups(KS);
  constexpr int R = scan_row_blocks(KIND, KS, RESID);
  constexpr bool two_waves = scan_two_waves(KIND, KS, RESID);
  if constexpr ((KIND == 0 || KIND == 2) && RESID && KS <= 12) {
    if (a.waves == 1) {
      if (V == 4) {
        hipLaunchKernelGGL((scan_mfma_kernel_1w<KIND, KS, A, R, 4, RESID>), dim3((unsigned)(a.nqb * nsplit)),
                           dim3(64), 0, st, a);
        return hipGetLastError();
      }
      if constexpr (KIND == 2) {  // the E4M3 copy keeps KP = 120 candidates (slots of 512)
        hipLaunchKernelGGL((scan_mfma_kernel_1w<KIND, KS, A, R, 8, RESID>), dim3((unsigned)(a.nqb * nsplit)),
                           dim3(64), 0, st, a);
        return hipGetLastError();
      }
      return hipErrorInvalidValue;
    }
  }
  if (a.waves == 1) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(a.nqb * nsplit)), block(256);
  if constexpr (scan_four_wg(KIND, KS)) {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_kernel_4wg<KIND, KS, A, R, 4, RESID>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((scan_mfma_kernel_4wg<KIND, KS, A, R, 8, RESID>), grid, block, 0, st, a);
  } else if constexpr (two_waves) {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_kernel<KIND, KS, A, R, 4, RESID>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((scan_mfma_kernel<KIND, KS, A, R, 8, RESID>), grid, block, 0, st, a);
  } else {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_kernel_wide<KIND, KS, A, R, 4, RESID>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((scan_mfma_kernel_wide<KIND, KS, A, R, 8, RESID>), grid, block, 0, st, a);
  }
  return hipGetLastError();
}

template <int KIND, int KS, bool RESID>
static hipError_t launch_batch_ks(const ScanBatchArgs& b, int V, int grid, hipStream_t st) {
  constexpr int A = scan_groups(KS);
  constexpr int R = scan_row_blocks(KIND, KS, RESID);
  constexpr bool two_waves = scan_two_waves(KIND, KS, RESID);
  if constexpr (scan_four_wg(KIND, KS)) {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_batch_kernel_4wg<KIND, KS, A, R, 4, RESID>), dim3(grid), dim3(256), 0, st, b);
    else
      hipLaunchKernelGGL((scan_mfma_batch_kernel_4wg<KIND, KS, A, R, 8, RESID>), dim3(grid), dim3(256), 0, st, b);
  } else if constexpr (two_waves) {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_batch_kernel<KIND, KS, A, R, 4, RESID>), dim3(grid), dim3(256), 0, st, b);
    else
      hipLaunchKernelGGL((scan_mfma_batch_kernel<KIND, KS, A, R, 8, RESID>), dim3(grid), dim3(256), 0, st, b);
  } else {
    if (V == 4)
      hipLaunchKernelGGL((scan_mfma_batch_kernel_wide<KIND, KS, A, R, 4, RESID>), dim3(grid), dim3(256), 0, st, b);
    else
      hipLaunchKernelGGL((scan_mfma_batch_kernel_wide<KIND, KS, A, R, 8, RESID>), dim3(grid), dim3(256), 0, st, b);
  }
  return hipGetLastError();
}

hipError_t launch_scan_mfma_batch(const ScanBatchArgs& b, int kind, int ks, int V, int grid, bool resid,
                                  hipStream_t st) {
  if (V != 4 && V != 8) return hipErrorInvalidValue;
  if (b.nitems <= 0) return hipSuccess;
  if (kind == 0 && resid) {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_batch_ks<0, K, true>(b, V, grid, st);
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8) KS_CASE(9)
      KS_CASE(10) KS_CASE(12) KS_CASE(14) KS_CASE(16) KS_CASE(20) KS_CASE(24) KS_CASE(28) KS_CASE(32)
      KS_CASE(40)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  if (kind == 0) {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_batch_ks<0, K, false>(b, V, grid, st);
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8) KS_CASE(9)
      KS_CASE(10) KS_CASE(12) KS_CASE(14) KS_CASE(16) KS_CASE(20) KS_CASE(24) KS_CASE(28) KS_CASE(32)
      KS_CASE(40) KS_CASE(48) KS_CASE(56) KS_CASE(64)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_batch_ks<1, K, false>(b, V, grid, st);
    KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(6) KS_CASE(8) KS_CASE(12) KS_CASE(16) KS_CASE(20)
    KS_CASE(24) KS_CASE(32)
#undef KS_CASE
    default:
      return hipErrorInvalidValue;
  }
}

int mfma_resid_max_ksteps() { return 40; }

namespace {

bool in_bucket(int kind, int ks) {
  if (kind < 0 || kind > 2) return false;
  const int* tab = kind == 0 ? kBf16Steps : (kind == 1 ? kF32Steps : kFp8Steps);
  const int nt = kind == 0   ? (int)(sizeof(kBf16Steps) / sizeof(int))
                 : kind == 1 ? (int)(sizeof(kF32Steps) / sizeof(int))
                             : (int)(sizeof(kFp8Steps) / sizeof(int));
  for (int i = 0; i < nt; ++i)
    if (tab[i] == ks) return true;
  return false;
}

struct Violations {
  int64_t count = 0;
  std::string first;
  void add(const char* what, long long v, long long lim) {
    if (count++ == 0) {
      char buf[200];
      std::snprintf(buf, sizeof(buf), "%s: %lld against a limit of %lld", what, v, lim);
      first = buf;
    }
  }
};

// The launch-independent part of the check: LDS staging of one shape.
void check_shape(int kind, int ks, int W, bool resid, Violations& v) {
  const int R = scan_row_blocks(kind, ks, resid);
  const int BN = 32 * R, CH = 2 * ks, ROWB = (CH + 1) * 16;
  const int WL = scan_wave_loads(ks, R, W), TB = scan_tile_bytes(ks, R, W);
  const uint32_t kp_bytes = uint32_t(ks) * 32;
  std::vector<uint32_t> src(size_t(WL) * 64 * W, 0xFFFFFFFFu);
  for (int j = 0; j < WL; ++j)
    for (int w = 0; w < W; ++w)
      for (int lane = 0; lane < 64; ++lane) {
        const long long dst = (long long)(j * W + w) * 1024 + lane * 16;  // inside one tile buffer
        if (dst + 16 > TB) v.add("LDS tile copy destination", dst + 16, TB);
        const uint32_t so = scan_stage_src(ks, R, W, j, w, lane);
        if (so + 16 > uint32_t(BN) * kp_bytes) v.add("tile copy source", so + 16, (long long)BN * kp_bytes);
        src[size_t(j * W + w) * 64 + lane] = so;
      }
  // every 16-byte fragment the matrix cores read holds (row, K step, lane half)
  for (int row = 0; row < BN; ++row)
    for (int s = 0; s < ks; ++s)
      for (int h = 0; h < 2; ++h) {
        const int byte = row * ROWB + s * 32 + h * 16;
        if (byte + 16 > TB || size_t(byte / 16) >= src.size()) {
          v.add("LDS fragment read", byte + 16, TB);
        } else if (src[byte / 16] != uint32_t(row) * kp_bytes + uint32_t(s * 32 + h * 16)) {
          v.add("staged chunk differs from the chunk read", src[byte / 16], (long long)row * kp_bytes + s * 32 + h * 16);
        }
      }
  // C inputs: lane l of wave w < HXW copies row t + 64 w + l (rows past BN
  // repeat the last) into word 64 w + l of the buffer's HXN; reads of float4 at
  // rb * 32 + q4 * 8 + half * 4 stay inside the words copied from rows < BN
  const int HXN = scan_hx_words(R), HXW = scan_hx_waves(R);
  if (HXW > W) v.add("C-input copying waves", HXW, W);
  if (64 * HXW > HXN) v.add("C-input copy destination", 64 * HXW, HXN);
  if ((R - 1) * 32 + 3 * 8 + 4 + 4 > std::min(HXN, 64 * std::min(HXW, W))) v.add("C-input read", (R - 1) * 32 + 32, HXN);
  // the thresholds ride in 64 words per wave and buffer; a wave reads g * 32 + col
  if (32 * scan_groups(ks) > 64) v.add("threshold read", 32 * scan_groups(ks), 64);
}

}  // namespace

int64_t scan_check_launch(const ScanMfmaArgs& a, int kind, int ks, int V, int nsplit, const ScanExtents& ex, int L,
                          std::string* first) {
  Violations v;
  const bool resid = a.resid != 0;
  const int W = a.waves == 1 ? 1 : 4;
  if (kind < 0 || kind > 2) v.add("operand kind", kind, 2);
  if (kind == 2 && !resid) v.add("E4M3 operand without the residual layout", kind, 0);
  if (!in_bucket(kind, ks) || (resid && (kind == 1 || ks > mfma_resid_max_ksteps()))) v.add("K steps", ks, 0);
  if (V != 4 && V != 8) v.add("slot width V", V, 8);
  if (W == 1 && !mfma_one_wave_supported(kind, ks, V, resid)) v.add("one-wave shape", ks, 12);
  if (v.count == 0) {
    // the shape part does not depend on the launch: once per shape
    static std::mutex mu;
    static std::map<int, std::pair<int64_t, std::string>> done;
    const int key = ((kind * 2 + (resid ? 1 : 0)) * 2 + (W == 1 ? 1 : 0)) * 128 + ks;
    std::lock_guard<std::mutex> lock(mu);
    auto it = done.find(key);
    if (it == done.end()) {
      Violations sv;
      check_shape(kind, ks, W, resid, sv);
      it = done.emplace(key, std::make_pair(sv.count, sv.first)).first;
    }
    if (it->second.first) {
      if (v.count == 0) v.first = it->second.second;
      v.count += it->second.first;
    }
  }
  if (v.count == 0) {
    const int A = scan_groups(ks), R = scan_row_blocks(kind, ks, resid), RC = scan_row_chunk(kind, ks, resid);
    const int NS = 64 * V, BN = 32 * R, QPB = W * 32 * A;
    const size_t kp_bytes = size_t(ks) * 32;
    const size_t es = kind == 0 ? 2 : (kind == 1 ? 4 : 1);
    if (size_t(a.Kp) * es != kp_bytes) v.add("operand row bytes", (long long)a.Kp * es, (long long)kp_bytes);
    if (a.KP < 1 || 2 * a.KP > NS) v.add("kept candidates KP", a.KP, NS / 2);
    // a merge leaves cnt <= Csub - 16 RC before a chunk of RC row blocks, whose
    // lanes then append at most 16 RC
    if (a.Csub < 16 * RC) v.add("append segment shorter than one chunk's appends", a.Csub, 16 * RC);
    if (a.KP + 2 * a.Csub > NS) v.add("kept list + append segments", a.KP + 2 * a.Csub, NS);
    if (a.rows_per_split <= 0 || a.rows_per_split % BN) v.add("rows per split, whole tiles", a.rows_per_split, BN);
    if (a.n_pad <= 0 || a.n_pad % BN) v.add("operand rows, whole tiles", a.n_pad, BN);
    if (a.nqb < 1 || a.nq_pad != int64_t(a.nqb) * QPB) v.add("query slots", a.nq_pad, (long long)a.nqb * QPB);
    if (nsplit < 1) v.add("splits", nsplit, 1);
    // per-query accesses grow with the slot: the last slot of the last split bounds them
    const int64_t qlast = int64_t(a.nqb) * QPB - 1;
    if (!resid && size_t(qlast + 1) * kp_bytes > ex.qop)
      v.add("query operand", (long long)((qlast + 1) * kp_bytes), (long long)ex.qop);
    if (size_t(qlast + 1) * 4 > ex.gthr) v.add("shared thresholds", (long long)(qlast + 1) * 4, (long long)ex.gthr);
    if (size_t(nsplit) * a.nq_pad * 4 > ex.lcnts)
      v.add("list counts", (long long)(size_t(nsplit) * a.nq_pad * 4), (long long)ex.lcnts);
    const size_t slot_end = scan_slot_index(nsplit - 1, a.nq_pad, qlast, NS, NS) * 8;
    if (slot_end > ex.slots) v.add("candidate slots", (long long)slot_end, (long long)ex.slots);
    // rows: the last tile of every split (every tile reads BN operand rows and C inputs)
    for (int s = 0; s < nsplit && v.count == 0; ++s) {
      int64_t r0, r1;
      scan_split_rows(s, a.rows_per_split, a.n_pad, r0, r1);
      if (r0 >= r1) continue;
      const int64_t tl = r0 + (r1 - 1 - r0) / BN * BN;
      if (size_t(tl + BN) * kp_bytes > ex.xop)
        v.add("operand rows", (long long)(size_t(tl + BN) * kp_bytes), (long long)ex.xop);
      if (size_t(tl + BN) * 4 > ex.hx) v.add("C inputs", (long long)(size_t(tl + BN) * 4), (long long)ex.hx);
      if (resid && size_t((tl >> 6) + 1) * 4 > ex.tile_list)
        v.add("block lists", (long long)((tl >> 6) + 1) * 4, (long long)ex.tile_list);
      // E4M3 row scales: a tile copies BN bytes as BN / 4 words from row tl
      if (kind == 2 && (size_t(tl + BN) > ex.xsc || (tl & 3) != 0 || (BN & 3) != 0))
        v.add("row scales", (long long)(tl + BN), (long long)ex.xsc);
    }
    if (resid) {
      if (L < 1) v.add("lists", L, 1);
      if (size_t(L) * a.nq_pad * 4 > ex.qtab)
        v.add("per-list query terms", (long long)(size_t(L) * a.nq_pad * 4), (long long)ex.qtab);
      if (size_t(L - 1) * a.ldc * 4 + size_t(a.d) * 4 > ex.cents)
        v.add("centroids", (long long)(size_t(L - 1) * a.ldc * 4 + size_t(a.d) * 4), (long long)ex.cents);
      // padding slots repeat query m - 1
      if (a.m < 1 || size_t(a.m - 1) * a.ld32 * 4 + size_t(a.d) * 4 > ex.q32)
        v.add("f32 queries", (long long)(size_t(a.m - 1) * a.ld32 * 4 + size_t(a.d) * 4), (long long)ex.q32);
      if (a.d > (kind == 2 ? 32 : 16) * ks || (a.ld32 & 7)) v.add("f32 query pitch", a.ld32, 8);
      if (kind == 2) {
        // prepared fragments: hi and lo rows of kp_bytes per (list, slot), one scale per (list, slot)
        if (a.L != L) v.add("fragment lists", a.L, L);
        const size_t frag_end = size_t(2) * L * a.nq_pad * kp_bytes;
        if (frag_end > ex.qfrag) v.add("query fragments", (long long)frag_end, (long long)ex.qfrag);
        if (size_t(L) * a.nq_pad * 4 > ex.qsc)
          v.add("query scales", (long long)(size_t(L) * a.nq_pad * 4), (long long)ex.qsc);
      }
    }
  }
  if (v.count && first) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), " (kind %d, %d K steps, V %d, %d waves, %d splits, %s)", kind, ks, V, W, nsplit,
                  resid ? "residual" : "plain");
    *first = v.first + buf;
  }
  return v.count;
}

hipError_t launch_scan_mfma(const ScanMfmaArgs& a, int kind, int ks, int V, int nsplit, hipStream_t st) {
  if (V != 4 && V != 8) return hipErrorInvalidValue;
  if (kind == 2) {
    if (!a.resid || !a.xsc) return hipErrorInvalidValue;
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_ks<2, K, true>(a, V, nsplit, st);
      KS_CASE(2) KS_CASE(4) KS_CASE(6) KS_CASE(8)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  if (kind == 0 && a.resid) {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_ks<0, K, true>(a, V, nsplit, st);
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8) KS_CASE(9)
      KS_CASE(10) KS_CASE(12) KS_CASE(14) KS_CASE(16) KS_CASE(20) KS_CASE(24) KS_CASE(28) KS_CASE(32)
      KS_CASE(40)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  if (kind == 0) {
    switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_ks<0, K, false>(a, V, nsplit, st);
      KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(5) KS_CASE(6) KS_CASE(7) KS_CASE(8) KS_CASE(9)
      KS_CASE(10) KS_CASE(12) KS_CASE(14) KS_CASE(16) KS_CASE(20) KS_CASE(24) KS_CASE(28) KS_CASE(32)
      KS_CASE(40) KS_CASE(48) KS_CASE(56) KS_CASE(64)
#undef KS_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  switch (ks) {
#define KS_CASE(K) \
  case K:          \
    return launch_ks<1, K, false>(a, V, nsplit, st);
    KS_CASE(1) KS_CASE(2) KS_CASE(3) KS_CASE(4) KS_CASE(6) KS_CASE(8) KS_CASE(12) KS_CASE(16) KS_CASE(20)
    KS_CASE(24) KS_CASE(32)
#undef KS_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace fa
