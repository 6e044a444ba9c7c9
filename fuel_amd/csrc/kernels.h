// Kernel argument blocks and launchers shared between the HIP translation units
// and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.h"

namespace fa {

// Exclusion semantics for self-queries (query slot q is data row qrow):
//   0 none; 1 fuel's for_each_knn (k+1, drop self, truncate: src/outlier/common.rs:27,41-44);
//   2 the Python linear-scan path (k, drop self: python/search.rs:485-488).
enum ExcludeMode { EXCL_NONE = 0, EXCL_FOR_EACH_KNN = 1, EXCL_PY_LINEAR = 2 };

// Grouped launches: one grid over the np problems of a device argument table
// (tab[p] of the kernel's own argument type); workgroup b belongs to problem p
// with start[p] <= b < start[p + 1] (start[0] = 0, start[np] = nblocks), so
// small problems share one launch instead of one each.
struct GroupTab {
  const void* tab;
  const int64_t* start;
  int np;
  int64_t nblocks;
};
__device__ __forceinline__ int group_problem(const GroupTab& g, int64_t b) {
  int lo = 0, hi = g.np;  // start[lo] <= b < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (g.start[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename F>
struct ScanExactArgs {
  const F* X;
  int64_t ldx;
  int64_t n;
  int d;
  const F* Q;
  int64_t ldq;
  const int64_t* qrows;  // slot -> query row (nullptr: identity)
  int64_t nq;            // number of query slots
  int metric;            // 0 euclidean, 1 sqeuclidean
  int arith;             // ARITH_AVX or ARITH_FALLBACK
  int KP;
  int64_t rows_per_split;
  typename Comp<F>::T* slots;  // [nsplit][nq][NS]
  int* lcnts;                  // [nsplit][nq]
  const typename Comp<F>::T* lo;  // optional [nq]: keep only candidates ordering after lo[slot]
};

// Order-preserving unsigned keys of a distance (u32 for f32, u64 for f64).
template <typename F>
struct OrdKey;
template <>
struct OrdKey<float> {
  using T = uint32_t;
};
template <>
struct OrdKey<double> {
  using T = uint64_t;
};
__device__ __forceinline__ uint32_t ord_key(float f) { return ord_f32(f); }
__device__ __forceinline__ uint64_t ord_key(double f) { return ord_f64(f); }
template <typename F>
__device__ __forceinline__ F unord_key(typename OrdKey<F>::T k);
template <>
__device__ __forceinline__ float unord_key<float>(uint32_t k) { return unord_f32(k); }
template <>
__device__ __forceinline__ double unord_key<double>(uint64_t k) { return unord_f64(k); }

template <typename F>
struct RangeArgs {
  const F* X;
  int64_t ldx;
  int64_t n;
  int d;
  const F* Q;
  int64_t ldq;
  int64_t nq;
  int64_t qbase;  // data row of query slot 0 (self-exclusion)
  int metric;
  int arith;      // ARITH_AVX or ARITH_FALLBACK
  F radius;
  const F* radii;        // optional per-query radius (query slot i: radii[i * radii_stride])
  int64_t radii_stride;
  const F* skips;        // optional per-query lower limit (keep distances >= skips[i])
  int exclude_self;
  int nsplit;
  int64_t rows_per_split;
  int64_t* counts;         // pass 0: [nq][nsplit]
  const int64_t* offsets;  // pass 1: exclusive scan of counts
  typename OrdKey<F>::T* out_keys;
  int64_t* out_idx;
};

template <typename F>
struct PairwiseArgs {
  const F* X;
  int64_t ldx;
  int64_t n;
  int d;
  int metric;
  int arith;  // ARITH_AVX or ARITH_FALLBACK
  int64_t rows_per_split;
  F* out;
};

template <typename F>
struct LofArgs {
  const int64_t* nb_idx;  // n x k (for_each_knn semantics, no padding)
  const F* nb_dist;
  int64_t n;
  int k;
  F* kdist;
  F* lrd;
  F* out;
};

template <typename F>
struct FinalizeArgs {
  const typename Comp<F>::T* slots;
  const int* lcnts;
  int64_t slot_stride;  // NS of the scan that produced the slots
  int nsplit;
  int64_t nq;
  int KK;
  int k;
  int mode;
  const int64_t* qrows;  // slot -> query row within this batch (nullptr: identity)
  int64_t qbase;         // data row of batch query 0 (self-exclusion)
  int64_t* out_idx;      // rows indexed by query row
  F* out_dist;
  // Instead of output rows: the slot's sorted top-KK composites go to
  // raw_out[slot * raw_stride + raw_off + e] (MAX-padded), their count is added
  // to cnt_out[slot], and lo_out[slot] receives the KK-th (MAX when fewer were
  // found): the round of a kNN whose k exceeds one slot.
  typename Comp<F>::T* raw_out;
  int64_t raw_stride;
  int64_t raw_off;
  typename Comp<F>::T* lo_out;
  int* cnt_out;
};

// Output rows of a kNN found in rounds: slot s holds cnt[s] sorted composites
// raw[s * stride ...]; fuel's exclusion / truncation / padding as write_knn_row.
template <typename F>
struct LongRowArgs {
  const typename Comp<F>::T* raw;
  int64_t stride;
  const int* cnt;
  int64_t nq;
  int k;
  int mode;
  const int64_t* qrows;
  int64_t qbase;
  int64_t* out_idx;
  F* out_dist;
};
template <typename F>
hipError_t launch_write_long_rows(const LongRowArgs<F>& a, hipStream_t st);

// Write one output row from a wave-resident, ascending, lexicographically sorted
// candidate set x (count valid entries) with fuel's exclusion/truncation/padding
// semantics (python/search.rs:485-488, :1195-1216; src/outlier/common.rs:41-44).
template <typename F, int V>
__device__ __forceinline__ void write_knn_row(const typename Comp<F>::T (&x)[V], int count, int64_t qrow,
                                              int k, int mode, int64_t* __restrict__ oidx,
                                              F* __restrict__ odist) {
  using C = Comp<F>;
  const int lane = __lane_id();
  // position of the query's own index (at most one entry)
  int sp = 64 * V;
  if (mode != EXCL_NONE) {
#pragma unroll
    for (int v = V - 1; v >= 0; --v) {
      const int e = v * 64 + lane;
      const bool m = e < count && (int64_t)C::idx(x[v]) == qrow;
      const uint64_t b = __ballot(m);
      if (b) sp = v * 64 + __builtin_ctzll(b);
    }
  }
  F kth = F(INFINITY);
  if (mode == EXCL_PY_LINEAR && k >= 1 && count >= k) kth = C::key(wave_get<typename C::T, V>(x, k - 1));
  int written = 0;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int e = v * 64 + lane;
    bool keep = e < count;
    if (mode != EXCL_NONE) keep = keep && e != sp;
    if (mode == EXCL_PY_LINEAR) keep = keep && C::key(x[v]) <= kth;
    const int j = e - (e > sp ? 1 : 0);
    if (keep && j < k) {
      oidx[j] = (int64_t)C::idx(x[v]);
      odist[j] = C::key(x[v]);
    }
    written += __popcll(__ballot(keep));
  }
  if (written > k) written = k;
  for (int j = written + lane; j < k; j += 64) {
    oidx[j] = -1;
    odist[j] = F(INFINITY);
  }
}

// ---- mixed-precision candidate path ------------------------------------------
// Index-level bound constants (nonnegative doubles stored as their bit patterns
// so that atomicMax on u64 orders them): [0] max ||x||^2, [1] max ||x~ - x||^2,
// [2] max ||x~||^2 (x~ = the MFMA operand copy), [3] max |fl32(||x||^2/2) -
// ||x||^2/2| and [4] max fl32(||x||^2/2) (the MFMA's C input).  flags (8 bytes after
// the 8 constants): [0] = 1 while every coordinate is an integer of magnitude
// <= 256, [1] = 1 while all are >= 0.
struct PrepDbArgs {
  const float* X;  // resident f32 rows (ldx)
  int64_t ldx;
  int64_t n;
  int64_t n_pad;
  int d;
  int Kp;
  int bf16;    // operand type: 1 bf16, 0 f32
  void* Xop;   // n_pad x Kp operand copy
  float* hx;   // n_pad: -fl32(||x||^2 / 2), MFMA_PAD_HXN for padding rows
  unsigned long long* consts;
  int* flags;
};

struct PrepQArgs {
  const float* Q;
  int64_t ldq;
  int64_t m;
  int64_t m_pad;
  int d;
  int Kp;
  int bf16;
  void* Qop;    // m_pad x Kp operand copy (bf16 or f32)
  float* Q32;   // m x ld32 aligned f32 copy for the exact rerank
  int64_t ld32;
  double* qn2;  // ||q||^2
  double* eq2;  // ||q~ - q||^2
  double* qt2;  // ||q~||^2
  int* qint;    // 1 if integer-valued with |v| <= 256
  uint32_t* gthr;  // m_pad shared thresholds (order-preserving bits), initialised here
  const int64_t* rows;  // optional: slot i reads query row rows[i] of Q
  const float* theta;   // optional: slot i's initial threshold (key units); +inf without
};

// fuel's math::pairwise_sqdist (k-means distance matrix): out[i * ldo + j] for
// rows i of P1 (m x d, pitch ld1) and rows j of P2 (n x d, pitch ld2).
template <typename F>
struct KmeansPairwiseArgs {
  const F* P1;
  int64_t ld1;
  int64_t m;
  const F* P2;
  int64_t ld2;
  int64_t n;
  int d;
  F* out;
  int64_t ldo;
};
template <typename F>
hipError_t launch_kmeans_pairwise(const KmeansPairwiseArgs<F>& a, hipStream_t st);

// Finite sentinels of the MFMA scan.  Thresholds start at KEY_LIMIT instead of
// +inf and padding rows carry PAD_HXN as their C input, so a padding row's
// result (PAD_HXN) never beats a threshold and every value stays finite, which
// the row tag in the low mantissa bits requires (a tagged infinity would be a
// NaN).  The scan is used only while every squared norm stays below NORM2_LIMIT,
// so that real keys stay far below KEY_LIMIT.
constexpr float MFMA_KEY_LIMIT = 1.0e38f;
constexpr float MFMA_PAD_HXN = -3.0e38f;
constexpr double MFMA_NORM2_LIMIT = 1.0e36;

struct ScanMfmaArgs {
  const void* Xop;
  const float* hx;
  int Kp;
  int64_t n_pad;
  const void* Qop;
  int64_t nq_pad;          // query slots (multiple of the workgroup's query count)
  int nqb;                 // query blocks
  int64_t rows_per_split;  // multiple of the row tile
  int KP;                  // candidates kept per (split, query)
  int Csub;                // per-lane-half append buffer capacity
  uint64_t* slots;         // [nsplit][nq_pad][NS]
  int* lcnts;              // [nsplit][nq_pad]
  uint32_t* gthr;          // [nq_pad] per-query threshold shared by the splits (order-preserving bits)
  // residual operand (bf16 only; resid = 0 otherwise): rows sorted by list,
  // every list padded to whole 64-row blocks; the B operand of a block of list
  // j is bf16(fl32(q - c_j)), rebuilt from Q32 when the list changes, and the
  // per-(list, query) term qtab[j * nq_pad + q] joins the C input
  int resid;
  const int* tile_list;    // list of each 64-row block
  const float* cents;      // L x ldc centroids
  int ldc;
  const float* Q32;        // m x ld32 f32 queries
  int64_t ld32;
  int64_t m;
  int d;
  const float* qtab;       // [L][nq_pad]
  int waves;               // waves per workgroup: 4, or 1 (see mfma_one_wave_supported); 0 means 4
  const uint8_t* xsc;      // E4M3 operand (kind 2): n_pad E8M0 row-scale codes
  const uint8_t* qfrag;    // E4M3: [2][L][nq_pad][Kp] hi, then lo codes of q - c_j (prep_qfrag_fp8)
  const int* qsc;          // E4M3: [L][nq_pad] E8M0 code of the hi scale
  int L;                   // E4M3: lists
};

// ---- residual (centroid-relative) operand ------------------------------------
// Nearest centroid of each row (ties: the smaller centroid index), and the
// number of rows per centroid.
struct AssignArgs {
  const float* X;
  int64_t ldx;
  int64_t n;
  int d;
  const float* cents;  // L x d
  int L;
  uint32_t* list;      // n
  uint32_t* counts;    // L (incremented)
};
hipError_t launch_assign(const AssignArgs& a, hipStream_t st);
// New centroid of every non-empty list: the f64 mean of its rows, in sorted
// row order (deterministic); empty lists keep their centroid.
hipError_t launch_list_means(const float* X, int64_t ldx, int d, const uint32_t* sorted_rows,
                             const int64_t* offsets, const uint32_t* counts, int L, float* cents,
                             hipStream_t st);
// Padded position of every row: rperm[pad_off[list] + rank] = row.
hipError_t launch_scatter_rperm(const uint32_t* sorted_rows, const uint32_t* sorted_list, int64_t n,
                                const int64_t* offsets, const int64_t* pad_offsets, uint32_t* rperm,
                                hipStream_t st);
hipError_t launch_iota_u32(uint32_t* p, int64_t n, hipStream_t st);
// Stable sort of (key, value) u32 pairs by key (rocPRIM radix sort, low `bits` bits).
hipError_t sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                          int64_t n, int bits, void* tmp, size_t& tmp_bytes, hipStream_t st);

// Residual operand rows: position p holds row rperm[p] (RPERM_NONE: padding)
// of list tile_list[p / 64] as r~ = bf16(fl32(x - c_j)), hxn = -fl32(||r||^2/2)
// with r = x - c_j exact in f64; per-list maxima (order-preserving u64 bits of
// non-negative doubles) lconsts[8 j + i]: [0] ||r||^2, [1] ||r~ - r||^2,
// [2] ||r~||^2, [3] |fl32(||r||^2/2) - ||r||^2/2| (+ the f64 error of the
// sum), [4] fl32(||r||^2/2).
constexpr uint32_t RPERM_NONE = 0xFFFFFFFFu;
struct PrepResidArgs {
  const float* X;
  int64_t ldx;
  int d;
  int Kp;
  int64_t n_rpad;
  const uint32_t* rperm;
  const int* tile_list;
  const float* cents;
  int ldc;
  void* Xop;      // n_rpad x Kp: bf16, or E4M3 codes of r 2^-e_p (fp8)
  uint8_t* xsc;   // fp8: n_rpad E8M0 codes of the per-row scales 2^e_p
  int fp8;
  float* hx;
  unsigned long long* lconsts;  // L x 8, zero before the launch
  unsigned long long* gconsts;  // [0] max ||x||^2 (the sentinel-range check); zero before the launch
};
hipError_t launch_prep_resid(const PrepResidArgs& a, hipStream_t st);
// C-input term of every (list, query slot): a float no smaller than
// -||q - c_j||^2 / 2 + e_j(q), e_j(q) bounding the residual product's
// operand rounding, the hxn rounding, the MFMA accumulation and the C-input
// addition, so that -(MFMA result) never exceeds ||q - x||^2 / 2 before the
// row tag.  s2[q * L + j] is ||q - c_j||^2 as the f32 pairwise kernel computes
// it (launch_kmeans_pairwise of the queries against the centroids).  Slots >= m
// repeat query m - 1.
// The same launch seeds each slot's shared threshold: with N_j an upper bound of
// ||q - c_j|| and RN_j = max ||r|| over list j, every row of list j has
// ||q - x||^2 / 2 <= (N_j + RN_j)^2 / 2, and its key is no larger (up to the
// row tag); so theta = min over lists with >= KP rows of that bound, rounded
// up, lies at or above the KP-th smallest key, and gthr = min(gthr, theta)
// drops only rows that cannot enter any split's kept list.
struct QListArgs {
  const float* s2;  // m x L
  int64_t m;
  int64_t nq_pad;
  int d;
  int fp8;          // the operand is the E4M3 copy, its B operand the hi + lo E4M3 split of q - c_j
  int L;
  const unsigned long long* lconsts;
  float* qtab;  // [L][nq_pad]
  const uint32_t* lcount;  // rows per list
  int KP;
  uint32_t* gthr;          // [nq_pad], initialised by the query preparation
};
hipError_t launch_prep_qlist(const QListArgs& a, hipStream_t st);
// Grouped forms over tables of PrepResidArgs (n_rpad / 64 workgroups each, bf16
// operand), QListArgs ((L nq_pad + 255) / 256 for the terms, then (nq_pad + 3)
// / 4 for the proven start, two tables of the same arguments) and
// SampleSeedArgs (m each).
hipError_t launch_prep_resid_group(const GroupTab& g, hipStream_t st);
hipError_t launch_prep_qlist_group(const GroupTab& terms, const GroupTab& theta, hipStream_t st);
hipError_t launch_seed_sample_group(const GroupTab& g, hipStream_t st);
// Sampled threshold of a residual scan, one wave per real query: the nearest
// list with at least KP rows (smallest s2, ties to the smaller index), its
// first min(rows, SEED_SAMPLE_ROWS) rows, each row's ||q - x||^2 / 2 in f64
// raised to an upper bound of its scan key (x (1 + 2^-16) + 2^-140: the key is
// a lower bound of that value up to the row tag), and the KP-th smallest of
// them one float step up: at least KP rows keep a key below it, so the scan
// drops nothing its lists need, and gthr falls to it when lower.
constexpr int SEED_SAMPLE_ROWS = 512;
struct SampleSeedArgs {
  const float* s2;          // m x L, ||q - c_j||^2
  int64_t m;
  int L;
  int KP;
  const uint32_t* lcount;   // rows per list
  const int64_t* loff;      // first position of every list (L + 1)
  const uint32_t* rperm;    // position -> row
  const float* X;           // rows, pitch ldx
  int64_t ldx;
  const float* Q32;         // queries, pitch ld32
  int64_t ld32;
  int d;
  uint32_t* gthr;
};
hipError_t launch_seed_sample(const SampleSeedArgs& a, hipStream_t st);
// B fragments of the E4M3 residual scan, one wave per (list j, query slot q;
// slots >= m repeat query m - 1): dq = fl32(q - c_j), e = fp8_scale_exp(max
// |dq|), hi = e4m3(dq 2^-e), lo = e4m3((dq - hi 2^e) 2^-(e-4)); byte k of a
// fragment row is coordinate k (zero beyond d), as in the operand rows.
struct QFragArgs {
  const float* Q32;
  int64_t ld32;
  int64_t m;
  int64_t nq_pad;
  int d;
  const float* cents;
  int ldc;
  int L;
  int Kp;          // bytes per fragment row (a multiple of 64)
  uint8_t* qfrag;  // [2][L][nq_pad][Kp]
  int* qsc;        // [L][nq_pad]
};
hipError_t launch_prep_qfrag_fp8(const QFragArgs& a, hipStream_t st);

struct RerankArgs {
  const uint64_t* slots;
  const int* lcnts;
  int nsplit;
  int64_t nq;
  int64_t nq_pad;
  int slot_ns;
  int KP;
  const float* X;
  int64_t ldx;
  const float* Q32;
  int64_t ld32;
  int d;
  int metric;
  int arith;  // rounding of the exact distances: ARITH_AVX or ARITH_FALLBACK
  int bf16;
  const double* qn2;
  const double* eq2;
  const double* qt2;
  const int* qint;
  const unsigned long long* consts;
  const int* flags;
  int KK;
  int k;
  int mode;
  int64_t qbase;
  const int64_t* rows;  // optional: slot i is query row rows[i] (output row and self index)
  int64_t* out_idx;
  float* out_dist;
  int* fail_count;
  int64_t* fail_list;  // receives the query rows (not slots) that did not certify
  const uint32_t* gthr;  // final shared thresholds of the scan, per slot
  float* fail_theta;     // optional: beside fail_list, a key threshold below which
                         // every row that can still enter the query's answer lies
  // residual operand: candidates are padded positions (rperm maps them to rows)
  // and keys are lower bounds of ||q - x||^2 / 2 up to the row tag
  int resid;
  const uint32_t* rperm;
};

// A kernel that stores v into count u32s (unlike hipMemsetAsync, safe to
// record into a captured graph).
hipError_t launch_fill_u32(uint32_t* p, int64_t count, uint32_t v, hipStream_t st);

// Persistent form of the MFMA scan over many problems of one shape (same kind,
// K steps and slot width): work item i of problem p (i in [0, nqb * nsplit))
// is query block i % nqb of split i / nqb; workgroups take items from a shared
// counter until none are left.
// The persistent scan's work items are cut into SCAN_XCDS ranges of about equal
// work; a workgroup starts on the range of its XCD (blockIdx.x mod SCAN_XCDS:
// workgroups are dispatched to the XCDs in turn), so the items of one split,
// which read the same rows, share that XCD's L2; a workgroup whose range is
// drained helps the others, so every item is taken exactly once.
constexpr int SCAN_XCDS = 8;
constexpr int SCAN_COUNTER_STRIDE = 64;  // words between the ranges' counters
struct ScanBatchArgs {
  const ScanMfmaArgs* probs;   // device array, one per problem
  const int64_t* item_start;   // nprob + 1 running totals of the problems' items
  int nprob;
  int64_t nitems;
  unsigned int* counter;       // SCAN_XCDS counters, SCAN_COUNTER_STRIDE words apart, zero before the launch
  int64_t range_start[SCAN_XCDS + 1];  // item ranges: [range_start[x], range_start[x + 1])
};

hipError_t launch_prep_db(const PrepDbArgs& a, hipStream_t st);
hipError_t launch_prep_queries(const PrepQArgs& a, hipStream_t st);

// count u32s at p set to v, one table entry per workgroup
struct FillEntry {
  uint32_t* p;
  int64_t count;
  uint32_t v;
  uint32_t pad;
};
hipError_t launch_fill_multi(const FillEntry* tab, int n, hipStream_t st);
// PrepDbArgs, (n_pad + 63) / 64 workgroups each, and PrepQArgs, (m_pad + 3) / 4
// each; thread_rows: one row per thread instead of one per wave, (n_pad + 255)
// / 256 and (m_pad + 255) / 256 workgroups each, for Kp <= 64 with rows 16-byte
// aligned and pitches (ldx, ldq, ld32) multiples of 4
constexpr int PREP_THREAD_ROWS_MAX_KP = 64;
hipError_t launch_prep_db_group(const GroupTab& g, bool thread_rows, hipStream_t st);
hipError_t launch_prep_queries_group(const GroupTab& g, bool thread_rows, hipStream_t st);
hipError_t launch_rerank_group(const GroupTab& g, int V, int arith, hipStream_t st);  // RerankArgs, nq each
// SeedArgs, m each; vmax = 64-entry registers for the largest parts * 2 * SEED_GROUPS
hipError_t launch_seed_select_group(const GroupTab& g, int vmax, hipStream_t st);
// kind 0: bf16 MFMA, 1: f32 MFMA, 2: E4M3 MFMA (residual operand only, with
// its row scales in a.xsc).  Returns hipErrorInvalidValue for an unsupported K
// step count.
hipError_t launch_scan_mfma(const ScanMfmaArgs& a, int kind, int ksteps, int V, int nsplit, hipStream_t st);
hipError_t launch_rerank(const RerankArgs& a, int V, hipStream_t st);
// grid = the number of persistent workgroups
hipError_t launch_scan_mfma_batch(const ScanBatchArgs& b, int kind, int ksteps, int V, int grid, bool resid,
                                  hipStream_t st);
// Supported K step counts of the MFMA scan (bf16: 16 dims per step; f32: 8 dims
// per step group; E4M3: 32 dims per step, an even count); the padded K is
// steps * 16 (bf16), steps * 8 (f32) or steps * 32 (E4M3).
int mfma_ksteps_bucket(int kind, int d);
// Largest K step count with a residual-operand variant of the bf16 scan.
int mfma_resid_max_ksteps();
// Launch shape: waves per workgroup is 4, or 1 where mfma_one_wave_supported
// (residual bf16 or E4M3 scans with V = 4 and at most 12 K steps; small batches).
int mfma_queries_per_block(int kind, int ksteps, int waves = 4);
int mfma_rows_per_tile(int kind, int ksteps, bool resid);
int mfma_workgroups_per_cu(int kind, int ksteps, int waves, bool resid);
bool mfma_one_wave_supported(int kind, int ksteps, int V, bool resid);

// Seeded thresholds of the plain MFMA scans.  Workgroup (query block qb, part
// p) runs the scan's own tile pass (same fragments, C inputs and MFMA order,
// so the keys are the scan's) over the sample tiles p, p + parts, ... < ntiles,
// tile i starting at row i * stride, and keeps no candidates: per lane the
// smallest key of each of SEED_GROUPS groups of rows.  The group minima of a
// query are keys of distinct rows, so the KP-th smallest of them bounds the
// query's KP-th smallest key from above; the select pass lowers the query's
// shared threshold to it.  Exactness never depends on the seed: the
// certificate uses the thresholds the scan actually applied.
constexpr int SEED_GROUPS = 16;  // per lane half
struct SeedArgs {
  const void* Xop;
  const float* hx;
  int Kp;
  const void* Qop;
  int64_t nq_pad;
  int nqb;
  int64_t stride;  // rows between sample tiles (a multiple of the tile)
  int64_t ntiles;  // sample tiles
  int parts;       // workgroups per query block
  float* mins;     // [nq_pad][parts][2][SEED_GROUPS] group minima (+inf when empty)
  int64_t m;       // real queries (slots beyond keep their threshold)
  int KP;
  uint32_t* gthr;  // [nq_pad] shared thresholds, lowered by the select pass
};
// The sample pass, then (select) the per-query selection; without it the
// caller runs launch_seed_select_group over the problems' SeedArgs.
hipError_t launch_seed_mfma(const SeedArgs& s, int kind, int ksteps, hipStream_t st, bool select = true);
// The sample passes of a table of SeedArgs (nqb * parts workgroups each; every
// entry already checked by its builder), one launch.
hipError_t launch_seed_mfma_group(const GroupTab& g, int kind, int ksteps, hipStream_t st);

// Bytes addressable from each pointer of a ScanMfmaArgs.
struct ScanExtents {
  size_t xop, hx, qop, slots, lcnts, gthr, tile_list, cents, q32, qtab, xsc, qfrag, qsc;
};
// Host-side check of one scan launch with the arguments launch_scan_mfma (or
// one problem of the persistent form) receives: every global address the
// kernel derives from `a` stays inside `ex`, every tile copy lands inside its
// LDS buffer and stages the row chunk the matrix-core reads of that LDS
// position expect, and a tile's appends fit the slot segments.  The residual
// operand's block lists are taken to lie in [0, L) (checked when the
// partition is built).  Returns the number of violations; *first (if not
// null) receives a description of the first.
int64_t scan_check_launch(const ScanMfmaArgs& a, int kind, int ksteps, int V, int nsplit, const ScanExtents& ex,
                          int L, std::string* first);

template <typename F>
hipError_t launch_scan_exact(const ScanExactArgs<F>& a, int qw, int nsplit, hipStream_t st);
template <typename F>
hipError_t launch_finalize_exact(const FinalizeArgs<F>& a, hipStream_t st);
template <typename F>
hipError_t launch_range(const RangeArgs<F>& a, int pass, int nsplit, hipStream_t st);
template <typename F>
hipError_t launch_pairwise(const PairwiseArgs<F>& a, int form, int nsplit, hipStream_t st);
template <typename F>
hipError_t launch_unord(const typename OrdKey<F>::T* keys, int64_t nnz, F* dist, hipStream_t st);
hipError_t launch_gather_offsets(const int64_t* qs_offsets, int64_t m, int nsplit, int64_t total, int64_t* out,
                                 hipStream_t st);
template <typename F>
hipError_t launch_lof(const LofArgs<F>& a, int knn_outlier, hipStream_t st);

// rocPRIM wrappers (sort.hip): exclusive scan of int64 counts; stable segmented
// sort of (order key, index) pairs.
hipError_t scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t count, void* tmp, size_t& tmp_bytes,
                              hipStream_t st);
template <typename K>
hipError_t segmented_sort_pairs(const K* keys_in, K* keys_out, const int64_t* vals_in, int64_t* vals_out,
                                int64_t nnz, int64_t nseg, const int64_t* offsets, void* tmp, size_t& tmp_bytes,
                                hipStream_t st);

}  // namespace fa
