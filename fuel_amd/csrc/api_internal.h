// Host-side internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fuel_amd.h"

namespace fa {

// A grow-only device allocation reused across calls (no allocation in steady state).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;  // false: p is borrowed from the caller and never freed here
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void release();
  hipError_t ensure(size_t want);
  void borrow(const void* ptr, size_t nbytes) {
    release();
    p = const_cast<void*>(ptr);
    bytes = nbytes;
    owned = false;
  }
  template <typename T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

int32_t fail(int32_t code, const std::string& msg);
int32_t hip_fail(hipError_t e, const char* what);
int32_t check_device(int dev);
int device_cu_count(int dev);
bool cancelled();
int choose_splits(int64_t qblocks, int64_t n, int cus, int64_t min_rows);
// Host rows of d elements (es bytes each) -> device rows of pitch ld elements:
// one contiguous host-to-device copy (through `stage` when the pitches differ),
// then the padding is applied by a device-side 2D copy.  A 2D copy straight from
// pageable host memory moves one row at a time.
hipError_t upload_rows(void* dst, int64_t ld, const void* host, int64_t rows, int d, size_t es, DevBuf& stage,
                       hipStream_t st);

// Query-side workspaces of one caller of the kNN path.  The index owns one,
// used by the synchronous entry points under the index mutex; each serving
// worker owns another, so batches on different streams run concurrently
// against one resident index (whose rows and operand copies are read-only).
struct QueryCtx {
  struct LevelWs {  // one per level of the certification chain
    DevBuf Qop, Q32, qnorm, qint, slots, lcnts, fail_count, fail_list, fail_theta, gthr, qtab, qs2, seed;
    DevBuf qfrag, qsc;  // E4M3 residual scan: per-(list, query) B fragments and scales
  } lv[3];
  DevBuf slots, lcnts;  // exact scan
  DevBuf raw, lo, rcnt;  // exact kNN in rounds (k beyond one slot)
  fuel_amd_stats stats{};
  std::vector<int64_t> fail_rows;  // query rows whose first-level certificate failed (last call)
  bool timed = false;  // record kernel timing events (the index's own context)
  bool publish = true;  // knn_run copies stats and fail rows to the index's last-call record
  void release() {
    for (auto& w : lv) {
      w.Qop.release();
      w.Q32.release();
      w.qnorm.release();
      w.qint.release();
      w.slots.release();
      w.lcnts.release();
      w.fail_count.release();
      w.fail_list.release();
      w.fail_theta.release();
      w.gthr.release();
      w.qtab.release();
      w.qs2.release();
      w.seed.release();
      w.qfrag.release();
      w.qsc.release();
    }
    slots.release();
    lcnts.release();
    raw.release();
    lo.release();
    rcnt.release();
  }
};

}  // namespace fa

struct fuel_amd_index {
  int device = 0;
  int dtype = 0;
  int metric = 0;
  int64_t n = 0;
  int d = 0;
  int64_t ld = 0;  // padded row stride of the resident copy (elements)
  int arith = 0;   // rounding of the restated distance: fa::ARITH_AVX or fa::ARITH_FALLBACK
  fuel_amd_opts opts{};
  hipStream_t stream = nullptr;
  std::mutex mu;
  fa::DevBuf X;                          // resident rows (n x ld)
  fa::DevBuf upload;                     // staging of host query rows whose pitch differs from ld
  fa::DevBuf hq, hout_idx, hout_dist;    // staging for host-pointer calls
  fa::DevBuf ws_a, ws_b, ws_c, ws_d, ws_e, ws_tmp;  // range / pairwise / outlier workspaces
  // mixed-precision candidate path: one operand copy per kind (0 bf16, 1 f32,
  // 2 bf16 residual, 3 E4M3 residual), prepared once (under prep_mu) on first
  // use, read-only afterwards
  struct MfmaOperand {
    std::atomic<bool> ready{false};
    bool usable = true;  // false when some squared norm reaches MFMA_NORM2_LIMIT
    int ks = 0;  // K steps
    int kp = 0;  // padded K (elements)
    int64_t n_pad = 0;  // operand rows (kinds 2, 3: lists padded to 64-row blocks)
    fa::DevBuf Xop, hx, consts;  // operand copy, ||x||^2/2 (kinds 2, 3: of the residual), bound constants + flags
    int L = 0;                   // kinds 2, 3: lists, centroids (L x d), position -> row, list of each
    fa::DevBuf cents, rperm, tile_list, lconsts, lcount;  // 64-row block, per-list bound constants, rows
    fa::DevBuf xsc;              // kind 3: E8M0 row-scale codes
    fa::DevBuf loff;             // kinds 2, 3: first position of every list
  } mop[4];
  std::mutex prep_mu;
  // FUEL_AMD_MODE_AUTO: the mode chosen by the first kNN call's probe (-1 until
  // then), with the probe's outcome
  std::atomic<int> auto_mode{-1};
  std::mutex auto_mu;
  int64_t auto_probe_rows = 0, auto_probe_fail = 0;
  fa::QueryCtx ctx;  // the synchronous entry points' query workspaces and stats
  // Concurrent kNN calls (fuel's SearchIndex is shared by many threads): the
  // index's own context, stream and staging serve one call at a time under
  // `mu`; a kNN call that finds them busy takes a lane of its own (context,
  // stream, staging), made on first need and kept for later calls.
  struct CallLane {
    fa::QueryCtx ctx;
    hipStream_t stream = nullptr;
    fa::DevBuf upload, hq, hout_idx, hout_dist;
  };
  std::mutex lanes_mu;
  std::vector<std::unique_ptr<CallLane>> lanes;  // every lane made
  std::vector<CallLane*> lanes_idle;
  // the last finished kNN call's statistics and failed rows (any context)
  std::mutex stats_mu;
  fuel_amd_stats last_stats{};
  std::vector<int64_t> last_fail_rows;
  // profiling: event pairs around the dominant kernel of each call
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::string timed_kernel;
  void release_all() {
    X.release();
    hq.release();
    hout_idx.release();
    hout_dist.release();
    ctx.release();
    for (auto& L : lanes) {
      if (L->stream) {
        (void)hipStreamSynchronize(L->stream);
        (void)hipStreamDestroy(L->stream);
        L->stream = nullptr;
      }
      L->ctx.release();
      L->upload.release();
      L->hq.release();
      L->hout_idx.release();
      L->hout_dist.release();
    }
    lanes_idle.clear();
    lanes.clear();
    ws_a.release();
    ws_b.release();
    ws_c.release();
    ws_d.release();
    ws_e.release();
    ws_tmp.release();
    for (auto& m : mop) {
      m.Xop.release();
      m.hx.release();
      m.consts.release();
      m.cents.release();
      m.rperm.release();
      m.tile_list.release();
      m.lconsts.release();
      m.lcount.release();
      m.xsc.release();
      m.loff.release();
      m.ready = false;
    }
  }
};

namespace fa {
// kNN of m device query rows on stream st with the workspaces of ctx (the
// index's own unless a serving worker passes its own).
int32_t knn_run(fuel_amd_index* ix, QueryCtx* ctx, const void* d_query, int64_t m, int64_t ldq, int k,
                int exclude, int64_t* d_out_idx, void* d_out_dist, hipStream_t st);
// The exact kNN of m device query rows in one pass, its candidate slots and
// list counts in the given device memory: allocates nothing and never waits for
// the stream.  Refuses (FUEL_AMD_ERR_INVALID) a call that does not fit.
int32_t knn_run_scratch(fuel_amd_index* ix, const void* d_query, int64_t m, int64_t ldq, int k, int exclude,
                        int64_t* d_out_idx, void* d_out_dist, hipStream_t st, void* scratch, size_t bytes);
// Prepare the operand copy of the mode the index's options select (a no-op for
// the exact mode); the serving workers call the kNN path only after this.
int32_t prepare_mode(fuel_amd_index* ix, hipStream_t st);
int32_t range_run(fuel_amd_index* ix, const void* query, int64_t m, double radius, int exclude_self,
                  fuel_amd_csr* out);
int32_t knn_ties_run(fuel_amd_index* ix, const void* query, int64_t m, int k, fuel_amd_csr* out);
// Every row with skip[i] <= distance <= cutoff[i] of m device query rows (pitch
// ix->ld; cutoff and skip: m device values of the index dtype), sorted by
// (distance, index), as a host CSR.
int32_t range_window_run(fuel_amd_index* ix, const void* d_q, int64_t m, const void* d_cutoff, const void* d_skip,
                         fuel_amd_csr* out);
// Exact range of every indexed row against the index (self included), host CSR.
int32_t range_self_run(fuel_amd_index* ix, double radius, fuel_amd_csr* out);
// Density clustering over range_self_run (cluster.cpp).  variant 0: dbscan,
// 1: parallel_dbscan.
int32_t dbscan_run(fuel_amd_index* ix, double eps, int64_t min_points, int variant, int64_t* labels);
int32_t optics_run(fuel_amd_index* ix, double max_eps, int64_t min_points, int64_t* ordering, void* reach,
                   void* core, int64_t* pred, int64_t* labels);
int32_t pairwise_run(fuel_amd_index* ix, int form, void* out);
int32_t outlier_run(fuel_amd_index* ix, int k, int knn_outlier, void* out_scores);
// Mixed-precision kNN (f32 indexes): kind 0 bf16 MFMA, 1 f32 MFMA, 2 bf16 MFMA
// on the residual operand.  Returns
// FUEL_AMD_ERR_UNSUPPORTED when the shape has no MFMA variant.
// With d_rows, slot i is query row d_rows[i] of Q and writes output row d_rows[i].
// kp = candidates kept per split (0: from the index's widening factor); level =
// position in the certification chain (which also selects the query
// workspace); d_theta = optional per-slot starting thresholds (key units), the
// certified second pass of queries that failed a previous level.
// probe = true runs the first level only: the queries whose certificate fails
// are counted (stats.fallback_queries) but not answered.
int32_t knn_mfma(fuel_amd_index* ix, QueryCtx* ctx, const float* Q, int64_t m, int64_t ldq, int k, int exclude,
                 int kind, int64_t* out_idx, float* out_dist, hipStream_t st, const int64_t* d_rows = nullptr,
                 int64_t qbase = 0, int level = 0, int kp = 0, const float* d_theta = nullptr, bool probe = false);
// The mode FUEL_AMD_MODE_AUTO stands for on this index (decided once, by a
// probe with this k on the first call): FUEL_AMD_MODE_BF16, _BF16R or _EXACT.
int32_t resolve_auto_mode(fuel_amd_index* ix, QueryCtx* ctx, int k, hipStream_t st, int* mode);
bool mfma_supported(const fuel_amd_index* ix, int kind, int k, int exclude);
// The candidate kind the kNN path takes for a resolved mode (not AUTO) at this
// k and exclusion: 0 bf16, 1 f32, 2 residual bf16, 3 residual E4M3, -1 the
// exact scan (modes without an MFMA variant for the shape fall back as listed).
int knn_kind(const fuel_amd_index* ix, int mode, int k, int exclude);
// Query-side device bytes the first level of knn_mfma allocates for m queries
// of `kind` at k (0 for the exact scan or an operand not yet prepared):
// candidate slots and split counts, the query operand or per-list terms, the
// f32 query copy, norms, thresholds, the failure list and the seed minima.
int64_t mfma_level0_bytes(fuel_amd_index* ix, int kind, int64_t m, int k, int exclude);
// Operand type of the scan for a candidate kind: 0 bf16 (kinds 0 and 2), 1 f32,
// 2 E4M3 (kind 3).
inline int mfma_operand_type(int kind) { return kind == 2 ? 0 : (kind == 3 ? 2 : kind); }
// Launch plan of the next chunk of an MFMA kNN call: `remaining` of m_total
// queries, operand type mk (0 bf16, 1 f32, 2 E4M3), residual operand or not, K steps,
// KP kept candidates, n_pad operand rows, cus compute units.
struct MfmaChunkPlan {
  int V, NS, csub, waves, qpb, rows_tile, wg_per_cu, nsplit;
  int64_t chunk, nqb, nq_pad, rps;
  size_t slots_bytes;  // candidate slots the chunk allocates
};
MfmaChunkPlan plan_mfma_chunk(int mk, bool resid, int ks, int64_t m_total, int64_t remaining, int KP, int64_t n_pad,
                              int cus);
// Seed pass of a plain MFMA scan (ntiles = 0: none): sample tiles, the rows
// between them and the workgroups per query block.
struct SeedPlan {
  int64_t ntiles = 0, stride = 0;
  int parts = 0;
};
SeedPlan plan_seed(int mk, int ks, bool resid, bool has_theta, int flags, int64_t n_pad, int rows_tile, int KP);
// Padded residual layout of lists with cnt[j] rows (returns the operand rows).
int64_t resid_layout(const std::vector<uint32_t>& cnt, std::vector<int64_t>& pad, std::vector<int>& tl);
// Residual operand layout of n rows (f32, pitch ld): centroids (L x d), padded
// position -> row (RPERM_NONE for padding) and the list of every 64-row block.
// loff receives the first position of every list (L + 1 entries).
int32_t build_partition(const float* X, int64_t ld, int64_t n, int d, hipStream_t st, DevBuf& cents,
                        DevBuf& rperm, DevBuf& tile_list, DevBuf& lcount, DevBuf& loff, int& L, int64_t& n_rpad);
// Exact path on an explicit list of query rows (the certification fallback).
int32_t knn_exact_rows(fuel_amd_index* ix, QueryCtx* ctx, const float* Q, int64_t ldq, const int64_t* d_rows,
                       int64_t nrows, int64_t qbase, int k, int exclude, int64_t* out_idx, float* out_dist,
                       hipStream_t st);
// Record the start (begin=true) or end of a timed kernel on `st` (no-op unless enabled).
void timing_mark(fuel_amd_index* ix, hipStream_t st, bool begin, const char* name);
}  // namespace fa
