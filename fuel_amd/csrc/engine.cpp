// Orchestration of the exact kNN, range, pairwise and outlier paths on one
// index: workspace sizing, query batching, split choice and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

namespace fa {

namespace {

constexpr int EXACT_V = 8;  // slot = 512 composites
constexpr int EXACT_NS = 64 * EXACT_V;
constexpr size_t SLOT_BUDGET = size_t(1) << 30;  // bytes of candidate slots per batch

int exact_qw(int64_t nq) { return nq >= 64 ? 16 : (nq >= 8 ? 8 : 4); }

// kNN with k beyond one slot's KK = EXACT_NS / 2 (fuel's KNNHeap takes any k,
// src/api/search.rs:80-151): the top-KK composites are found in rounds of RK =
// EXACT_NS / 2, each an exact scan that keeps only the candidates ordering
// after the previous round's last composite; the rounds' sorted lists then
// concatenate to the sorted top KK, which the row writer packs like
// write_knn_row.
template <typename F>
int32_t knn_exact_rounds(fuel_amd_index* ix, QueryCtx* ctx, const F* Q, int64_t m, int64_t ldq, int k, int mode,
                         int64_t* out_idx, F* out_dist, hipStream_t st, const int64_t* qrows, int64_t qbase) {
  using T = typename Comp<F>::T;
  const int KK = k + (mode != EXCL_NONE ? 1 : 0);
  constexpr int RK = EXACT_NS / 2;
  const int rounds = (KK + RK - 1) / RK;
  const int64_t stride = int64_t(rounds) * RK;
  const int cus = device_cu_count(ix->device);
  int64_t done = 0;
  while (done < m) {
    if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
    int64_t chunk = m - done;
    const int qw = exact_qw(chunk);
    int nsplit = choose_splits((chunk + qw - 1) / qw, ix->n, cus, 512);
    const size_t per_q = size_t(nsplit) * EXACT_NS * sizeof(T) + size_t(stride) * sizeof(T);
    chunk = std::min<int64_t>(chunk, std::max<int64_t>(qw, int64_t(SLOT_BUDGET / per_q)));
    hipError_t e = ctx->slots.ensure(size_t(nsplit) * EXACT_NS * sizeof(T) * chunk);
    if (e == hipSuccess) e = ctx->lcnts.ensure(sizeof(int) * size_t(nsplit) * chunk);
    if (e == hipSuccess) e = ctx->raw.ensure(sizeof(T) * size_t(stride) * chunk);
    if (e == hipSuccess) e = ctx->lo.ensure(sizeof(T) * size_t(chunk));
    if (e == hipSuccess) e = ctx->rcnt.ensure(sizeof(int) * size_t(chunk));
    if (e == hipSuccess) e = launch_fill_u32(ctx->rcnt.as<uint32_t>(), chunk, 0u, st);
    if (e != hipSuccess) return hip_fail(e, "round workspaces");
    int64_t rps = (ix->n + nsplit - 1) / nsplit;
    rps = std::max<int64_t>(256, (rps + 255) & ~int64_t(255));
    nsplit = (int)std::max<int64_t>(1, (ix->n + rps - 1) / rps);
    for (int r = 0; r < rounds; ++r) {
      if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
      ScanExactArgs<F> sa{};
      sa.X = ix->X.as<F>();
      sa.ldx = ix->ld;
      sa.n = ix->n;
      sa.d = ix->d;
      sa.Q = qrows ? Q : Q + done * ldq;
      sa.ldq = ldq;
      sa.qrows = qrows ? qrows + done : nullptr;
      sa.nq = chunk;
      sa.metric = ix->metric;
      sa.arith = ix->arith;
      sa.KP = RK;
      sa.rows_per_split = rps;
      sa.slots = ctx->slots.as<T>();
      sa.lcnts = ctx->lcnts.as<int>();
      sa.lo = r > 0 ? ctx->lo.as<T>() : nullptr;
      if (ctx->timed) timing_mark(ix, st, true, "scan_exact_kernel");
      e = launch_scan_exact<F>(sa, qw, nsplit, st);
      if (ctx->timed) timing_mark(ix, st, false, nullptr);
      if (e != hipSuccess) return hip_fail(e, "scan_exact launch");
      FinalizeArgs<F> fa_{};
      fa_.slots = sa.slots;
      fa_.lcnts = sa.lcnts;
      fa_.slot_stride = EXACT_NS;
      fa_.nsplit = nsplit;
      fa_.nq = chunk;
      fa_.KK = RK;
      fa_.k = RK;
      fa_.mode = EXCL_NONE;
      fa_.raw_out = ctx->raw.as<T>();
      fa_.raw_stride = stride;
      fa_.raw_off = int64_t(r) * RK;
      fa_.lo_out = ctx->lo.as<T>();
      fa_.cnt_out = ctx->rcnt.as<int>();
      e = launch_finalize_exact<F>(fa_, st);
      if (e != hipSuccess) return hip_fail(e, "finalize launch");
    }
    LongRowArgs<F> la{};
    la.raw = ctx->raw.as<T>();
    la.stride = stride;
    la.cnt = ctx->rcnt.as<int>();
    la.nq = chunk;
    la.k = k;
    la.mode = mode;
    la.qrows = qrows ? qrows + done : nullptr;
    la.qbase = qrows ? qbase : qbase + done;
    la.out_idx = qrows ? out_idx : out_idx + done * k;
    la.out_dist = qrows ? out_dist : out_dist + done * k;
    e = launch_write_long_rows<F>(la, st);
    if (e != hipSuccess) return hip_fail(e, "kNN rows");
    ctx->stats.splits = nsplit;
    ctx->stats.kp = KK;
    done += chunk;
  }
  return FUEL_AMD_OK;
}

// Exact kNN of m query slots.  With qrows, slot i is query row qrows[i] of Q
// (a device array) and its result goes to output row qrows[i].
template <typename F>
int32_t knn_exact(fuel_amd_index* ix, QueryCtx* ctx, const F* Q, int64_t m, int64_t ldq, int k, int mode,
                  int64_t* out_idx, F* out_dist, hipStream_t st, const int64_t* qrows = nullptr, int64_t qbase = 0) {
  using T = typename Comp<F>::T;
  const int KK = k + (mode != EXCL_NONE ? 1 : 0);
  if (KK > EXACT_NS / 2) return knn_exact_rounds<F>(ix, ctx, Q, m, ldq, k, mode, out_idx, out_dist, st, qrows, qbase);
  const int cus = device_cu_count(ix->device);
  int64_t done = 0;
  while (done < m) {
    if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
    int64_t chunk = m - done;
    const int qw = exact_qw(chunk);
    int nsplit = choose_splits((chunk + qw - 1) / qw, ix->n, cus, 512);
    const size_t per_q = size_t(nsplit) * EXACT_NS * sizeof(T);
    chunk = std::min<int64_t>(chunk, std::max<int64_t>(qw, int64_t(SLOT_BUDGET / per_q)));
    hipError_t e = ctx->slots.ensure(per_q * chunk);
    if (e == hipSuccess) e = ctx->lcnts.ensure(sizeof(int) * size_t(nsplit) * chunk);
    if (e != hipSuccess) return hip_fail(e, "slot workspace");
    ScanExactArgs<F> sa{};
    sa.X = ix->X.as<F>();
    sa.ldx = ix->ld;
    sa.n = ix->n;
    sa.d = ix->d;
    sa.Q = qrows ? Q : Q + done * ldq;
    sa.ldq = ldq;
    sa.qrows = qrows ? qrows + done : nullptr;
    sa.nq = chunk;
    sa.metric = ix->metric;
    sa.arith = ix->arith;
    sa.KP = KK;
    int64_t rps = (ix->n + nsplit - 1) / nsplit;
    rps = std::max<int64_t>(256, (rps + 255) & ~int64_t(255));
    sa.rows_per_split = rps;
    nsplit = (int)std::max<int64_t>(1, (ix->n + rps - 1) / rps);
    sa.slots = ctx->slots.as<T>();
    sa.lcnts = ctx->lcnts.as<int>();
    if (ctx->timed) timing_mark(ix, st, true, "scan_exact_kernel");
    e = launch_scan_exact<F>(sa, qw, nsplit, st);
    if (ctx->timed) timing_mark(ix, st, false, nullptr);
    if (e != hipSuccess) return hip_fail(e, "scan_exact launch");
    FinalizeArgs<F> fa_{};
    fa_.slots = sa.slots;
    fa_.lcnts = sa.lcnts;
    fa_.slot_stride = EXACT_NS;
    fa_.nsplit = nsplit;
    fa_.nq = chunk;
    fa_.KK = KK;
    fa_.k = k;
    fa_.mode = mode;
    fa_.qrows = qrows ? qrows + done : nullptr;
    fa_.qbase = qrows ? qbase : qbase + done;
    fa_.out_idx = qrows ? out_idx : out_idx + done * k;
    fa_.out_dist = qrows ? out_dist : out_dist + done * k;
    e = launch_finalize_exact<F>(fa_, st);
    if (e != hipSuccess) return hip_fail(e, "finalize launch");
    ctx->stats.splits = nsplit;
    ctx->stats.kp = KK;
    done += chunk;
  }
  return FUEL_AMD_OK;
}

// Range search of m query rows already in device memory (pitch ix->ld), with a
// common radius or per-query radii (d_radii[i * radii_stride]) and optional
// per-query lower limits (d_skips[i]: keep distances >= it), into a host CSR.
template <typename F>
int32_t range_core(fuel_amd_index* ix, const F* d_q, int64_t m, double radius, const F* d_radii,
                   int64_t radii_stride, int exclude_self, fuel_amd_csr* out, const F* d_skips = nullptr) {
  using K = typename OrdKey<F>::T;
  hipStream_t st = ix->stream;
  const size_t es = sizeof(F);
  out->m = m;
  out->offsets = (int64_t*)std::malloc(sizeof(int64_t) * (m + 1));
  if (!out->offsets) return fail(FUEL_AMD_ERR_NOMEM, "host allocation failed");
  if (m == 0) {
    out->offsets[0] = 0;
    return FUEL_AMD_OK;
  }
  const int QW = 16;
  const int cus = device_cu_count(ix->device);
  int nsplit = choose_splits((m + QW - 1) / QW, ix->n, cus, 1024);
  int64_t rps = (ix->n + nsplit - 1) / nsplit;
  rps = std::max<int64_t>(256, (rps + 255) & ~int64_t(255));
  nsplit = (int)std::max<int64_t>(1, (ix->n + rps - 1) / rps);
  const size_t ncs = size_t(m) * nsplit;
  hipError_t e = ix->ws_a.ensure(sizeof(int64_t) * (ncs + 1));           // counts
  if (e == hipSuccess) e = ix->ws_b.ensure(sizeof(int64_t) * (ncs + 1));  // offsets
  if (e == hipSuccess) e = ix->ws_c.ensure(sizeof(int64_t) * (m + 1));    // per-query offsets
  if (e != hipSuccess) return hip_fail(e, "range workspace");
  RangeArgs<F> ra{};
  ra.X = ix->X.as<F>();
  ra.ldx = ix->ld;
  ra.n = ix->n;
  ra.d = ix->d;
  ra.Q = d_q;
  ra.ldq = ix->ld;
  ra.nq = m;
  ra.qbase = 0;
  ra.metric = ix->metric;
  ra.arith = ix->arith;
  ra.radius = (F)radius;  // python/search.rs:785 (`radius as f32`)
  ra.radii = d_radii;
  ra.radii_stride = radii_stride;
  ra.skips = d_skips;
  ra.exclude_self = exclude_self ? 1 : 0;
  ra.nsplit = nsplit;
  ra.rows_per_split = rps;
  ra.counts = ix->ws_a.as<int64_t>();
  e = launch_range<F>(ra, 0, nsplit, st);
  if (e != hipSuccess) return hip_fail(e, "range count launch");
  size_t tmp_bytes = 0;
  e = scan_exclusive_i64(ra.counts, ix->ws_b.as<int64_t>(), (int64_t)ncs, nullptr, tmp_bytes, st);
  if (e == hipSuccess) e = ix->ws_tmp.ensure(tmp_bytes);
  if (e == hipSuccess)
    e = scan_exclusive_i64(ra.counts, ix->ws_b.as<int64_t>(), (int64_t)ncs, ix->ws_tmp.p, tmp_bytes, st);
  int64_t tail[2] = {0, 0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(&tail[0], ix->ws_b.as<int64_t>() + ncs - 1, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&tail[1], ra.counts + ncs - 1, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "range count");
  const int64_t nnz = tail[0] + tail[1];
  out->nnz = nnz;
  out->index = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(1, nnz));
  out->dist = std::malloc(es * std::max<int64_t>(1, nnz));
  if (!out->index || !out->dist) return fail(FUEL_AMD_ERR_NOMEM, "host allocation failed");
  e = launch_gather_offsets(ix->ws_b.as<int64_t>(), m, nsplit, nnz, ix->ws_c.as<int64_t>(), st);
  if (e != hipSuccess) return hip_fail(e, "range offsets");
  if (nnz > 0) {
    if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
    if (nnz > int64_t(0xFFFFFFFFll)) return fail(FUEL_AMD_ERR_UNSUPPORTED, "more than 2^32 range results");
    // keys, idx (unsorted) then sorted
    const size_t key_bytes = (sizeof(K) * size_t(nnz) + 15) & ~size_t(15);
    e = ix->ws_d.ensure(key_bytes + 8 * size_t(nnz));
    if (e == hipSuccess) e = ix->ws_e.ensure(key_bytes + 8 * size_t(nnz));
    if (e != hipSuccess) return hip_fail(e, "range result workspace");
    K* k_in = ix->ws_d.as<K>();
    int64_t* i_in = reinterpret_cast<int64_t*>(ix->ws_d.as<char>() + key_bytes);
    K* k_out = ix->ws_e.as<K>();
    int64_t* i_out = reinterpret_cast<int64_t*>(ix->ws_e.as<char>() + key_bytes);
    ra.offsets = ix->ws_b.as<int64_t>();
    ra.out_keys = k_in;
    ra.out_idx = i_in;
    e = launch_range<F>(ra, 1, nsplit, st);
    if (e != hipSuccess) return hip_fail(e, "range fill launch");
    tmp_bytes = 0;
    e = segmented_sort_pairs<K>(k_in, k_out, i_in, i_out, nnz, m, ix->ws_c.as<int64_t>(), nullptr, tmp_bytes, st);
    if (e == hipSuccess) e = ix->ws_tmp.ensure(tmp_bytes);
    if (e != hipSuccess) return hip_fail(e, "range sort (workspace query)");
    e = segmented_sort_pairs<K>(k_in, k_out, i_in, i_out, nnz, m, ix->ws_c.as<int64_t>(), ix->ws_tmp.p, tmp_bytes,
                                st);
    if (e != hipSuccess) return hip_fail(e, "range sort (segmented radix sort)");
    e = launch_unord<F>(k_out, nnz, reinterpret_cast<F*>(k_in), st);
    if (e != hipSuccess) return hip_fail(e, "range sort (key decode)");
    e = hipMemcpyAsync(out->index, i_out, 8 * size_t(nnz), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out->dist, k_in, es * size_t(nnz), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return hip_fail(e, "range result download");
  }
  e = hipMemcpyAsync(out->offsets, ix->ws_c.p, sizeof(int64_t) * (m + 1), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "range download");
  return FUEL_AMD_OK;
}

template <typename F>
int32_t range_impl(fuel_amd_index* ix, const void* query, int64_t m, double radius, int exclude_self,
                   fuel_amd_csr* out) {
  hipError_t e = ix->hq.ensure(size_t(std::max<int64_t>(m, 1)) * ix->ld * sizeof(F));
  if (e != hipSuccess) return hip_fail(e, "query buffer");
  if (m > 0 && ix->d > 0) {
    e = upload_rows(ix->hq.p, ix->ld, query, m, ix->d, sizeof(F), ix->upload, ix->stream);
    if (e != hipSuccess) return hip_fail(e, "query upload");
  }
  return range_core<F>(ix, ix->hq.as<F>(), m, radius, nullptr, 0, exclude_self, out);
}

// KNNHeap's result (src/api/search.rs:80-151): every point whose key is <= the
// k-th key, sorted by (key, index) -- the exact k-th key of each query from the
// kNN path, then an inclusive range search at that radius.
template <typename F>
int32_t knn_ties_impl(fuel_amd_index* ix, const void* query, int64_t m, int k, fuel_amd_csr* out) {
  const size_t es = sizeof(F);
  hipStream_t st = ix->stream;
  hipError_t e = ix->hq.ensure(size_t(std::max<int64_t>(m, 1)) * ix->ld * es);
  if (e != hipSuccess) return hip_fail(e, "query buffer");
  if (k == 0 || m == 0) {  // k = 0 -> empty lists (src/api/search.rs:106-108)
    out->m = m;
    out->nnz = 0;
    out->offsets = (int64_t*)std::calloc(size_t(m + 1), sizeof(int64_t));
    out->index = (int64_t*)std::malloc(8);
    out->dist = std::malloc(es);
    if (!out->offsets || !out->index || !out->dist) return fail(FUEL_AMD_ERR_NOMEM, "host allocation failed");
    return FUEL_AMD_OK;
  }
  if (ix->d > 0) {
    e = upload_rows(ix->hq.p, ix->ld, query, m, ix->d, es, ix->upload, st);
    if (e != hipSuccess) return hip_fail(e, "query upload");
  }
  e = ix->hout_idx.ensure(size_t(m) * k * 8);
  if (e == hipSuccess) e = ix->hout_dist.ensure(size_t(m) * k * es);
  if (e != hipSuccess) return hip_fail(e, "kNN buffers");
  int32_t rc = knn_run(ix, &ix->ctx, ix->hq.p, m, ix->ld, k, EXCL_NONE, ix->hout_idx.as<int64_t>(), ix->hout_dist.p, st);
  if (rc) return rc;
  // the k-th key of every query (+inf when fewer than k points exist: then all of them)
  return range_core<F>(ix, ix->hq.as<F>(), m, 0.0, ix->hout_dist.as<F>() + (k - 1), k, 0, out);
}

template <typename F>
int32_t pairwise_impl(fuel_amd_index* ix, int form, void* out) {
  hipStream_t st = ix->stream;
  const int64_t n = ix->n;
  const size_t count = form == 0 ? size_t(n) * n : (n >= 2 ? size_t(n) * (n - 1) / 2 : 0);
  if (count == 0) return FUEL_AMD_OK;
  hipError_t e = ix->ws_a.ensure(count * sizeof(F));
  if (e != hipSuccess) return hip_fail(e, "pairwise output");
  const int cus = device_cu_count(ix->device);
  int nsplit = choose_splits((n + 15) / 16, n, cus, 1024);
  int64_t rps = (n + nsplit - 1) / nsplit;
  rps = std::max<int64_t>(256, (rps + 255) & ~int64_t(255));
  nsplit = (int)std::max<int64_t>(1, (n + rps - 1) / rps);
  PairwiseArgs<F> pa{};
  pa.X = ix->X.as<F>();
  pa.ldx = ix->ld;
  pa.n = n;
  pa.d = ix->d;
  pa.metric = ix->metric;
  pa.arith = ix->arith;
  pa.rows_per_split = rps;
  pa.out = ix->ws_a.as<F>();
  e = launch_pairwise<F>(pa, form, nsplit, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out, pa.out, count * sizeof(F), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "pairwise");
  return FUEL_AMD_OK;
}

template <typename F>
int32_t outlier_impl(fuel_amd_index* ix, int k, int knn_outlier, void* out_scores) {
  hipStream_t st = ix->stream;
  const int64_t n = ix->n;
  if (n == 0) return FUEL_AMD_OK;
  const int k_eff = (int)std::min<int64_t>(k, n - 1);  // src/outlier/lof.rs:25
  hipError_t e = ix->ws_a.ensure(8 * size_t(n) * std::max(1, k_eff));
  if (e == hipSuccess) e = ix->ws_b.ensure(sizeof(F) * size_t(n) * std::max(1, k_eff));
  if (e == hipSuccess) e = ix->ws_c.ensure(sizeof(F) * size_t(n) * 3);
  if (e != hipSuccess) return hip_fail(e, "outlier workspace");
  if (k_eff > 0) {
    int32_t rc = knn_run(ix, &ix->ctx, ix->X.p, n, ix->ld, k_eff, EXCL_FOR_EACH_KNN, ix->ws_a.as<int64_t>(), ix->ws_b.p, st);
    if (rc) return rc;
  }
  LofArgs<F> la{};
  la.nb_idx = ix->ws_a.as<int64_t>();
  la.nb_dist = ix->ws_b.as<F>();
  la.n = n;
  la.k = k_eff;
  la.kdist = ix->ws_c.as<F>();
  la.lrd = la.kdist + n;
  la.out = la.lrd + n;
  e = launch_lof<F>(la, knn_outlier, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out_scores, la.out, sizeof(F) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "outlier scores");
  return FUEL_AMD_OK;
}

}  // namespace

// The exact kNN of m query rows in one pass whose candidate slots and list
// counts live in caller-provided device memory (`scratch`, `bytes`): no
// allocation, no host synchronization.  The split count is the one knn_exact
// would choose, reduced until the pass fits; a pass that does not fit with one
// split is refused.
template <typename F>
int32_t knn_exact_in_scratch(fuel_amd_index* ix, const F* Q, int64_t m, int64_t ldq, int k, int mode, int64_t* out_idx,
                             F* out_dist, hipStream_t st, void* scratch, size_t bytes) {
  using T = typename Comp<F>::T;
  const int KK = k + (mode != EXCL_NONE ? 1 : 0);
  if (KK > EXACT_NS / 2)
    return fail(FUEL_AMD_ERR_UNSUPPORTED,
                "the allocation-free kNN takes k + exclusion <= " + std::to_string(EXACT_NS / 2));
  const int qw = exact_qw(m);
  int nsplit = choose_splits((m + qw - 1) / qw, ix->n, device_cu_count(ix->device), 512);
  auto need = [&](int s) {
    return ((size_t(s) * EXACT_NS * sizeof(T) * m + 255) & ~size_t(255)) + sizeof(int) * size_t(s) * m;
  };
  while (nsplit > 1 && need(nsplit) > bytes) nsplit = (nsplit + 1) / 2;
  if (need(nsplit) > bytes)
    return fail(FUEL_AMD_ERR_INVALID, "the allocation-free kNN holds at most " +
                                          std::to_string(bytes / (EXACT_NS * sizeof(T) + sizeof(int))) +
                                          " queries per call (" + std::to_string(m) + " given)");
  int64_t rps = (ix->n + nsplit - 1) / nsplit;
  rps = std::max<int64_t>(256, (rps + 255) & ~int64_t(255));
  nsplit = (int)std::max<int64_t>(1, (ix->n + rps - 1) / rps);
  ScanExactArgs<F> sa{};
  sa.X = ix->X.as<F>();
  sa.ldx = ix->ld;
  sa.n = ix->n;
  sa.d = ix->d;
  sa.Q = Q;
  sa.ldq = ldq;
  sa.nq = m;
  sa.metric = ix->metric;
  sa.arith = ix->arith;
  sa.KP = KK;
  sa.rows_per_split = rps;
  // every (split, query) list count is written by its workgroup before the
  // finalize reads it, and slots are read only below those counts: nothing in
  // the scratch needs clearing between calls
  sa.slots = static_cast<T*>(scratch);
  sa.lcnts = reinterpret_cast<int*>(static_cast<char*>(scratch) +
                                    ((size_t(nsplit) * EXACT_NS * sizeof(T) * m + 255) & ~size_t(255)));
  hipError_t e = launch_scan_exact<F>(sa, qw, nsplit, st);
  if (e != hipSuccess) return hip_fail(e, "scan_exact launch");
  FinalizeArgs<F> fa_{};
  fa_.slots = sa.slots;
  fa_.lcnts = sa.lcnts;
  fa_.slot_stride = EXACT_NS;
  fa_.nsplit = nsplit;
  fa_.nq = m;
  fa_.KK = KK;
  fa_.k = k;
  fa_.mode = mode;
  fa_.qbase = 0;
  fa_.out_idx = out_idx;
  fa_.out_dist = out_dist;
  e = launch_finalize_exact<F>(fa_, st);
  if (e != hipSuccess) return hip_fail(e, "finalize launch");
  return FUEL_AMD_OK;
}

int32_t knn_run_scratch(fuel_amd_index* ix, const void* d_query, int64_t m, int64_t ldq, int k, int exclude,
                        int64_t* d_out_idx, void* d_out_dist, hipStream_t st, void* scratch, size_t bytes) {
  if (m == 0 || k == 0) return FUEL_AMD_OK;
  if (!d_query || !d_out_idx || !d_out_dist) return fail(FUEL_AMD_ERR_INVALID, "NULL buffer");
  return ix->dtype == FUEL_AMD_F32
             ? knn_exact_in_scratch<float>(ix, (const float*)d_query, m, ldq, k, exclude, d_out_idx,
                                           (float*)d_out_dist, st, scratch, bytes)
             : knn_exact_in_scratch<double>(ix, (const double*)d_query, m, ldq, k, exclude, d_out_idx,
                                            (double*)d_out_dist, st, scratch, bytes);
}

int32_t knn_exact_rows(fuel_amd_index* ix, QueryCtx* ctx, const float* Q, int64_t ldq, const int64_t* d_rows,
                       int64_t nrows, int64_t qbase, int k, int exclude, int64_t* out_idx, float* out_dist,
                       hipStream_t st) {
  return knn_exact<float>(ix, ctx, Q, nrows, ldq, k, exclude, out_idx, out_dist, st, d_rows, qbase);
}

int knn_kind(const fuel_amd_index* ix, int mode, int k, int exclude) {
  int kind = mode == FUEL_AMD_MODE_BF16    ? 0
             : mode == FUEL_AMD_MODE_FP32  ? 1
             : mode == FUEL_AMD_MODE_BF16R ? 2
             : mode == FUEL_AMD_MODE_FP8R  ? 3
                                           : -1;
  if (kind == 3 && !mfma_supported(ix, 3, k, exclude)) kind = 2;  // no E4M3 variant this wide
  if (kind == 2 && !mfma_supported(ix, 2, k, exclude)) kind = 0;  // no residual variant this wide
  if (kind >= 0 && !mfma_supported(ix, kind, k, exclude)) kind = -1;
  return kind;
}

int32_t knn_run(fuel_amd_index* ix, QueryCtx* ctx, const void* d_query, int64_t m, int64_t ldq, int k,
                int exclude, int64_t* d_out_idx, void* d_out_dist, hipStream_t st) {
  if (m == 0 || k == 0) return FUEL_AMD_OK;
  if (!d_query || !d_out_idx || !d_out_dist) return fail(FUEL_AMD_ERR_INVALID, "NULL buffer");
  auto t0 = std::chrono::steady_clock::now();
  int mode = ix->opts.mode;
  if (mode == FUEL_AMD_MODE_AUTO) {
    int32_t rc = resolve_auto_mode(ix, ctx, k, st, &mode);
    if (rc) return rc;
  }
  ctx->stats = fuel_amd_stats{};
  ctx->stats.queries = m;
  ctx->fail_rows.clear();
  const int kind = knn_kind(ix, mode, k, exclude);
  int32_t rc;
  if (kind >= 0) {
    static const int kModeOfKind[4] = {FUEL_AMD_MODE_BF16, FUEL_AMD_MODE_FP32, FUEL_AMD_MODE_BF16R,
                                       FUEL_AMD_MODE_FP8R};
    ctx->stats.mode_used = kModeOfKind[kind];
    rc = knn_mfma(ix, ctx, (const float*)d_query, m, ldq, k, exclude, kind, d_out_idx, (float*)d_out_dist, st);
  } else if (ix->dtype == FUEL_AMD_F32) {
    ctx->stats.mode_used = FUEL_AMD_MODE_EXACT;
    rc = knn_exact<float>(ix, ctx, (const float*)d_query, m, ldq, k, exclude, d_out_idx, (float*)d_out_dist, st);
  } else {
    ctx->stats.mode_used = FUEL_AMD_MODE_EXACT;
    rc = knn_exact<double>(ix, ctx, (const double*)d_query, m, ldq, k, exclude, d_out_idx, (double*)d_out_dist,
                           st);
  }
  ctx->stats.ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (ctx->publish) {
    std::lock_guard<std::mutex> lock(ix->stats_mu);
    ix->last_stats = ctx->stats;
    ix->last_fail_rows = ctx->fail_rows;
  }
  return rc;
}

int32_t range_run(fuel_amd_index* ix, const void* query, int64_t m, double radius, int exclude_self,
                  fuel_amd_csr* out) {
  int32_t rc = ix->dtype == FUEL_AMD_F32 ? range_impl<float>(ix, query, m, radius, exclude_self, out)
                                         : range_impl<double>(ix, query, m, radius, exclude_self, out);
  if (rc) fuel_amd_csr_free(out);
  return rc;
}

int32_t range_self_run(fuel_amd_index* ix, double radius, fuel_amd_csr* out) {
  int32_t rc = ix->dtype == FUEL_AMD_F32
                   ? range_core<float>(ix, ix->X.as<float>(), ix->n, radius, nullptr, 0, 0, out)
                   : range_core<double>(ix, ix->X.as<double>(), ix->n, radius, nullptr, 0, 0, out);
  if (rc) fuel_amd_csr_free(out);
  return rc;
}

int32_t range_window_run(fuel_amd_index* ix, const void* d_q, int64_t m, const void* d_cutoff, const void* d_skip,
                         fuel_amd_csr* out) {
  int32_t rc = ix->dtype == FUEL_AMD_F32
                   ? range_core<float>(ix, (const float*)d_q, m, 0.0, (const float*)d_cutoff, 1, 0, out,
                                       (const float*)d_skip)
                   : range_core<double>(ix, (const double*)d_q, m, 0.0, (const double*)d_cutoff, 1, 0, out,
                                        (const double*)d_skip);
  if (rc) fuel_amd_csr_free(out);
  return rc;
}

int32_t knn_ties_run(fuel_amd_index* ix, const void* query, int64_t m, int k, fuel_amd_csr* out) {
  int32_t rc = ix->dtype == FUEL_AMD_F32 ? knn_ties_impl<float>(ix, query, m, k, out)
                                         : knn_ties_impl<double>(ix, query, m, k, out);
  if (rc) fuel_amd_csr_free(out);
  return rc;
}

int32_t pairwise_run(fuel_amd_index* ix, int form, void* out) {
  return ix->dtype == FUEL_AMD_F32 ? pairwise_impl<float>(ix, form, out) : pairwise_impl<double>(ix, form, out);
}

int32_t outlier_run(fuel_amd_index* ix, int k, int knn_outlier, void* out_scores) {
  return ix->dtype == FUEL_AMD_F32 ? outlier_impl<float>(ix, k, knn_outlier, out_scores)
                                   : outlier_impl<double>(ix, k, knn_outlier, out_scores);
}

}  // namespace fa
