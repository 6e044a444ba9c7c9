// Serving loop over one resident index (continuous batching).
//
// Requests carry their own query rows and k.  Worker threads, each with its
// own stream, query workspaces (QueryCtx) and pinned staging, take turns at
// the waiting queue: a free worker admits, in arrival order, every waiting
// request with the head's k while the batch stays within max_batch queries and
// the estimated query-side device memory of all in-flight batches stays within
// hbm_budget.  It packs the admitted rows into one launch of the kNN path,
// and scatters the rows of the result back to the requests.  Requests that
// arrive while one batch runs are taken by the next free worker, so their
// kernels overlap with the running batch on the device.
//
// The kNN path answers every query row from that row alone (per-query
// thresholds, per-query certification), so a request's answer is the same
// whichever batch it joins: exactly what fuel_amd_knn returns for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

namespace fa {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0;
  int64_t m = 0;
  int k = 0;
  std::vector<char> q;  // m x d rows, index dtype
  std::vector<int64_t> idx;
  std::vector<char> dist;
  Clock::time_point t_submit, t_done;
  int32_t rc = FUEL_AMD_OK;
  std::string err;
  int64_t batch_queries = 0;
  bool done = false;
  bool claimed = false;  // a wait call owns the ticket
};

// Page-locked host memory, grown on demand (allocated, never registered).
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t ensure(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e == hipSuccess) bytes = want;
    return e;
  }
};

struct Worker {
  QueryCtx ctx;
  hipStream_t stream = nullptr;
  PinnedBuf hq, hidx, hdist;
  DevBuf dq, didx, ddist, stage;
  std::thread th;
};

}  // namespace
}  // namespace fa

struct fuel_amd_server {
  fuel_amd_index* ix = nullptr;
  int64_t max_batch = 4096;
  int64_t budget = int64_t(4) << 30;
  size_t es = 4;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<fa::Request*> queue;                   // waiting, arrival order
  std::unordered_map<uint64_t, fa::Request*> live;  // every ticket not yet waited for
  uint64_t next_id = 1;
  int64_t reserved = 0;  // estimated bytes of the in-flight batches
  int inflight = 0;
  bool stop = false;
  // statistics since the last reset: p50/p99 over the latencies of the last
  // LAT_WINDOW requests (a ring), max and mean over all of them
  static constexpr size_t LAT_WINDOW = size_t(1) << 20;
  std::vector<double> lat_ms;
  size_t lat_next = 0;  // ring position once the window is full
  double lat_max = 0.0, lat_sum = 0.0;
  int64_t n_req = 0, n_q = 0, n_batches = 0, max_bq = 0, n_failed = 0;
  double busy_s = 0.0;
  fa::Clock::time_point busy_since;
  std::vector<std::unique_ptr<fa::Worker>> workers;
};

namespace fa {
namespace {

// Query-side device bytes a batch of m queries may hold: the query copies of
// the certification levels, the outputs and the candidate slots (the scan
// spreads a small batch over many splits, so slots take at least ~64 MiB).
int64_t batch_bytes(const fuel_amd_server* s, int64_t m, int k) {
  // query rows and outputs (device and pinned staging), an allowance per query
  // for the certification chain's levels, and the first level's workspaces
  // (candidate slots dominate: ~130 MB for 64 queries against 10M rows)
  const int64_t per_q = s->ix->ld * int64_t(s->es) * 3 + int64_t(k) * (8 + int64_t(s->es)) + (int64_t(16) << 10);
  int mode = s->ix->opts.mode;
  if (mode == FUEL_AMD_MODE_AUTO) mode = s->ix->auto_mode.load(std::memory_order_acquire);
  return m * per_q + (int64_t(64) << 20) + mfma_level0_bytes(s->ix, knn_kind(s->ix, mode, k, 0), m, k, 0);
}

int32_t run_batch(fuel_amd_server* s, Worker* w, const std::vector<Request*>& batch, int64_t bm, int k) {
  fuel_amd_index* ix = s->ix;
  const size_t es = s->es;
  const size_t row_bytes = size_t(ix->d) * es;
  hipError_t e = w->hq.ensure(std::max<size_t>(1, bm * row_bytes));
  if (e == hipSuccess) e = w->hidx.ensure(size_t(bm) * k * 8);
  if (e == hipSuccess) e = w->hdist.ensure(size_t(bm) * k * es);
  if (e != hipSuccess) return hip_fail(e, "pinned staging");
  size_t off = 0;
  for (Request* r : batch) {
    if (!r->q.empty()) std::memcpy(static_cast<char*>(w->hq.p) + off, r->q.data(), r->q.size());
    off += size_t(r->m) * row_bytes;
  }
  e = w->dq.ensure(size_t(bm) * ix->ld * es);
  if (e == hipSuccess) e = w->didx.ensure(size_t(bm) * k * 8);
  if (e == hipSuccess) e = w->ddist.ensure(size_t(bm) * k * es);
  if (e != hipSuccess) return hip_fail(e, "batch buffers");
  if (ix->d > 0) {
    e = upload_rows(w->dq.p, ix->ld, w->hq.p, bm, ix->d, es, w->stage, w->stream);
    if (e != hipSuccess) return hip_fail(e, "query upload");
  }
  int32_t rc = knn_run(ix, &w->ctx, w->dq.p, bm, ix->ld, k, EXCL_NONE, w->didx.as<int64_t>(), w->ddist.p, w->stream);
  if (rc) return rc;
  e = hipMemcpyAsync(w->hidx.p, w->didx.p, size_t(bm) * k * 8, hipMemcpyDeviceToHost, w->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(w->hdist.p, w->ddist.p, size_t(bm) * k * es, hipMemcpyDeviceToHost, w->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
  if (e != hipSuccess) return hip_fail(e, "result download");
  size_t row = 0;
  for (Request* r : batch) {
    const size_t cnt = size_t(r->m) * k;
    r->idx.assign(static_cast<const int64_t*>(w->hidx.p) + row * k, static_cast<const int64_t*>(w->hidx.p) + row * k + cnt);
    r->dist.assign(static_cast<const char*>(w->hdist.p) + row * k * es,
                   static_cast<const char*>(w->hdist.p) + (row * k + cnt) * es);
    row += size_t(r->m);
  }
  return FUEL_AMD_OK;
}

void finish(fuel_amd_server* s, Request* r, int32_t rc, const std::string& err, int64_t bq, Clock::time_point t) {
  r->rc = rc;
  r->err = err;
  r->t_done = t;
  r->batch_queries = bq;
  r->done = true;
  s->n_req += 1;
  s->n_q += r->m;
  if (rc) s->n_failed += 1;
  const double ms = std::chrono::duration<double, std::milli>(t - r->t_submit).count();
  if (s->lat_ms.size() < fuel_amd_server::LAT_WINDOW) {
    s->lat_ms.push_back(ms);
  } else {
    s->lat_ms[s->lat_next] = ms;
    s->lat_next = (s->lat_next + 1) % fuel_amd_server::LAT_WINDOW;
  }
  s->lat_max = std::max(s->lat_max, ms);
  s->lat_sum += ms;
}

void worker_main(fuel_amd_server* s, Worker* w) {
  (void)hipSetDevice(s->ix->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
    if (s->queue.empty()) {
      if (s->stop) break;
      continue;
    }
    std::vector<Request*> batch;
    int64_t bm = 0;
    const int k = s->queue.front()->k;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
      Request* r = *it;
      if (r->k != k) {
        ++it;
        continue;
      }
      const int64_t nm = bm + r->m;
      if (nm > s->max_batch || s->reserved + batch_bytes(s, nm, k) > s->budget) break;
      batch.push_back(r);
      bm = nm;
      it = s->queue.erase(it);
    }
    if (batch.empty()) {
      if (s->inflight == 0) {
        // the head does not fit the budget even alone
        Request* r = s->queue.front();
        s->queue.pop_front();
        finish(s, r, FUEL_AMD_ERR_NOMEM, "request exceeds the server's HBM budget", 0, Clock::now());
        s->cv_done.notify_all();
      } else {
        s->cv_work.wait(lk);  // a finishing batch releases its reservation
      }
      continue;
    }
    const int64_t resv = batch_bytes(s, bm, k);
    s->reserved += resv;
    if (s->inflight++ == 0) s->busy_since = Clock::now();
    lk.unlock();
    const int32_t rc = run_batch(s, w, batch, bm, k);
    const std::string err = rc ? std::string(fuel_amd_last_error()) : std::string();
    const Clock::time_point t = Clock::now();
    lk.lock();
    s->reserved -= resv;
    if (--s->inflight == 0) s->busy_s += std::chrono::duration<double>(t - s->busy_since).count();
    for (Request* r : batch) finish(s, r, rc, err, bm, t);
    s->n_batches += 1;
    s->max_bq = std::max(s->max_bq, bm);
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace
}  // namespace fa

using namespace fa;

extern "C" {

int32_t fuel_amd_server_create(fuel_amd_index* ix, const fuel_amd_server_opts* opts, fuel_amd_server** out) {
  if (!ix || !out) return fail(FUEL_AMD_ERR_INVALID, "index and out must not be NULL");
  *out = nullptr;
  fuel_amd_server_opts o{};
  if (opts) o = *opts;
  if (o.workers < 0 || o.workers > 8) return fail(FUEL_AMD_ERR_INVALID, "workers must be in 0..8");
  if (o.max_batch < 0 || o.hbm_budget < 0) return fail(FUEL_AMD_ERR_INVALID, "negative max_batch or hbm_budget");
  {
    std::lock_guard<std::mutex> lock(ix->mu);
    int32_t rc = check_device(ix->device);
    if (rc) return rc;
    rc = prepare_mode(ix, ix->stream);
    if (rc) return rc;
  }
  auto s = std::make_unique<fuel_amd_server>();
  s->ix = ix;
  s->es = ix->dtype == FUEL_AMD_F32 ? 4 : 8;
  if (o.max_batch > 0) s->max_batch = o.max_batch;
  if (o.hbm_budget > 0) s->budget = o.hbm_budget;
  const int nw = o.workers > 0 ? o.workers : 2;
  for (int i = 0; i < nw; ++i) {
    auto w = std::make_unique<Worker>();
    w->ctx.publish = false;  // served calls leave the index's last-call record alone
    hipError_t e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      fuel_amd_server_destroy(s.release());
      return hip_fail(e, "worker stream");
    }
    s->workers.push_back(std::move(w));
  }
  for (auto& w : s->workers) w->th = std::thread(worker_main, s.get(), w.get());
  *out = s.release();
  return FUEL_AMD_OK;
}

int32_t fuel_amd_server_submit(fuel_amd_server* s, const void* query, int64_t m, int32_t k, uint64_t* ticket) {
  if (!s || !ticket) return fail(FUEL_AMD_ERR_INVALID, "server and ticket must not be NULL");
  if (m < 0 || k < 0) return fail(FUEL_AMD_ERR_INVALID, "m and k must be non-negative");
  if (m > s->max_batch) return fail(FUEL_AMD_ERR_INVALID, "m exceeds the server's max_batch");
  const size_t bytes = size_t(m) * s->ix->d * s->es;
  if (bytes > 0 && !query) return fail(FUEL_AMD_ERR_INVALID, "query must not be NULL");
  auto* r = new (std::nothrow) Request();
  if (!r) return fail(FUEL_AMD_ERR_NOMEM, "host allocation failed");
  r->m = m;
  r->k = k;
  if (bytes) r->q.assign(static_cast<const char*>(query), static_cast<const char*>(query) + bytes);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      delete r;
      return fail(FUEL_AMD_ERR_INVALID, "server is shutting down");
    }
    r->id = s->next_id++;
    s->live[r->id] = r;
    if (m == 0 || k == 0) {
      r->done = true;
      r->t_done = r->t_submit;
    } else {
      s->queue.push_back(r);
    }
    *ticket = r->id;
  }
  s->cv_work.notify_one();
  return FUEL_AMD_OK;
}

int32_t fuel_amd_server_wait(fuel_amd_server* s, uint64_t ticket, int64_t* out_idx, void* out_dist, double* latency_ms,
                             int64_t* batch_queries) {
  if (!s) return fail(FUEL_AMD_ERR_INVALID, "server must not be NULL");
  Request* r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return fail(FUEL_AMD_ERR_INVALID, "unknown ticket");
    r = it->second;
    // one wait per ticket: a second caller would race the first for the request
    if (r->claimed) return fail(FUEL_AMD_ERR_INVALID, "ticket is already being waited for");
    r->claimed = true;
    s->cv_done.wait(lk, [&] { return r->done; });
    // by key: submits during the wait may have rehashed the map
    s->live.erase(ticket);
  }
  std::unique_ptr<Request> hold(r);
  if (r->rc) return fail(r->rc, r->err);
  const size_t cnt = size_t(r->m) * r->k;
  if (cnt && (!out_idx || !out_dist)) return fail(FUEL_AMD_ERR_INVALID, "NULL output buffer");
  if (cnt) {
    std::memcpy(out_idx, r->idx.data(), cnt * 8);
    std::memcpy(out_dist, r->dist.data(), cnt * s->es);
  }
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (batch_queries) *batch_queries = r->batch_queries;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_server_stats(fuel_amd_server* s, int32_t reset, fuel_amd_serve_stats* out) {
  if (!s || !out) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  fuel_amd_serve_stats st{};
  st.requests = s->n_req;
  st.queries = s->n_q;
  st.batches = s->n_batches;
  st.max_batch_queries = s->max_bq;
  st.failed = s->n_failed;
  st.busy_s = s->busy_s;
  if (s->inflight > 0) st.busy_s += std::chrono::duration<double>(Clock::now() - s->busy_since).count();
  if (!s->lat_ms.empty()) {
    std::vector<double> v = s->lat_ms;
    const size_t n = v.size();
    auto pct = [&](double p) {
      // nearest rank
      size_t r = (size_t)std::ceil(p * double(n));
      r = std::min(std::max<size_t>(r, 1), n) - 1;
      std::nth_element(v.begin(), v.begin() + r, v.end());
      return v[r];
    };
    st.p50_ms = pct(0.50);
    st.p99_ms = pct(0.99);
    st.max_ms = s->lat_max;
    st.mean_ms = s->n_req > 0 ? s->lat_sum / double(s->n_req) : 0.0;
  }
  *out = st;
  if (reset) {
    s->lat_ms.clear();
    s->lat_next = 0;
    s->lat_max = s->lat_sum = 0.0;
    s->n_req = s->n_q = s->n_batches = s->max_bq = s->n_failed = 0;
    s->busy_s = 0.0;
    if (s->inflight > 0) s->busy_since = Clock::now();
  }
  return FUEL_AMD_OK;
}

int32_t fuel_amd_server_destroy(fuel_amd_server* s) {
  if (!s) return FUEL_AMD_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& w : s->workers)
    if (w->th.joinable()) w->th.join();
  (void)hipSetDevice(s->ix->device);
  for (auto& w : s->workers) {
    if (w->stream) {
      (void)hipStreamSynchronize(w->stream);
      (void)hipStreamDestroy(w->stream);
    }
    w->ctx.release();
  }
  for (auto& kv : s->live) delete kv.second;
  delete s;
  return FUEL_AMD_OK;
}

}  // extern "C"
