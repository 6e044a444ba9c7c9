// C-ABI host layer: index lifetime, argument checking, error/cancel conventions,
// query batching, workspace reuse and kernel orchestration.  Declarations and
// the reference interface each entry point replaces are in include/fuel_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fuel_amd.h"
#include "api_internal.h"
#include "kernels.h"

using namespace fa;

namespace fa {

thread_local std::string g_last_error;
std::atomic<bool> g_cancel{false};

int32_t fail(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int32_t hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory) return fail(FUEL_AMD_ERR_NOMEM, m);
  return fail(FUEL_AMD_ERR_DEVICE, m);
}

DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
  if (p && owned) hipFree(p);
  p = nullptr;
  bytes = 0;
  owned = true;
}
hipError_t DevBuf::ensure(size_t want) {
  if (want <= bytes && p) return hipSuccess;
  release();
  size_t b = std::max<size_t>(want, 256);
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return e;
  }
  bytes = b;
  return hipSuccess;
}

int32_t check_device(int dev) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return fail(FUEL_AMD_ERR_DEVICE, "no HIP device available (fuel_amd runs on the GPU only)");
  }
  if (dev < 0 || dev >= count) return fail(FUEL_AMD_ERR_INVALID, "device ordinal out of range");
  e = hipSetDevice(dev);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return FUEL_AMD_OK;
}

int device_cu_count(int dev) {
  static std::atomic<int> cached[64];  // zero-initialised; callers on several threads
  if (dev >= 0 && dev < 64) {
    const int c = cached[dev].load(std::memory_order_relaxed);
    if (c) return c;
  }
  int cu = 256;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    cu = 256;
  }
  if (dev >= 0 && dev < 64) cached[dev].store(cu, std::memory_order_relaxed);
  return cu;
}

bool cancelled() { return g_cancel.load(std::memory_order_relaxed); }

hipError_t upload_rows(void* dst, int64_t ld, const void* host, int64_t rows, int d, size_t es, DevBuf& stage,
                       hipStream_t st) {
  const size_t row_bytes = size_t(d) * es;
  if (rows <= 0 || d <= 0) return hipSuccess;
  if (ld == d) return hipMemcpyAsync(dst, host, size_t(rows) * row_bytes, hipMemcpyHostToDevice, st);
  hipError_t e = stage.ensure(size_t(rows) * row_bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(stage.p, host, size_t(rows) * row_bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(dst, size_t(ld) * es, stage.p, row_bytes, row_bytes, (size_t)rows, hipMemcpyDeviceToDevice,
                         st);
  return e;
}

// Choose the DB split count so that (query blocks x splits) fills the GPU.
int choose_splits(int64_t qblocks, int64_t n, int cus, int64_t min_rows) {
  int64_t target = int64_t(cus) * 8;
  int64_t s = (target + qblocks - 1) / std::max<int64_t>(1, qblocks);
  int64_t maxs = std::max<int64_t>(1, n / std::max<int64_t>(1, min_rows));
  s = std::max<int64_t>(1, std::min<int64_t>({s, maxs, 1024}));
  return (int)s;
}

void timing_mark(fuel_amd_index* ix, hipStream_t st, bool begin, const char* name) {
  if (!ix->timing) return;
  if (ix->ev_used >= ix->ev_pool.size()) return;  // pool exhausted: stop recording
  if (begin && (ix->ev_used & 1)) return;
  if (!begin && !(ix->ev_used & 1)) return;
  if (hipEventRecord(ix->ev_pool[ix->ev_used], st) == hipSuccess) {
    ++ix->ev_used;
    if (begin) ix->timed_kernel = name;
  } else {
    (void)hipGetLastError();
  }
}

}  // namespace fa

namespace fa {
template <typename F>
int32_t kmeans_pairwise_device(const F* p1, int64_t m, int64_t ld1, const F* p2, int64_t n, int64_t ld2, int d,
                               F* out, int64_t ldo, hipStream_t st) {
  KmeansPairwiseArgs<F> a{};
  a.P1 = p1;
  a.ld1 = ld1;
  a.m = m;
  a.P2 = p2;
  a.ld2 = ld2;
  a.n = n;
  a.d = d;
  a.out = out;
  a.ldo = ldo;
  hipError_t e = launch_kmeans_pairwise<F>(a, st);
  if (e != hipSuccess) return hip_fail(e, "pairwise_sqdist launch");
  return FUEL_AMD_OK;
}

int32_t check_sqdist_args(int64_t m, int64_t n, int32_t d, int32_t dtype) {
  if (m < 0 || n < 0 || d < 0) return fail(FUEL_AMD_ERR_INVALID, "sizes must be non-negative");
  if (dtype != FUEL_AMD_F32 && dtype != FUEL_AMD_F64) return fail(FUEL_AMD_ERR_INVALID, "dtype must be F32 or F64");
  return FUEL_AMD_OK;
}
}  // namespace fa

// ---------------------------------------------------------------------------
extern "C" {

const char* fuel_amd_version(void) { return "fuel_amd 0.1.0 (gfx950)"; }

int32_t fuel_amd_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return count;
}

const char* fuel_amd_last_error(void) { return g_last_error.c_str(); }
void fuel_amd_request_cancel(void) { g_cancel.store(true); }
void fuel_amd_clear_cancel(void) { g_cancel.store(false); }

static int32_t index_create_common(const void* data, bool on_device, int64_t n, int32_t d, int64_t ld,
                                   int32_t dtype, int32_t metric, const fuel_amd_opts* opts,
                                   fuel_amd_index** out) {
  g_last_error.clear();
  if (!out) return fail(FUEL_AMD_ERR_INVALID, "out must not be NULL");
  *out = nullptr;
  if (n < 0 || d < 0) return fail(FUEL_AMD_ERR_INVALID, "n and d must be non-negative");
  if (n > 0xFFFFFFFFll - 1) return fail(FUEL_AMD_ERR_UNSUPPORTED, "more than 2^32-2 rows");
  if (dtype != FUEL_AMD_F32 && dtype != FUEL_AMD_F64) return fail(FUEL_AMD_ERR_INVALID, "dtype must be F32 or F64");
  if (metric != FUEL_AMD_EUCLIDEAN && metric != FUEL_AMD_SQEUCLIDEAN)
    return fail(FUEL_AMD_ERR_UNSUPPORTED,
                "only the 'euclidean' and 'sqeuclidean' distances run on the linear-scan engine");
  if (n > 0 && !data) return fail(FUEL_AMD_ERR_INVALID, "data must not be NULL");
  fuel_amd_opts o{};
  o.mode = FUEL_AMD_MODE_AUTO;
  o.widen = 0;
  if (opts) o = *opts;
  if (o.widen < 0) o.widen = 0;  // 0: the mode's default (4; 16 for FP8R)
  if (o.mode < FUEL_AMD_MODE_AUTO || o.mode > FUEL_AMD_MODE_FP8R) return fail(FUEL_AMD_ERR_INVALID, "bad mode");
  if (o.flags & ~(FUEL_AMD_FLAG_FALLBACK_ARITH | FUEL_AMD_FLAG_NO_SEED))
    return fail(FUEL_AMD_ERR_INVALID, "unknown option flags");
  int32_t rc = check_device(o.device);
  if (rc) return rc;
  auto* ix = new (std::nothrow) fuel_amd_index();
  if (!ix) return fail(FUEL_AMD_ERR_NOMEM, "host allocation failed");
  ix->device = o.device;
  ix->dtype = dtype;
  ix->metric = metric;
  ix->n = n;
  ix->d = d;
  ix->opts = o;
  ix->arith = (o.flags & FUEL_AMD_FLAG_FALLBACK_ARITH) ? fa::ARITH_FALLBACK : fa::ARITH_AVX;
  ix->ctx.timed = true;
  const size_t es = dtype == FUEL_AMD_F32 ? 4 : 8;
  // rows padded to a 32-byte multiple so the exact kernels can use 16-byte loads
  ix->ld = dtype == FUEL_AMD_F32 ? ((int64_t(d) + 7) & ~int64_t(7)) : ((int64_t(d) + 3) & ~int64_t(3));
  if (ix->ld == 0) ix->ld = dtype == FUEL_AMD_F32 ? 8 : 4;
  hipError_t e = hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ix;
    return hip_fail(e, "hipStreamCreate");
  }
  e = ix->X.ensure(size_t(std::max<int64_t>(n, 1)) * ix->ld * es);
  if (e == hipSuccess) e = hipMemsetAsync(ix->X.p, 0, size_t(std::max<int64_t>(n, 1)) * ix->ld * es, ix->stream);
  if (e == hipSuccess && n > 0 && d > 0) {
    if (on_device) {
      e = hipMemcpy2DAsync(ix->X.p, ix->ld * es, data, ld * es, size_t(d) * es, (size_t)n, hipMemcpyDeviceToDevice,
                           ix->stream);
    } else {
      fa::DevBuf stage;
      e = fa::upload_rows(ix->X.p, ix->ld, data, n, d, es, stage, ix->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ix->stream);  // before `stage` is released
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ix->stream);
  if (e != hipSuccess) {
    int32_t r = hip_fail(e, "index upload");
    fuel_amd_index_destroy(ix);
    return r;
  }
  *out = ix;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_index_create(const void* data, int64_t n, int32_t d, int32_t dtype, int32_t metric,
                              const fuel_amd_opts* opts, fuel_amd_index** out) {
  return index_create_common(data, false, n, d, d, dtype, metric, opts, out);
}

int32_t fuel_amd_index_create_device(const void* d_data, int64_t n, int32_t d, int64_t ld, int32_t dtype,
                                     int32_t metric, const fuel_amd_opts* opts, fuel_amd_index** out) {
  if (ld < d) return fail(FUEL_AMD_ERR_INVALID, "ld must be >= d");
  return index_create_common(d_data, true, n, d, ld, dtype, metric, opts, out);
}

int32_t fuel_amd_index_destroy(fuel_amd_index* ix) {
  if (!ix) return FUEL_AMD_OK;
  (void)hipSetDevice(ix->device);
  if (ix->stream) hipStreamSynchronize(ix->stream);
  ix->release_all();
  for (auto& ev : ix->ev_pool) hipEventDestroy(ev);
  ix->ev_pool.clear();
  if (ix->stream) hipStreamDestroy(ix->stream);
  delete ix;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_timing_enable(fuel_amd_index* ix, int32_t enable) {
  if (!ix) return fail(FUEL_AMD_ERR_INVALID, "index must not be NULL");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  if (enable && ix->ev_pool.empty()) {
    ix->ev_pool.resize(4096);
    for (auto& ev : ix->ev_pool) {
      hipError_t e = hipEventCreate(&ev);
      if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    }
  }
  ix->timing = enable != 0;
  ix->ev_used = 0;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_timing_read(fuel_amd_index* ix, double* kernel_ms, int64_t* launches, char* name,
                             int32_t name_len) {
  if (!ix || !kernel_ms || !launches) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  double total = 0.0;
  int64_t cnt = 0;
  for (size_t i = 0; i + 1 < ix->ev_used; i += 2) {
    hipError_t e = hipEventSynchronize(ix->ev_pool[i + 1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ix->ev_pool[i], ix->ev_pool[i + 1]);
    if (e != hipSuccess) return hip_fail(e, "event timing");
    total += ms;
    ++cnt;
  }
  *kernel_ms = total;
  *launches = cnt;
  if (name && name_len > 0) {
    std::strncpy(name, ix->timed_kernel.c_str(), (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  ix->ev_used = 0;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_last_stats(const fuel_amd_index* ix, fuel_amd_stats* out) {
  if (!ix || !out) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  auto* mix = const_cast<fuel_amd_index*>(ix);
  std::lock_guard<std::mutex> lock(mix->stats_mu);
  *out = ix->last_stats;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_last_fallback_rows(const fuel_amd_index* ix, int64_t* rows, int64_t cap, int64_t* count) {
  if (!ix || !count || (cap > 0 && !rows) || cap < 0) return fail(FUEL_AMD_ERR_INVALID, "bad argument");
  auto* mix = const_cast<fuel_amd_index*>(ix);
  std::lock_guard<std::mutex> lock(mix->stats_mu);
  const auto& v = ix->last_fail_rows;
  *count = (int64_t)v.size();
  for (int64_t i = 0; i < cap && i < (int64_t)v.size(); ++i) rows[i] = v[i];
  return FUEL_AMD_OK;
}

namespace {

// The context a kNN call runs on: the index's own (with its lock held) when
// free, otherwise an idle lane (made when none is idle), returned on release.
struct CallCtx {
  fuel_amd_index* ix = nullptr;
  std::unique_lock<std::mutex> own;
  fuel_amd_index::CallLane* lane = nullptr;
  fa::QueryCtx* ctx = nullptr;
  hipStream_t stream = nullptr;
  fa::DevBuf *upload = nullptr, *hq = nullptr, *hout_idx = nullptr, *hout_dist = nullptr;

  int32_t acquire(fuel_amd_index* index) {
    ix = index;
    own = std::unique_lock<std::mutex>(ix->mu, std::try_to_lock);
    if (own.owns_lock()) {
      ctx = &ix->ctx;
      stream = ix->stream;
      upload = &ix->upload;
      hq = &ix->hq;
      hout_idx = &ix->hout_idx;
      hout_dist = &ix->hout_dist;
      return FUEL_AMD_OK;
    }
    {
      std::lock_guard<std::mutex> lock(ix->lanes_mu);
      if (!ix->lanes_idle.empty()) {
        lane = ix->lanes_idle.back();
        ix->lanes_idle.pop_back();
      }
    }
    if (!lane) {
      auto L = std::make_unique<fuel_amd_index::CallLane>();
      hipError_t e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
      if (e != hipSuccess) return hip_fail(e, "call lane stream");
      lane = L.get();
      std::lock_guard<std::mutex> lock(ix->lanes_mu);
      ix->lanes.push_back(std::move(L));
    }
    ctx = &lane->ctx;
    stream = lane->stream;
    upload = &lane->upload;
    hq = &lane->hq;
    hout_idx = &lane->hout_idx;
    hout_dist = &lane->hout_dist;
    return FUEL_AMD_OK;
  }
  ~CallCtx() {
    if (lane) {
      std::lock_guard<std::mutex> lock(ix->lanes_mu);
      ix->lanes_idle.push_back(lane);
    }
  }
};

}  // namespace

// ---- kNN --------------------------------------------------------------------

int32_t fuel_amd_knn_device(fuel_amd_index* ix, const void* d_query, int64_t m, int64_t ldq, int32_t k,
                            int32_t exclude_self, int64_t* d_out_idx, void* d_out_dist, void* stream) {
  g_last_error.clear();
  if (!ix) return fail(FUEL_AMD_ERR_INVALID, "index must not be NULL");
  if (m < 0 || k < 0) return fail(FUEL_AMD_ERR_INVALID, "m and k must be non-negative");
  if (exclude_self < 0 || exclude_self > 2) return fail(FUEL_AMD_ERR_INVALID, "bad exclude_self");
  if (ldq < ix->d) return fail(FUEL_AMD_ERR_INVALID, "ldq must be >= d");
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  CallCtx cc;
  rc = cc.acquire(ix);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : cc.stream;
  return knn_run(ix, cc.ctx, d_query, m, ldq, k, exclude_self, d_out_idx, d_out_dist, st);
}

int32_t fuel_amd_knn(fuel_amd_index* ix, const void* query, int64_t m, int32_t k, int32_t exclude_self,
                     int64_t* out_idx, void* out_dist) {
  g_last_error.clear();
  if (!ix) return fail(FUEL_AMD_ERR_INVALID, "index must not be NULL");
  if (m < 0 || k < 0) return fail(FUEL_AMD_ERR_INVALID, "m and k must be non-negative");
  if (exclude_self < 0 || exclude_self > 2) return fail(FUEL_AMD_ERR_INVALID, "bad exclude_self");
  if (m == 0 || k == 0) return FUEL_AMD_OK;
  if (!query || !out_idx || !out_dist) return fail(FUEL_AMD_ERR_INVALID, "NULL buffer");
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  CallCtx cc;
  rc = cc.acquire(ix);
  if (rc) return rc;
  const size_t es = ix->dtype == FUEL_AMD_F32 ? 4 : 8;
  hipStream_t st = cc.stream;
  hipError_t e = cc.hq->ensure(size_t(m) * ix->ld * es);
  if (e == hipSuccess) e = cc.hout_idx->ensure(size_t(m) * k * 8);
  if (e == hipSuccess) e = cc.hout_dist->ensure(size_t(m) * k * es);
  if (e != hipSuccess) return hip_fail(e, "query buffers");
  if (ix->d > 0) {
    e = fa::upload_rows(cc.hq->p, ix->ld, query, m, ix->d, es, *cc.upload, st);
    if (e != hipSuccess) return hip_fail(e, "query upload");
  }
  rc = knn_run(ix, cc.ctx, cc.hq->p, m, ix->ld, k, exclude_self, (int64_t*)cc.hout_idx->p, cc.hout_dist->p, st);
  if (rc) return rc;
  e = hipMemcpyAsync(out_idx, cc.hout_idx->p, size_t(m) * k * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out_dist, cc.hout_dist->p, size_t(m) * k * es, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "result download");
  return FUEL_AMD_OK;
}

// ---- range ------------------------------------------------------------------

void fuel_amd_csr_free(fuel_amd_csr* c) {
  if (!c) return;
  std::free(c->offsets);
  std::free(c->index);
  std::free(c->dist);
  c->offsets = nullptr;
  c->index = nullptr;
  c->dist = nullptr;
  c->nnz = 0;
  c->m = 0;
}

int32_t fuel_amd_range(fuel_amd_index* ix, const void* query, int64_t m, double radius, int32_t exclude_self,
                       fuel_amd_csr* out) {
  g_last_error.clear();
  if (!ix || !out) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  std::memset(out, 0, sizeof(*out));
  if (m < 0) return fail(FUEL_AMD_ERR_INVALID, "m must be non-negative");
  if (m > 0 && !query) return fail(FUEL_AMD_ERR_INVALID, "query must not be NULL");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  return range_run(ix, query, m, radius, exclude_self, out);
}

int32_t fuel_amd_knn_ties(fuel_amd_index* ix, const void* query, int64_t m, int32_t k, fuel_amd_csr* out) {
  g_last_error.clear();
  if (!ix || !out) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  std::memset(out, 0, sizeof(*out));
  if (m < 0 || k < 0) return fail(FUEL_AMD_ERR_INVALID, "m and k must be non-negative");
  if (m > 0 && !query) return fail(FUEL_AMD_ERR_INVALID, "query must not be NULL");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  return knn_ties_run(ix, query, m, k, out);
}

// ---- density clustering -------------------------------------------------------

int32_t fuel_amd_dbscan(fuel_amd_index* ix, double eps, int64_t min_points, int32_t variant, int64_t* out_labels) {
  g_last_error.clear();
  if (!ix || (!out_labels && ix->n > 0)) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  if (!(eps >= 0.0)) return fail(FUEL_AMD_ERR_INVALID, "eps must be non-negative");
  if (variant != 0 && variant != 1) return fail(FUEL_AMD_ERR_INVALID, "variant must be 0 (dbscan) or 1 (parallel)");
  if (min_points < 0) return fail(FUEL_AMD_ERR_INVALID, "min_points must be non-negative");
  if (variant == 1 && min_points == 0) return fail(FUEL_AMD_ERR_INVALID, "min_points must be greater than 0");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  return dbscan_run(ix, eps, min_points, variant, out_labels);
}

int32_t fuel_amd_optics(fuel_amd_index* ix, double max_eps, int64_t min_points, int64_t* out_ordering,
                        void* out_reachability, void* out_core_distance, int64_t* out_predecessor,
                        int64_t* out_labels) {
  g_last_error.clear();
  if (!ix) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  if (ix->n > 0 && (!out_ordering || !out_reachability || !out_core_distance || !out_predecessor || !out_labels))
    return fail(FUEL_AMD_ERR_INVALID, "NULL output");
  if (!(max_eps >= 0.0)) return fail(FUEL_AMD_ERR_INVALID, "eps must be non-negative");
  if (min_points <= 0) return fail(FUEL_AMD_ERR_INVALID, "min_points must be greater than 0");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  return optics_run(ix, max_eps, min_points, out_ordering, out_reachability, out_core_distance, out_predecessor,
                    out_labels);
}

// ---- k-means distance matrix (math::pairwise_sqdist) -----------------------


int32_t fuel_amd_pairwise_sqdist_device(const void* d_points1, int64_t nrows, int64_t ld1, const void* d_points2,
                                        int64_t ncols, int64_t ld2, int32_t d, int32_t dtype, void* d_out,
                                        int64_t ldo, void* stream) {
  g_last_error.clear();
  int32_t rc = check_sqdist_args(nrows, ncols, d, dtype);
  if (rc) return rc;
  if (nrows == 0 || ncols == 0) return FUEL_AMD_OK;
  if (!d_points1 || !d_points2 || !d_out) return fail(FUEL_AMD_ERR_INVALID, "NULL buffer");
  if (ld1 < d || ld2 < d || ldo < ncols) return fail(FUEL_AMD_ERR_INVALID, "row pitch smaller than the row");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) (void)hipGetLastError();
  rc = check_device(dev);  // the caller's current device
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return dtype == FUEL_AMD_F32
             ? kmeans_pairwise_device<float>((const float*)d_points1, nrows, ld1, (const float*)d_points2, ncols,
                                             ld2, d, (float*)d_out, ldo, st)
             : kmeans_pairwise_device<double>((const double*)d_points1, nrows, ld1, (const double*)d_points2, ncols,
                                              ld2, d, (double*)d_out, ldo, st);
}

int32_t fuel_amd_pairwise_sqdist(const void* points1, int64_t nrows, const void* points2, int64_t ncols, int32_t d,
                                 int32_t dtype, void* out) {
  g_last_error.clear();
  int32_t rc = check_sqdist_args(nrows, ncols, d, dtype);
  if (rc) return rc;
  if (nrows == 0 || ncols == 0) return FUEL_AMD_OK;
  if (!out || (d > 0 && (!points1 || !points2))) return fail(FUEL_AMD_ERR_INVALID, "NULL buffer");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) (void)hipGetLastError();
  rc = check_device(dev);
  if (rc) return rc;
  const size_t es = dtype == FUEL_AMD_F32 ? 4 : 8;
  const int64_t dd = d > 0 ? d : 1;
  DevBuf b1, b2, bo;
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  e = b1.ensure(size_t(nrows) * dd * es);
  if (e == hipSuccess) e = b2.ensure(size_t(ncols) * dd * es);
  if (e == hipSuccess) e = bo.ensure(size_t(nrows) * ncols * es);
  if (e == hipSuccess && d > 0) e = hipMemcpyAsync(b1.p, points1, size_t(nrows) * d * es, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && d > 0) e = hipMemcpyAsync(b2.p, points2, size_t(ncols) * d * es, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    hipStreamDestroy(st);
    return hip_fail(e, "pairwise_sqdist buffers");
  }
  rc = dtype == FUEL_AMD_F32
           ? kmeans_pairwise_device<float>(b1.as<float>(), nrows, dd, b2.as<float>(), ncols, dd, d, bo.as<float>(),
                                           ncols, st)
           : kmeans_pairwise_device<double>(b1.as<double>(), nrows, dd, b2.as<double>(), ncols, dd, d,
                                            bo.as<double>(), ncols, st);
  if (rc == FUEL_AMD_OK) {
    e = hipMemcpyAsync(out, bo.p, size_t(nrows) * ncols * es, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_fail(e, "pairwise_sqdist download");
  }
  hipStreamDestroy(st);
  return rc;
}

// ---- pairwise ---------------------------------------------------------------

int32_t fuel_amd_pairwise(const void* data, int64_t n, int32_t d, int32_t dtype, int32_t metric, int32_t form,
                          void* out) {
  g_last_error.clear();
  if (form != FUEL_AMD_FORM_SQUARE && form != FUEL_AMD_FORM_CONDENSED)
    return fail(FUEL_AMD_ERR_INVALID, "form must be SQUARE or CONDENSED");
  fuel_amd_index* ix = nullptr;
  int32_t rc = fuel_amd_index_create(data, n, d, dtype, metric, nullptr, &ix);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(ix->mu);
    rc = pairwise_run(ix, form, out);
  }
  fuel_amd_index_destroy(ix);
  return rc;
}

int32_t fuel_amd_index_pairwise(fuel_amd_index* ix, int32_t form, void* out) {
  g_last_error.clear();
  if (!ix || (!out && ix->n > 1)) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  if (form != FUEL_AMD_FORM_SQUARE && form != FUEL_AMD_FORM_CONDENSED)
    return fail(FUEL_AMD_ERR_INVALID, "form must be SQUARE or CONDENSED");
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  return pairwise_run(ix, form, out);
}

// ---- outlier scores -----------------------------------------------------------

int32_t fuel_amd_lof(fuel_amd_index* ix, int32_t k, void* out_scores) {
  g_last_error.clear();
  if (!ix || (!out_scores && ix->n > 0)) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  if (k <= 0) return fail(FUEL_AMD_ERR_INVALID, "k must be greater than 0");  // src/outlier/lof.rs:22
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  return outlier_run(ix, k, 0, out_scores);
}

int32_t fuel_amd_knn_outlier(fuel_amd_index* ix, int32_t k, void* out_scores) {
  g_last_error.clear();
  if (!ix || (!out_scores && ix->n > 0)) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  if (k <= 0) return fail(FUEL_AMD_ERR_INVALID, "k must be greater than 0");  // src/outlier/knn.rs:22
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  return outlier_run(ix, k, 1, out_scores);
}

}  // extern "C"
