// fuel's LinearScanPrioritySearcher (src/search/linear_scan.rs:70-235) for m
// queries of one index.  Slot q is the searcher of query row q.  Its state is
// the reference's: cutoff (initially +inf), skip (initially 0), an
// initialized flag, a position, and the candidate list -- every row with
// skip <= d <= cutoff, sorted by (distance, index) -- which the reference builds
// on the first next() (ensure_initialized, :105-130).
//
// The candidate lists come from the GPU range kernel with per-query limits
// (one launch for a group of slots).  When a slot is materialized, the next
// slots of the handle that hold no list yet are materialized with it (up to
// GROUP): a list built with the limits of that moment stays valid as long as
// the skip is unchanged and the cutoff has not grown, since next() and
// all_lower_bound() stop at the first candidate beyond the current cutoff.
// The reference's lazy state stays observable as before: a slot counts as
// initialized only from its first next(), so all_lower_bound() of a slot that
// was never advanced returns 0, as in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

struct fuel_amd_priority {
  fuel_amd_index* ix = nullptr;
  int64_t m = 0;
  size_t es = 4;
  fa::DevBuf Q;                    // m x ld query rows
  fa::DevBuf stage, lim;           // a group's gathered rows and their limits
  struct Slot {
    double cutoff = std::numeric_limits<double>::infinity();
    double skip = 0.0;
    bool initialized = false;      // the reference's flag
    bool have = false;             // a materialized list is held
    double list_cutoff = 0.0, list_skip = 0.0;  // the limits it was built with
    int64_t position = 0;
    std::vector<int64_t> idx;
    std::vector<double> dist;      // values of the index dtype
  };
  std::vector<Slot> slots;
  std::mutex mu;
};

namespace fa {
namespace {

constexpr int64_t GROUP = 64;

double as_dtype(const fuel_amd_priority* p, double v) { return p->es == 4 ? (double)(float)v : v; }

bool list_valid(const fuel_amd_priority::Slot& s) {
  return s.have && s.list_skip == s.skip && s.list_cutoff >= s.cutoff;
}

// Materialize slot q and the next slots without a valid list (up to GROUP).
int32_t materialize(fuel_amd_priority* p, int64_t q) {
  fuel_amd_index* ix = p->ix;
  std::vector<int64_t> group;
  for (int64_t i = q; i < p->m && (int64_t)group.size() < GROUP; ++i)
    if (i == q || !p->slots[i].have) group.push_back(i);
  const int64_t g = (int64_t)group.size();
  const size_t row = size_t(ix->ld) * p->es;
  std::lock_guard<std::mutex> lock(ix->mu);
  int32_t rc = check_device(ix->device);
  if (rc) return rc;
  hipStream_t st = ix->stream;
  hipError_t e = p->stage.ensure(row * g);
  if (e == hipSuccess) e = p->lim.ensure(2 * p->es * g);
  std::vector<char> lim(2 * p->es * g);
  for (int64_t j = 0; j < g && e == hipSuccess; ++j) {
    const auto& s = p->slots[group[j]];
    if (p->es == 4) {
      const float c = (float)s.cutoff, k = (float)s.skip;
      std::memcpy(lim.data() + 4 * j, &c, 4);
      std::memcpy(lim.data() + 4 * (g + j), &k, 4);
    } else {
      std::memcpy(lim.data() + 8 * j, &s.cutoff, 8);
      std::memcpy(lim.data() + 8 * (g + j), &s.skip, 8);
    }
    e = hipMemcpyAsync(p->stage.as<char>() + row * j, p->Q.as<char>() + row * group[j], row, hipMemcpyDeviceToDevice,
                       st);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(p->lim.p, lim.data(), lim.size(), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "priority searcher staging");
  fuel_amd_csr csr{};
  rc = range_window_run(ix, p->stage.p, g, p->lim.p, p->lim.as<char>() + p->es * g, &csr);
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // limits were copied from the stack
  if (rc) return rc;
  if (e != hipSuccess) {
    fuel_amd_csr_free(&csr);
    return hip_fail(e, "priority searcher");
  }
  for (int64_t j = 0; j < g; ++j) {
    auto& s = p->slots[group[j]];
    const int64_t a = csr.offsets[j], b = csr.offsets[j + 1];
    s.idx.assign(csr.index + a, csr.index + b);
    s.dist.resize(size_t(b - a));
    for (int64_t t = a; t < b; ++t)
      s.dist[t - a] = p->es == 4 ? (double)static_cast<const float*>(csr.dist)[t] : static_cast<const double*>(csr.dist)[t];
    s.have = true;
    s.list_cutoff = s.cutoff;
    s.list_skip = s.skip;
  }
  fuel_amd_csr_free(&csr);
  return FUEL_AMD_OK;
}

// ensure_initialized (src/search/linear_scan.rs:105-130)
int32_t ensure_initialized(fuel_amd_priority* p, int64_t q) {
  auto& s = p->slots[q];
  if (s.initialized) return FUEL_AMD_OK;
  if (!list_valid(s)) {
    int32_t rc = materialize(p, q);
    if (rc) return rc;
  }
  s.position = 0;
  s.initialized = true;
  return FUEL_AMD_OK;
}

void write_dist(const fuel_amd_priority* p, double v, void* out) {
  if (!out) return;
  if (p->es == 4) {
    const float f = (float)v;
    std::memcpy(out, &f, 4);
  } else {
    std::memcpy(out, &v, 8);
  }
}

int32_t check_slot(fuel_amd_priority* p, int64_t q) {
  if (!p) return fail(FUEL_AMD_ERR_INVALID, "searcher must not be NULL");
  if (q < 0 || q >= p->m) return fail(FUEL_AMD_ERR_INVALID, "query slot out of range");
  return FUEL_AMD_OK;
}

}  // namespace
}  // namespace fa

using namespace fa;

extern "C" {

int32_t fuel_amd_priority_create(fuel_amd_index* ix, const void* query, int64_t m, fuel_amd_priority** out) {
  if (!ix || !out) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (m < 0 || (m > 0 && !query && ix->d > 0)) return fail(FUEL_AMD_ERR_INVALID, "bad query rows");
  auto* p = new (std::nothrow) fuel_amd_priority();
  if (!p) return fail(FUEL_AMD_ERR_NOMEM, "host allocation failed");
  p->ix = ix;
  p->m = m;
  p->es = ix->dtype == FUEL_AMD_F32 ? 4 : 8;
  p->slots.resize(size_t(m));
  {
    std::lock_guard<std::mutex> lock(ix->mu);
    int32_t rc = check_device(ix->device);
    hipError_t e = rc ? hipSuccess : p->Q.ensure(size_t(std::max<int64_t>(m, 1)) * ix->ld * p->es);
    if (!rc && e == hipSuccess) e = hipMemsetAsync(p->Q.p, 0, p->Q.bytes, ix->stream);
    if (!rc && e == hipSuccess && m > 0 && ix->d > 0) e = upload_rows(p->Q.p, ix->ld, query, m, ix->d, p->es, ix->upload, ix->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(ix->stream);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "query upload");
    if (rc) {
      delete p;
      return rc;
    }
  }
  *out = p;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_priority_destroy(fuel_amd_priority* p) {
  if (!p) return FUEL_AMD_OK;
  (void)hipSetDevice(p->ix->device);
  delete p;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_priority_reset(fuel_amd_priority* p, int64_t q, double cutoff, double skip) {
  int32_t rc = check_slot(p, q);
  if (rc) return rc;
  const double c = as_dtype(p, cutoff), k = as_dtype(p, skip);
  // debug_assert!(skip >= 0) and debug_assert!(cutoff >= skip) (:206-207)
  if (!(k >= 0.0) || !(c >= k)) return fail(FUEL_AMD_ERR_INVALID, "need 0 <= skip <= cutoff");
  std::lock_guard<std::mutex> lock(p->mu);
  auto& s = p->slots[q];
  s.position = 0;
  s.initialized = false;
  s.cutoff = c;
  s.skip = k;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_priority_decrease_cutoff(fuel_amd_priority* p, int64_t q, double threshold) {
  int32_t rc = check_slot(p, q);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(p->mu);
  auto& s = p->slots[q];
  const double t = as_dtype(p, threshold);
  // debug_assert!(threshold <= self.cutoff) (:193)
  if (!(t <= s.cutoff)) return fail(FUEL_AMD_ERR_INVALID, "the cutoff may only decrease");
  s.cutoff = t;
  return FUEL_AMD_OK;
}

int32_t fuel_amd_priority_next(fuel_amd_priority* p, int64_t q, int32_t (*skip_point)(void* ctx, int64_t index),
                               void* ctx, int64_t* out_index, void* out_dist) {
  int32_t rc = check_slot(p, q);
  if (rc) return rc;
  if (!out_index) return fail(FUEL_AMD_ERR_INVALID, "out_index must not be NULL");
  std::lock_guard<std::mutex> lock(p->mu);
  rc = ensure_initialized(p, q);
  if (rc) return rc;
  auto& s = p->slots[q];
  // find_next (:132-155)
  while (s.position < (int64_t)s.idx.size()) {
    const int64_t at = s.position++;
    const double d = s.dist[at];
    if (d > s.cutoff) break;
    if (d < s.skip) continue;
    if (skip_point && skip_point(ctx, s.idx[at])) continue;
    *out_index = s.idx[at];
    write_dist(p, d, out_dist);
    return FUEL_AMD_OK;
  }
  *out_index = -1;
  write_dist(p, std::numeric_limits<double>::infinity(), out_dist);
  return FUEL_AMD_OK;
}

int32_t fuel_amd_priority_all_lower_bound(fuel_amd_priority* p, int64_t q, void* out) {
  int32_t rc = check_slot(p, q);
  if (rc) return rc;
  if (!out) return fail(FUEL_AMD_ERR_INVALID, "out must not be NULL");
  std::lock_guard<std::mutex> lock(p->mu);
  const auto& s = p->slots[q];
  // all_lower_bound_internal (:157-175)
  double v = 0.0;
  if (s.initialized) {
    v = std::numeric_limits<double>::infinity();
    for (int64_t i = s.position; i < (int64_t)s.idx.size(); ++i) {
      const double d = s.dist[i];
      if (d > s.cutoff) break;
      if (d >= s.skip) {
        v = d;
        break;
      }
    }
  }
  write_dist(p, v, out);
  return FUEL_AMD_OK;
}

}  // extern "C"
