// Density clustering over the GPU range search: DBSCAN, parallel DBSCAN and
// OPTICS.  One GPU call computes every point's eps-neighbourhood (exact range on
// the restated distance, self included, sorted by (distance, index)); the
// label / ordering walks then run on the host in the reference's order.
//
//  dbscan           src/cluster/dbscan.rs:14-83 -- seeds in point order, each
//                   cluster expanded completely before the next seed; a border
//                   point keeps the first cluster that reaches it.  The order in
//                   which one expansion visits its frontier does not change the
//                   labels, so a stack replaces the reference's hash set.
//  parallel_dbscan  src/cluster/parallel_dbscan.rs:16-113 -- cores (>= min_points
//                   neighbours), connected components of neighbouring cores
//                   numbered by their smallest core index, a border point takes
//                   the label of the first core in its sorted neighbour list.
//  optics           src/cluster/optics.rs:56-133, :306-377 -- seeds popped by
//                   (reachability, index), stale seeds skipped, core distance =
//                   the min_points-th smallest neighbour distance, DBSCAN-style
//                   label extraction at eps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <queue>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

namespace fa {

namespace {

constexpr int64_t NOISE = -1;
constexpr int64_t UNVISITED = -2;

struct CsrGuard {
  fuel_amd_csr c{};
  ~CsrGuard() { fuel_amd_csr_free(&c); }
};

void dbscan_labels(const fuel_amd_csr& c, int64_t n, int64_t min_points, int64_t* labels) {
  const int64_t* off = c.offsets;
  const int64_t* idx = c.index;
  std::fill(labels, labels + n, UNVISITED);
  std::vector<int64_t> stack;
  int64_t cluster = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (labels[p] != UNVISITED) continue;
    if (off[p + 1] - off[p] < min_points) {
      labels[p] = NOISE;
      continue;
    }
    labels[p] = cluster;
    stack.assign(idx + off[p], idx + off[p + 1]);
    while (!stack.empty()) {
      const int64_t q = stack.back();
      stack.pop_back();
      if (labels[q] == NOISE) labels[q] = cluster;  // border point
      if (labels[q] != UNVISITED) continue;
      labels[q] = cluster;
      if (off[q + 1] - off[q] >= min_points) {
        for (int64_t j = off[q]; j < off[q + 1]; ++j) {
          const int64_t r = idx[j];
          if (labels[r] == NOISE) labels[r] = cluster;
          if (labels[r] == UNVISITED) stack.push_back(r);
        }
      }
    }
    ++cluster;
  }
}

struct UnionFind {
  std::vector<int64_t> parent;
  explicit UnionFind(int64_t n) : parent(n) {
    for (int64_t i = 0; i < n; ++i) parent[i] = i;
  }
  int64_t find(int64_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  }
  void unite(int64_t a, int64_t b) {
    a = find(a);
    b = find(b);
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
};

void parallel_dbscan_labels(const fuel_amd_csr& c, int64_t n, int64_t min_points, int64_t* labels) {
  const int64_t* off = c.offsets;
  const int64_t* idx = c.index;
  std::vector<int64_t> core_id(n, -1);
  int64_t ncores = 0;
  for (int64_t p = 0; p < n; ++p)
    if (off[p + 1] - off[p] >= min_points) core_id[p] = ncores++;
  if (ncores == 0) {
    std::fill(labels, labels + n, NOISE);
    return;
  }
  UnionFind uf(ncores);
  for (int64_t p = 0; p < n; ++p) {
    if (core_id[p] < 0) continue;
    for (int64_t j = off[p]; j < off[p + 1]; ++j)
      if (core_id[idx[j]] >= 0) uf.unite(core_id[p], core_id[idx[j]]);
  }
  // components numbered by their first core in core order
  std::vector<int64_t> root_label(ncores, -1), label_of_core(ncores);
  int64_t next = 0;
  for (int64_t k = 0; k < ncores; ++k) {
    const int64_t r = uf.find(k);
    if (root_label[r] < 0) root_label[r] = next++;
    label_of_core[k] = root_label[r];
  }
  for (int64_t p = 0; p < n; ++p) {
    if (core_id[p] >= 0) {
      labels[p] = label_of_core[core_id[p]];
      continue;
    }
    int64_t assigned = NOISE;
    for (int64_t j = off[p]; j < off[p + 1] && assigned == NOISE; ++j)
      if (core_id[idx[j]] >= 0) assigned = label_of_core[core_id[idx[j]]];
    labels[p] = assigned;
  }
}

template <typename F>
void optics_walk(const fuel_amd_csr& c, int64_t n, int64_t min_points, F eps, int64_t* ordering, F* reach,
                 F* core, int64_t* pred, int64_t* labels) {
  const int64_t* off = c.offsets;
  const int64_t* idx = c.index;
  const F* dist = reinterpret_cast<const F*>(c.dist);
  const F inf = std::numeric_limits<F>::infinity();
  std::vector<char> processed(n, 0), has_core(n, 0);
  std::fill(reach, reach + n, inf);
  std::fill(pred, pred + n, int64_t(-1));
  for (int64_t i = 0; i < n; ++i) core[i] = std::numeric_limits<F>::quiet_NaN();
  // min-heap on (reachability, index)
  struct Cand {
    F r;
    int64_t i;
  };
  auto later = [](const Cand& a, const Cand& b) { return a.r > b.r || (a.r == b.r && a.i > b.i); };
  std::priority_queue<Cand, std::vector<Cand>, decltype(later)> seeds(later);
  int64_t nord = 0;
  // neighbour lists are sorted by distance: the min_points-th smallest is entry min_points-1
  auto core_of = [&](int64_t p, F& cd) {
    if (off[p + 1] - off[p] < min_points) return false;
    cd = dist[off[p] + min_points - 1];
    return true;
  };
  auto update = [&](int64_t p, F cd) {
    for (int64_t j = off[p]; j < off[p + 1]; ++j) {
      const int64_t q = idx[j];
      if (q == p || processed[q]) continue;
      const F nr = std::max(cd, dist[j]);
      if (nr < reach[q]) {
        reach[q] = nr;
        pred[q] = p;
        seeds.push(Cand{nr, q});
      }
    }
  };
  for (int64_t p = 0; p < n; ++p) {
    if (processed[p]) continue;
    processed[p] = 1;
    ordering[nord++] = p;
    F cd;
    if (!core_of(p, cd)) continue;
    core[p] = cd;
    has_core[p] = 1;
    while (!seeds.empty()) seeds.pop();
    update(p, cd);
    while (!seeds.empty()) {
      const Cand s = seeds.top();
      seeds.pop();
      if (processed[s.i]) continue;
      if (s.r > reach[s.i]) continue;  // stale entry
      processed[s.i] = 1;
      ordering[nord++] = s.i;
      F cq;
      if (core_of(s.i, cq)) {
        core[s.i] = cq;
        has_core[s.i] = 1;
        update(s.i, cq);
      }
    }
  }
  // DBSCAN-style extraction at eps (optics.rs:345-377)
  int64_t cluster = -1;
  for (int64_t o = 0; o < n; ++o) {
    const int64_t p = ordering[o];
    const F cd = has_core[p] ? core[p] : inf;
    if (reach[p] > eps)
      labels[p] = cd <= eps ? ++cluster : NOISE;
    else
      labels[p] = cluster >= 0 ? cluster : NOISE;
  }
}

}  // namespace

int32_t dbscan_run(fuel_amd_index* ix, double eps, int64_t min_points, int variant, int64_t* labels) {
  if (ix->n == 0) return FUEL_AMD_OK;
  CsrGuard g;
  int32_t rc = range_self_run(ix, eps, &g.c);
  if (rc) return rc;
  if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
  if (variant == 0) {
    dbscan_labels(g.c, ix->n, min_points, labels);
  } else {
    parallel_dbscan_labels(g.c, ix->n, min_points, labels);
  }
  return FUEL_AMD_OK;
}

int32_t optics_run(fuel_amd_index* ix, double max_eps, int64_t min_points, int64_t* ordering, void* reach,
                   void* core, int64_t* pred, int64_t* labels) {
  if (ix->n == 0) return FUEL_AMD_OK;
  CsrGuard g;
  int32_t rc = range_self_run(ix, max_eps, &g.c);
  if (rc) return rc;
  if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
  if (ix->dtype == FUEL_AMD_F32)
    optics_walk<float>(g.c, ix->n, min_points, (float)max_eps, ordering, (float*)reach, (float*)core, pred, labels);
  else
    optics_walk<double>(g.c, ix->n, min_points, max_eps, ordering, (double*)reach, (double*)core, pred, labels);
  return FUEL_AMD_OK;
}

}  // namespace fa
