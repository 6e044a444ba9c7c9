// ORACLE — test infrastructure only.
//
// A CPU restatement of fuel's exact linear-scan kNN / range / pairwise path, used
// as the checker by tests/, by __graft_entry__.smoke() and as the timed CPU
// baseline in bench.py.  Nothing in the product library (fuel_amd/) links, loads
// or calls this file.  Every function cites the reference source it restates
// (paths are relative to the fuel repository root).
//
// Build: oracle/Makefile  (g++ -O2 -mavx2 -mfma -ffp-contract=off, never fast-math)
//
// Pinning (how each part is checked, see tests/test_oracle.py):
//   sqdist_f32 / sqdist_f64 / ....... pinned bit for bit by an exact-rational
//   sqdist_fallback                   emulator of each IEEE operation
//                                     (tests/test_oracle_arith.py) on scale-mixed
//                                     inputs where unfused lanes, 4-vs-8 lanes, one
//                                     sequential accumulator, another lane fold or
//                                     another tail form each change the bits; also
//                                     the exact pairwise matrices of
//                                     src/math/mod.rs:672-690, an exact-integer
//                                     closed form and brute-force fp64 within the
//                                     rounding bound.
//   KNNHeap / linear_scan_knn ....... pinned: src/search/linear_scan.rs:252-265
//                                     (order [1,0,2] with a tie), the brute-force
//                                     semantics of src/search/vptree/tests/shared.rs:20-42
//                                     (sort + take_while <= kth), integer grids.
//   linear_scan_range ............... pinned: src/search/linear_scan.rs:268-279.
//   for_each_knn + LOF + kNN-outlier  pinned: ELKI LOF-10 / KNN-10 columns of
//                                     testdata/reference-outlier-scores.csv at 1e-6
//                                     (src/outlier/lof.rs:192-211, src/outlier/knn.rs:78-96)
//                                     and the closed form of src/outlier/lof.rs:119-138.
//   pairwise dense / condensed ...... pinned: symmetry, zero diagonal, the
//                                     condensed order i(i-1)/2+j of
//                                     src/api/condensed_distance_matrix.rs:94-101.
//   precomputed subset_with_ties .... pinned: src/search/precomputed.rs:136-165.
//   dbscan / parallel_dbscan ........ pinned: the sklearn toy labels of
//                                     src/cluster/dbscan.rs tests (eps 1, min_points
//                                     1-4) and the two-cluster case; the two variants
//                                     agree where border points are unambiguous.
//   optics .......................... pinned: the ELKI ordering and reachability of
//                                     src/cluster/optics.rs:587-743 (60 points,
//                                     eps = inf, min_points 5) and the sklearn
//                                     processing order of :554-585.
//   kmeans_pairwise ................. pinned: the exact matrices of
//                                     src/math/mod.rs:672-690 and an exact-rational
//                                     emulation of the sequential fused chain.
//   extract_xi ...................... pinned: the two-valley plot of
//                                     src/cluster/optics.rs:512-552.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <queue>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// a1: squared_euclidean_distance_f32_avx_fma, src/distance/squared_euclidean.rs:13-39
// One 8-lane accumulator interleaved over d, a sequential (left-fold) lane sum,
// then a fused multiply-add tail.  Rust's f32 `Sum` is a left fold whose start
// value is a signed zero; adding it to a non-negative lane value is exact.
float sqdist_f32_avx(const float* a, const float* b, int64_t d) {
  const int64_t sd = d & ~int64_t(7);
  __m256 acc = _mm256_setzero_ps();
  for (int64_t i = 0; i < sd; i += 8) {
    __m256 va = _mm256_loadu_ps(a + i);
    __m256 vb = _mm256_loadu_ps(b + i);
    __m256 diff = _mm256_sub_ps(va, vb);
    acc = _mm256_fmadd_ps(diff, diff, acc);
  }
  alignas(32) float tmp[8];
  _mm256_store_ps(tmp, acc);
  float sum = -0.0f;
  for (int j = 0; j < 8; ++j) sum = sum + tmp[j];
  for (int64_t i = sd; i < d; ++i) {
    float diff = a[i] - b[i];
    sum = std::fma(diff, diff, sum);
  }
  return sum;
}

// a2: squared_euclidean_distance_f64_avx_fma, src/distance/squared_euclidean.rs:41-67
double sqdist_f64_avx(const double* a, const double* b, int64_t d) {
  const int64_t sd = d & ~int64_t(3);
  __m256d acc = _mm256_setzero_pd();
  for (int64_t i = 0; i < sd; i += 4) {
    __m256d va = _mm256_loadu_pd(a + i);
    __m256d vb = _mm256_loadu_pd(b + i);
    __m256d diff = _mm256_sub_pd(va, vb);
    acc = _mm256_fmadd_pd(diff, diff, acc);
  }
  alignas(32) double tmp[4];
  _mm256_store_pd(tmp, acc);
  double sum = -0.0;
  for (int j = 0; j < 4; ++j) sum = sum + tmp[j];
  for (int64_t i = sd; i < d; ++i) {
    double diff = a[i] - b[i];
    sum = std::fma(diff, diff, sum);
  }
  return sum;
}

// Scalar restatement of a1 with explicit lanes (no intrinsics); used to cross-check
// the vector form above.  Same arithmetic, same order.
float sqdist_f32_lanes(const float* a, const float* b, int64_t d) {
  const int64_t sd = d & ~int64_t(7);
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = 0; i < sd; i += 8)
    for (int j = 0; j < 8; ++j) {
      float t = a[i + j] - b[i + j];
      acc[j] = std::fma(t, t, acc[j]);
    }
  float sum = -0.0f;
  for (int j = 0; j < 8; ++j) sum = sum + acc[j];
  for (int64_t i = sd; i < d; ++i) {
    float t = a[i] - b[i];
    sum = std::fma(t, t, sum);
  }
  return sum;
}

// a3: squared_euclidean_distance_fallback, src/distance/squared_euclidean.rs:99-133
// (taken only on hosts without AVX+FMA; kept for completeness).  Unfused.
template <typename N>
N sqdist_fallback(const N* a, const N* b, int64_t d) {
  const int64_t sd = d & ~int64_t(7);
  N vsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = 0; i < sd; i += 8)
    for (int j = 0; j < 8; ++j) {
      N diff = a[i + j] - b[i + j];
      N p = diff * diff;
      vsum[j] = p + vsum[j];
    }
  N s = 0;
  for (int j = 0; j < 8; ++j) s = s + vsum[j];
  for (int64_t i = sd; i < d; ++i) {
    N diff = a[i] - b[i];
    N p = diff * diff;
    s = p + s;
  }
  return s;
}

inline float sqdist(const float* a, const float* b, int64_t d) { return sqdist_f32_avx(a, b, d); }
inline double sqdist(const double* a, const double* b, int64_t d) { return sqdist_f64_avx(a, b, d); }

// a4/a5: Euclidean = sqrt of a1 (src/distance/euclidean.rs:7-9, IEEE correctly
// rounded); SquaredEuclidean = a1 (src/distance/squared_euclidean.rs:137-157).
// metric bit 0: 0 = euclidean / l2, 1 = sqeuclidean / squared_euclidean.
// metric bit 4 (ARITH_FALLBACK): the distance takes a3's rounding, the branch of
// squared_euclidean_distance (src/distance/squared_euclidean.rs:71-97) that fuel
// runs on hosts without AVX+FMA (every non-x86 host), instead of a1/a2.
constexpr int ARITH_FALLBACK = 16;
template <typename F>
inline F distance(const F* a, const F* b, int64_t d, int metric) {
  F s = (metric & ARITH_FALLBACK) ? sqdist_fallback<F>(a, b, d) : sqdist(a, b, d);
  return (metric & 1) == 0 ? std::sqrt(s) : s;
}

// ---------------------------------------------------------------------------
// a7: DistPair and its total order, src/api/search.rs:10-30.  Distance compared
// with partial_cmp (NaN -> Equal), then index.
template <typename F>
struct DistPair {
  F distance;
  int64_t index;
};

template <typename F>
inline int cmp_pair(const DistPair<F>& x, const DistPair<F>& y) {
  if (x.distance < y.distance) return -1;
  if (x.distance > y.distance) return 1;
  if (x.index < y.index) return -1;
  if (x.index > y.index) return 1;
  return 0;
}

template <typename F>
struct PairLess {
  bool operator()(const DistPair<F>& x, const DistPair<F>& y) const { return cmp_pair(x, y) < 0; }
};

// a8: KNNHeap, src/api/search.rs:80-151.  A max-heap of k pairs plus the list of
// pairs tied with the current k-th distance.
template <typename F>
struct KNNHeap {
  size_t k;
  std::vector<DistPair<F>> heap;  // std heap with PairLess: front() is the max
  std::vector<DistPair<F>> ties;
  F k_distance;

  explicit KNNHeap(size_t k_) : k(k_), k_distance(INFINITY) { heap.reserve(k_); }

  void insert(const DistPair<F>& pair) {
    if (k == 0) return;
    if (heap.size() < k) {
      heap.push_back(pair);
      std::push_heap(heap.begin(), heap.end(), PairLess<F>());
      if (heap.size() == k) k_distance = heap.front().distance;
      return;
    }
    // partial_cmp(...).unwrap_or(Equal)
    int c = (pair.distance < k_distance) ? -1 : (pair.distance > k_distance ? 1 : 0);
    if (c < 0) {
      std::pop_heap(heap.begin(), heap.end(), PairLess<F>());
      DistPair<F> removed = heap.back();
      heap.back() = pair;
      std::push_heap(heap.begin(), heap.end(), PairLess<F>());
      F new_k = heap.front().distance;
      if (new_k < k_distance) {
        ties.clear();
      } else if (removed.distance == new_k) {
        ties.push_back(removed);
      }
      k_distance = new_k;
    } else if (c == 0) {
      ties.push_back(pair);
    }
  }

  std::vector<DistPair<F>> into_vec() {
    std::vector<DistPair<F>> out = heap;
    out.insert(out.end(), ties.begin(), ties.end());
    std::sort(out.begin(), out.end(), PairLess<F>());
    return out;
  }
};

// a9: linear_scan_knn, src/api/search.rs:169-184 (points visited in index order).
template <typename F>
std::vector<DistPair<F>> linear_scan_knn(const F* X, int64_t n, int64_t d, const F* q, int64_t k,
                                         int metric) {
  if (k == 0 || n == 0) return {};
  KNNHeap<F> heap((size_t)k);
  for (int64_t i = 0; i < n; ++i) heap.insert({distance(q, X + i * d, d, metric), i});
  return heap.into_vec();
}

// a9: linear_scan_range, src/api/search.rs:154-166 (keep d <= radius, sort).
template <typename F>
std::vector<DistPair<F>> linear_scan_range(const F* X, int64_t n, int64_t d, const F* q, F radius,
                                           int metric) {
  std::vector<DistPair<F>> out;
  for (int64_t i = 0; i < n; ++i) {
    F dist = distance(q, X + i * d, d, metric);
    if (dist <= radius) out.push_back({dist, i});
  }
  std::sort(out.begin(), out.end(), PairLess<F>());
  return out;
}

// Dynamic parallel-for over [0, count), like rayon's par_try_map over query rows
// (src/api/parallel.rs:64-100); collect preserves order because each item writes
// its own slot.
template <typename Fn>
void parallel_for(int64_t count, int nthreads, Fn fn) {
  if (nthreads <= 1 || count <= 1) {
    for (int64_t i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next(0);
  const int64_t chunk = 16;
  std::vector<std::thread> pool;
  int nt = (int)std::min<int64_t>(nthreads, count);
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        int64_t s = next.fetch_add(chunk);
        if (s >= count) break;
        int64_t e = std::min(count, s + chunk);
        for (int64_t i = s; i < e; ++i) fn(i);
      }
    });
  for (auto& th : pool) th.join();
}

// Exclusion semantics for self-queries:
//   0: none
//   1: fuel's for_each_knn (src/outlier/common.rs:27, 41-44): search k+1, drop the
//      query's own index, truncate to k.
//   2: the Python linear-scan path (python/search.rs:485-488): search k, drop the
//      query's own index (no truncation needed; packing truncates to k).
template <typename F>
std::vector<DistPair<F>> knn_one(const F* X, int64_t n, int64_t d, const F* q, int64_t qi,
                                 int64_t k, int metric, int exclude) {
  if (exclude == 1) {
    auto r = linear_scan_knn(X, n, d, q, k + 1, metric);
    r.erase(std::remove_if(r.begin(), r.end(), [&](const DistPair<F>& p) { return p.index == qi; }),
            r.end());
    if ((int64_t)r.size() > k) r.resize(k);
    return r;
  }
  auto r = linear_scan_knn(X, n, d, q, k, metric);
  if (exclude == 2)
    r.erase(std::remove_if(r.begin(), r.end(), [&](const DistPair<F>& p) { return p.index == qi; }),
            r.end());
  return r;
}

// a12/a14: linear_scan_knn_batch + knn_results_to_arrays, python/search.rs:471-491 and
// :1195-1216.  Rows truncated to k, padded with index -1 and distance +inf.
template <typename F>
void knn_batch(const F* X, int64_t n, int64_t d, const F* Q, int64_t m, int64_t k, int metric,
               int exclude, int nthreads, int64_t* out_idx, F* out_dist) {
  parallel_for(m, nthreads, [&](int64_t i) {
    auto r = knn_one(X, n, d, Q + i * d, i, k, metric, exclude);
    for (int64_t j = 0; j < k; ++j) {
      if (j < (int64_t)r.size()) {
        out_idx[i * k + j] = r[j].index;
        out_dist[i * k + j] = r[j].distance;
      } else {
        out_idx[i * k + j] = -1;
        out_dist[i * k + j] = (F)INFINITY;
      }
    }
  });
}

struct Csr {
  int64_t* offsets;  // m + 1
  int64_t* index;    // nnz
  void* dist;        // nnz of F
  int64_t nnz;
};

template <typename F>
void to_csr(std::vector<std::vector<DistPair<F>>>& rows, Csr* out) {
  int64_t m = (int64_t)rows.size();
  out->offsets = (int64_t*)std::malloc(sizeof(int64_t) * (m + 1));
  out->offsets[0] = 0;
  for (int64_t i = 0; i < m; ++i) out->offsets[i + 1] = out->offsets[i] + (int64_t)rows[i].size();
  out->nnz = out->offsets[m];
  out->index = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(1, out->nnz));
  F* dist = (F*)std::malloc(sizeof(F) * std::max<int64_t>(1, out->nnz));
  out->dist = dist;
  for (int64_t i = 0; i < m; ++i)
    for (size_t j = 0; j < rows[i].size(); ++j) {
      out->index[out->offsets[i] + j] = rows[i][j].index;
      dist[out->offsets[i] + j] = rows[i][j].distance;
    }
}

// KNNHeap output with all ties (no truncation): linear_scan_knn as returned to Rust
// callers (src/api/search.rs:169-184); used for PrecomputedKnnSearcher-style tables.
template <typename F>
void knn_ties_batch(const F* X, int64_t n, int64_t d, const F* Q, int64_t m, int64_t k, int metric,
                    int nthreads, Csr* out) {
  std::vector<std::vector<DistPair<F>>> rows((size_t)m);
  parallel_for(m, nthreads,
               [&](int64_t i) { rows[i] = linear_scan_knn(X, n, d, Q + i * d, k, metric); });
  to_csr(rows, out);
}

// a12/a14: linear_scan_range_batch + range_results_to_py, python/search.rs:494-514,
// :1218-1235.  The radius arrives as f64 and is cast to F (python/search.rs:785).
template <typename F>
void range_batch(const F* X, int64_t n, int64_t d, const F* Q, int64_t m, double radius,
                 int metric, int exclude_self, int nthreads, Csr* out) {
  std::vector<std::vector<DistPair<F>>> rows((size_t)m);
  F r = (F)radius;
  parallel_for(m, nthreads, [&](int64_t i) {
    auto v = linear_scan_range(X, n, d, Q + i * d, r, metric);
    if (exclude_self)
      v.erase(std::remove_if(v.begin(), v.end(), [&](const DistPair<F>& p) { return p.index == i; }),
              v.end());
    rows[i] = std::move(v);
  });
  to_csr(rows, out);
}

// a16: compute_pairwise_dense, src/api/square_distance_matrix.rs:12-42: compute
// f(p_i, p_j) for j >= i and mirror it to [j][i].
template <typename F>
void pairwise_dense(const F* X, int64_t n, int64_t d, int metric, int nthreads, F* out) {
  parallel_for(n, nthreads, [&](int64_t i) {
    for (int64_t j = i; j < n; ++j) {
      F v = distance(X + i * d, X + j * d, d, metric);
      out[i * n + j] = v;
      out[j * n + i] = v;
    }
  });
}

// a16: compute_pairwise_condensed, src/api/condensed_distance_matrix.rs:22-54:
// pairs (i, j), 0 <= j < i < n, row-major at index i(i-1)/2 + j, value f(p_i, p_j).
template <typename F>
void pairwise_condensed(const F* X, int64_t n, int64_t d, int metric, int nthreads, F* out) {
  if (n < 2) return;
  parallel_for(n - 1, nthreads, [&](int64_t r) {
    int64_t i = r + 1;
    int64_t base = i * (i - 1) / 2;
    for (int64_t j = 0; j < i; ++j) out[base + j] = distance(X + i * d, X + j * d, d, metric);
  });
}

// a18: for_each_knn, src/outlier/common.rs:12-49 (include_self = false): query index
// idx with k_query = k+1, drop idx, truncate to k.
template <typename F>
std::vector<std::vector<DistPair<F>>> for_each_knn(const F* X, int64_t n, int64_t d, int64_t k_eff,
                                                   int metric, int nthreads) {
  std::vector<std::vector<DistPair<F>>> rows((size_t)n);
  parallel_for(n, nthreads,
               [&](int64_t i) { rows[i] = knn_one(X, n, d, X + i * d, i, k_eff, metric, 1); });
  return rows;
}

// a19: local_outlier_factor, src/outlier/lof.rs:14-81.  Sums are sequential in
// neighbour order in F (Rust's `Sum` left fold).
template <typename F>
void lof(const F* X, int64_t n, int64_t d, int64_t k, int metric, int nthreads, F* scores) {
  int64_t k_eff = std::min<int64_t>(k, n > 0 ? n - 1 : 0);
  if (k_eff == 0) {
    for (int64_t i = 0; i < n; ++i) scores[i] = (F)1;
    return;
  }
  auto nb = for_each_knn(X, n, d, k_eff, metric, nthreads);
  std::vector<F> kdist((size_t)n), lrd((size_t)n);
  for (int64_t i = 0; i < n; ++i) kdist[i] = nb[i].empty() ? (F)0 : nb[i].back().distance;
  for (int64_t i = 0; i < n; ++i) {
    if (nb[i].empty()) {
      lrd[i] = (F)INFINITY;
      continue;
    }
    F s = (F)-0.0;
    for (auto& p : nb[i]) s = s + std::max(kdist[p.index], p.distance);
    lrd[i] = s > (F)0 ? (F)nb[i].size() / s : (F)INFINITY;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (nb[i].empty() || std::isinf(lrd[i])) {
      scores[i] = (F)1;
      continue;
    }
    F s = (F)-0.0;
    for (auto& p : nb[i]) s = s + lrd[p.index];
    scores[i] = s / (lrd[i] * (F)nb[i].size());
  }
}

// k_nearest_neighbors_outlier, src/outlier/knn.rs:14-32: distance to the k-th
// neighbour excluding self (0 when the neighbourhood is empty).
template <typename F>
void knn_outlier(const F* X, int64_t n, int64_t d, int64_t k, int metric, int nthreads, F* scores) {
  int64_t k_eff = std::min<int64_t>(k, n > 0 ? n - 1 : 0);
  auto nb = for_each_knn(X, n, d, k_eff, metric, nthreads);
  for (int64_t i = 0; i < n; ++i) scores[i] = nb[i].empty() ? (F)0 : nb[i].back().distance;
}

// PrecomputedKnnSearcher::subset_with_ties, src/search/precomputed.rs:60-67: from a
// KNNHeap row, keep the prefix with distance <= the k-th entry's distance.
template <typename F>
int64_t subset_with_ties(const F* dist, int64_t len, int64_t k) {
  if (k == 0 || len == 0) return 0;
  F kth = dist[std::min(k, len) - 1];
  int64_t c = 0;
  while (c < len && dist[c] <= kth) ++c;
  return c;
}

// ---------------------------------------------------------------------------
// a17: math::pairwise_sqdist, src/math/mod.rs:562-579 (the k-means distance
// matrix).  d >= 4: the AVX2 micro-kernel (src/math/avx2.rs:888-963,
// :1036-1082) gives every output one sequential fused chain
// acc = fma(a_k - b_k, a_k - b_k, acc) from +0; d < 4: scalar::sqdist
// (src/math/scalar.rs:10-16), the rounded squares summed left to right.
template <typename F>
void kmeans_pairwise(const F* A, int64_t m, const F* B, int64_t n, int64_t d, int nthreads, F* out) {
  parallel_for(m, nthreads, [&](int64_t i) {
    for (int64_t j = 0; j < n; ++j) {
      F acc = 0;
      for (int64_t k = 0; k < d; ++k) {
        F t = A[i * d + k] - B[j * d + k];
        if (d >= 4)
          acc = std::fma(t, t, acc);
        else
          acc = acc + t * t;
      }
      out[i * n + j] = acc;
    }
  });
}

// ---------------------------------------------------------------------------
// Density clustering over the linear-scan range search.
//
// dbscan, src/cluster/dbscan.rs:14-83: a hash-set frontier, drained by taking an
// arbitrary element (the reference's frontier.iter().next()).
template <typename F>
void dbscan(const F* X, int64_t n, int64_t d, double eps, int64_t min_points, int metric, int64_t* labels) {
  const int64_t UNVISITED = -2, NOISE = -1;
  F e = (F)eps;
  for (int64_t i = 0; i < n; ++i) labels[i] = UNVISITED;
  int64_t cluster_id = 0;
  std::unordered_set<int64_t> frontier;
  std::vector<int64_t> neighbors;
  for (int64_t p = 0; p < n; ++p) {
    if (labels[p] != UNVISITED) continue;
    frontier.clear();
    for (auto& pr : linear_scan_range(X, n, d, X + p * d, e, metric)) frontier.insert(pr.index);
    if ((int64_t)frontier.size() < min_points) {
      labels[p] = NOISE;
      continue;
    }
    labels[p] = cluster_id;
    while (!frontier.empty()) {
      int64_t cur = *frontier.begin();
      frontier.erase(frontier.begin());
      if (labels[cur] == NOISE) labels[cur] = cluster_id;
      if (labels[cur] != UNVISITED) continue;
      labels[cur] = cluster_id;
      neighbors.clear();
      for (auto& pr : linear_scan_range(X, n, d, X + cur * d, e, metric)) neighbors.push_back(pr.index);
      if ((int64_t)neighbors.size() >= min_points) {
        for (int64_t nb : neighbors) {
          if (labels[nb] == NOISE) labels[nb] = cluster_id;
          if (labels[nb] == UNVISITED) frontier.insert(nb);
        }
      }
    }
    ++cluster_id;
  }
}

// parallel_dbscan, src/cluster/parallel_dbscan.rs:16-113 (sequentially): core
// flags, union of neighbouring cores, labels in order of each component's first
// core, border points take the first core of their range result.
template <typename F>
void parallel_dbscan(const F* X, int64_t n, int64_t d, double eps, int64_t min_points, int metric,
                     int nthreads, int64_t* labels) {
  F e = (F)eps;
  std::vector<std::vector<DistPair<F>>> nb((size_t)n);
  parallel_for(n, nthreads, [&](int64_t i) { nb[i] = linear_scan_range(X, n, d, X + i * d, e, metric); });
  std::vector<int64_t> core_ids(n, -1), cores;
  for (int64_t i = 0; i < n; ++i)
    if ((int64_t)nb[i].size() >= min_points) {
      core_ids[i] = (int64_t)cores.size();
      cores.push_back(i);
    }
  if (cores.empty()) {
    for (int64_t i = 0; i < n; ++i) labels[i] = -1;
    return;
  }
  std::vector<int64_t> parent(cores.size()), rank(cores.size(), 0);
  for (size_t i = 0; i < parent.size(); ++i) parent[i] = (int64_t)i;
  auto find = [&](int64_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  auto unite = [&](int64_t a, int64_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (rank[a] < rank[b]) std::swap(a, b);
    parent[b] = a;
    if (rank[a] == rank[b]) ++rank[a];
  };
  for (size_t c = 0; c < cores.size(); ++c)
    for (auto& pr : nb[cores[c]])
      if (core_ids[pr.index] >= 0) unite((int64_t)c, core_ids[pr.index]);
  std::vector<int64_t> root_label(cores.size(), -1), label_of(cores.size());
  int64_t next = 0;
  for (size_t c = 0; c < cores.size(); ++c) {
    int64_t r = find((int64_t)c);
    if (root_label[r] < 0) root_label[r] = next++;
    label_of[c] = root_label[r];
  }
  for (int64_t i = 0; i < n; ++i) {
    if (core_ids[i] >= 0) {
      labels[i] = label_of[core_ids[i]];
      continue;
    }
    int64_t a = -1;
    for (auto& pr : nb[i]) {
      if (a != -1) break;
      if (core_ids[pr.index] >= 0) a = label_of[core_ids[pr.index]];
    }
    labels[i] = a;
  }
}

// optics, src/cluster/optics.rs:56-133, :306-377: the seeds heap orders by
// reachability then index (ReachCandidate's reversed Ord, :28-35).
template <typename F>
void optics(const F* X, int64_t n, int64_t d, double eps, int64_t min_points, int metric, int64_t* ordering,
            F* reach, F* core, int64_t* pred, int64_t* labels) {
  F e = (F)eps;
  const F inf = std::numeric_limits<F>::infinity();
  std::vector<char> processed(n, 0), is_core(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    reach[i] = inf;
    core[i] = std::numeric_limits<F>::quiet_NaN();
    pred[i] = -1;
  }
  struct RC {
    F r;
    int64_t i;
    bool operator<(const RC& o) const {  // max-heap on the reversed order
      if (o.r < r) return true;
      if (r < o.r) return false;
      return o.i < i;
    }
  };
  std::priority_queue<RC> seeds;
  int64_t no = 0;
  auto region = [&](int64_t p) {
    std::vector<std::pair<int64_t, F>> v;
    for (auto& pr : linear_scan_range(X, n, d, X + p * d, e, metric)) v.push_back({pr.index, pr.distance});
    return v;
  };
  auto core_distance = [&](std::vector<std::pair<int64_t, F>>& v, F& out) {
    if ((int64_t)v.size() < min_points) return false;
    std::nth_element(v.begin(), v.begin() + (min_points - 1), v.end(),
                     [](const std::pair<int64_t, F>& a, const std::pair<int64_t, F>& b) { return a.second < b.second; });
    out = v[min_points - 1].second;
    return true;
  };
  auto update = [&](const std::vector<std::pair<int64_t, F>>& v, int64_t p, F cd) {
    for (auto& pr : v) {
      if (pr.first == p || processed[pr.first]) continue;
      F nr = std::max(cd, pr.second);
      if (nr < reach[pr.first]) {
        reach[pr.first] = nr;
        pred[pr.first] = p;
        seeds.push(RC{nr, pr.first});
      }
    }
  };
  for (int64_t p = 0; p < n; ++p) {
    if (processed[p]) continue;
    auto nbs = region(p);
    processed[p] = 1;
    ordering[no++] = p;
    F cd;
    if (core_distance(nbs, cd)) {
      core[p] = cd;
      is_core[p] = 1;
      while (!seeds.empty()) seeds.pop();
      update(nbs, p, cd);
      while (!seeds.empty()) {
        RC c = seeds.top();
        seeds.pop();
        if (processed[c.i]) continue;
        if (c.r > reach[c.i]) continue;
        auto cn = region(c.i);
        processed[c.i] = 1;
        ordering[no++] = c.i;
        F cc;
        if (core_distance(cn, cc)) {
          core[c.i] = cc;
          is_core[c.i] = 1;
          update(cn, c.i, cc);
        }
      }
    }
  }
  int64_t cid = -1;
  for (int64_t o = 0; o < n; ++o) {
    int64_t p = ordering[o];
    F r = reach[p], c = is_core[p] ? core[p] : inf;
    if (r > e)
      labels[p] = c <= e ? ++cid : -1;
    else
      labels[p] = cid >= 0 ? cid : -1;
  }
}

// extract_xi_labels, src/cluster/optics.rs:148-304, :379-453: steep-down areas
// kept in a list, cluster intervals from each steep-up area, the predecessor
// correction, then the smallest intervals claim their points first.
template <typename F>
void extract_xi(const int64_t* ordering, const F* reach, const int64_t* pred, int64_t n, double xi_in,
                int64_t min_points, int64_t* labels) {
  for (int64_t i = 0; i < n; ++i) labels[i] = -1;
  if (n == 0) return;
  const F xi = (F)xi_in, ixi = (F)1 - xi;
  const F inf = std::numeric_limits<F>::infinity();
  std::vector<F> r(n);
  std::vector<int64_t> pos(n, -1);
  for (int64_t o = 0; o < n; ++o) {
    r[o] = reach[ordering[o]];
    pos[ordering[o]] = o;
  }
  auto steep_up = [&](int64_t i, F f) { return std::isfinite(r[i]) && (i + 1 >= n || r[i] <= r[i + 1] * f); };
  auto steep_down = [&](int64_t i, F f) { return i + 1 < n && std::isfinite(r[i + 1]) && r[i + 1] <= r[i] * f; };
  auto next_r = [&](int64_t i) { return i + 1 < n ? r[i + 1] : inf; };
  struct SDA {
    int64_t start;
    F maximum, mib;
  };
  auto filter = [&](F mib, std::vector<SDA>& set) {
    std::vector<SDA> keep;
    for (auto s : set) {
      if (s.maximum * ixi <= mib) continue;
      if (mib > s.mib) s.mib = mib;
      keep.push_back(s);
    }
    set.swap(keep);
  };
  auto pred_filter = [&](int64_t cstart, int64_t cend) {
    if (cend >= n) return n - 1;
    F sv = r[cstart];
    while (cend > cstart) {
      int64_t p = ordering[cend];
      if (r[cend] < sv) break;
      int64_t pp = pred[p];
      if (pp >= 0 && pos[pp] >= cstart && pos[pp] < cend) break;
      --cend;
    }
    return cend;
  };
  F mib = 0;
  std::vector<SDA> sdaset;
  std::vector<std::pair<int64_t, int64_t>> intervals;
  int64_t scan = 0;
  while (scan < n) {
    mib = std::max(mib, r[scan]);
    if (steep_down(scan, ixi)) {
      filter(mib, sdaset);
      F startval = r[scan];
      mib = 0;
      int64_t startsteep = scan, endsteep = scan;
      ++scan;
      while (scan < n) {
        if (steep_down(scan, ixi))
          endsteep = scan;
        else if (!steep_down(scan, (F)1) || scan - endsteep > min_points)
          break;
        ++scan;
      }
      sdaset.push_back(SDA{startsteep, startval, (F)0});
      continue;
    }
    if (steep_up(scan, ixi)) {
      filter(mib, sdaset);
      int64_t endsteep = scan;
      mib = r[scan];
      F esuccr = next_r(scan);
      while (std::isfinite(esuccr) && scan < n) {
        ++scan;
        if (scan >= n) break;
        if (steep_up(scan, ixi)) {
          endsteep = scan;
          mib = r[scan];
          esuccr = next_r(scan);
        } else if (!steep_up(scan, (F)1) || scan - endsteep > min_points) {
          break;
        }
      }
      if (std::isinf(esuccr) && scan < n) ++scan;
      for (auto it = sdaset.rbegin(); it != sdaset.rend(); ++it) {
        int64_t cstart = it->start, cend = pred_filter(it->start, endsteep);
        F eu = cend + 1 < n ? r[cend + 1] : inf;
        if (it->mib > std::min(it->maximum, eu) * ixi) continue;
        if (it->maximum * ixi >= eu) {
          while (cstart < cend && r[cstart + 1] * ixi > eu) ++cstart;
        } else if (eu * ixi >= it->maximum) {
          while (cend > cstart && r[cend] * ixi > it->maximum) --cend;
        }
        cend = pred_filter(cstart, cend);
        if (cend + 1 - cstart < min_points) continue;
        intervals.push_back({cstart, cend});
      }
      continue;
    }
    ++scan;
  }
  std::stable_sort(intervals.begin(), intervals.end(), [](const auto& a, const auto& b) {
    return std::make_pair(a.second - a.first, a.first) < std::make_pair(b.second - b.first, b.first);
  });
  std::vector<char> claimed(n, 0);
  int64_t cid = 0;
  for (auto& iv : intervals) {
    bool any = false;
    for (int64_t o = iv.first; o <= iv.second; ++o) {
      int64_t p = ordering[o];
      if (!claimed[p]) {
        claimed[p] = 1;
        labels[p] = cid;
        any = true;
      }
    }
    if (any) ++cid;
  }
}

// a10: LinearScanPrioritySearcher, src/search/linear_scan.rs:70-235: one
// query's searcher.  ensure_initialized (:105-130) materialises every point with
// skip <= d <= cutoff, sorted by DistPair order; find_next (:132-155),
// all_lower_bound_internal (:157-175), decrease_cutoff (:177-180) and
// reset_with_limits (:200-209) follow the reference line by line.
template <typename F>
struct PrioritySearcher {
  const F* X;
  int64_t n, d;
  std::vector<F> q;
  int metric;
  std::vector<DistPair<F>> candidates;
  size_t position = 0;
  F cutoff = (F)INFINITY;
  F skip = 0;
  bool initialized = false;

  void reset_state() {
    position = 0;
    cutoff = (F)INFINITY;
    skip = 0;
    initialized = false;
  }
  void ensure_initialized() {
    if (initialized) return;
    std::vector<DistPair<F>> c;
    c.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
      const F dd = distance(q.data(), X + i * d, d, metric);
      if (dd < skip || dd > cutoff) continue;
      c.push_back({dd, i});
    }
    std::sort(c.begin(), c.end(), PairLess<F>());
    candidates.swap(c);
    position = 0;
    initialized = true;
  }
  template <typename Filter>
  bool find_next(Filter skip_point, DistPair<F>& out) {
    ensure_initialized();
    while (position < candidates.size()) {
      const DistPair<F> cand = candidates[position];
      position += 1;
      if (cand.distance > cutoff) return false;
      if (cand.distance < skip) continue;
      if (skip_point(cand.index)) continue;
      out = cand;
      return true;
    }
    return false;
  }
  F all_lower_bound() const {
    if (!initialized) return 0;
    for (size_t i = position; i < candidates.size(); ++i) {
      const F dd = candidates[i].distance;
      if (dd > cutoff) break;
      if (dd >= skip) return dd;
    }
    return (F)INFINITY;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// C entry points for ctypes (tests, smoke and the CPU-baseline leg only).
extern "C" {

float oracle_sqdist_f32(const float* a, const float* b, int64_t d) { return sqdist_f32_avx(a, b, d); }
float oracle_sqdist_f32_lanes(const float* a, const float* b, int64_t d) {
  return sqdist_f32_lanes(a, b, d);
}
double oracle_sqdist_f64(const double* a, const double* b, int64_t d) { return sqdist_f64_avx(a, b, d); }
float oracle_sqdist_f32_fallback(const float* a, const float* b, int64_t d) {
  return sqdist_fallback<float>(a, b, d);
}
double oracle_sqdist_f64_fallback(const double* a, const double* b, int64_t d) {
  return sqdist_fallback<double>(a, b, d);
}

// Row-wise restated values of one query against every point (no sqrt).
void oracle_sqdist_rows_f32(const float* X, int64_t n, int64_t d, const float* q, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = sqdist_f32_avx(q, X + i * d, d);
}

// FP64 brute-force squared distance (the 1e-5-relative distance oracle): the
// coordinates are widened to double and summed in order; each step has
// relative error <= 2^-53, far below the f32 tolerances it checks.
void oracle_sqdist_fp64_pairs(const float* X, int64_t d, const float* Q, const int64_t* qi,
                              const int64_t* xi, int64_t count, double* out) {
  for (int64_t t = 0; t < count; ++t) {
    const float* a = Q + qi[t] * d;
    const float* b = X + xi[t] * d;
    double s = 0.0;
    for (int64_t j = 0; j < d; ++j) {
      double diff = (double)a[j] - (double)b[j];
      s += diff * diff;
    }
    out[t] = s;
  }
}

void oracle_knn_f32(const float* X, int64_t n, int64_t d, const float* Q, int64_t m, int64_t k,
                    int metric, int exclude, int nthreads, int64_t* out_idx, float* out_dist) {
  knn_batch<float>(X, n, d, Q, m, k, metric, exclude, nthreads, out_idx, out_dist);
}
void oracle_knn_f64(const double* X, int64_t n, int64_t d, const double* Q, int64_t m, int64_t k,
                    int metric, int exclude, int nthreads, int64_t* out_idx, double* out_dist) {
  knn_batch<double>(X, n, d, Q, m, k, metric, exclude, nthreads, out_idx, out_dist);
}

void oracle_knn_ties_f32(const float* X, int64_t n, int64_t d, const float* Q, int64_t m, int64_t k,
                         int metric, int nthreads, Csr* out) {
  knn_ties_batch<float>(X, n, d, Q, m, k, metric, nthreads, out);
}
void oracle_knn_ties_f64(const double* X, int64_t n, int64_t d, const double* Q, int64_t m,
                         int64_t k, int metric, int nthreads, Csr* out) {
  knn_ties_batch<double>(X, n, d, Q, m, k, metric, nthreads, out);
}

void oracle_range_f32(const float* X, int64_t n, int64_t d, const float* Q, int64_t m, double radius,
                      int metric, int exclude_self, int nthreads, Csr* out) {
  range_batch<float>(X, n, d, Q, m, radius, metric, exclude_self, nthreads, out);
}
void oracle_range_f64(const double* X, int64_t n, int64_t d, const double* Q, int64_t m,
                      double radius, int metric, int exclude_self, int nthreads, Csr* out) {
  range_batch<double>(X, n, d, Q, m, radius, metric, exclude_self, nthreads, out);
}

void oracle_csr_free(Csr* c) {
  std::free(c->offsets);
  std::free(c->index);
  std::free(c->dist);
  c->offsets = nullptr;
  c->index = nullptr;
  c->dist = nullptr;
}

void oracle_pairwise_f32(const float* X, int64_t n, int64_t d, int metric, int form, int nthreads,
                         float* out) {
  if (form == 0)
    pairwise_dense<float>(X, n, d, metric, nthreads, out);
  else
    pairwise_condensed<float>(X, n, d, metric, nthreads, out);
}
void oracle_pairwise_f64(const double* X, int64_t n, int64_t d, int metric, int form, int nthreads,
                         double* out) {
  if (form == 0)
    pairwise_dense<double>(X, n, d, metric, nthreads, out);
  else
    pairwise_condensed<double>(X, n, d, metric, nthreads, out);
}

void oracle_lof_f32(const float* X, int64_t n, int64_t d, int64_t k, int metric, int nthreads,
                    float* scores) {
  lof<float>(X, n, d, k, metric, nthreads, scores);
}
void oracle_lof_f64(const double* X, int64_t n, int64_t d, int64_t k, int metric, int nthreads,
                    double* scores) {
  lof<double>(X, n, d, k, metric, nthreads, scores);
}
void oracle_knn_outlier_f32(const float* X, int64_t n, int64_t d, int64_t k, int metric,
                            int nthreads, float* scores) {
  knn_outlier<float>(X, n, d, k, metric, nthreads, scores);
}
void oracle_knn_outlier_f64(const double* X, int64_t n, int64_t d, int64_t k, int metric,
                            int nthreads, double* scores) {
  knn_outlier<double>(X, n, d, k, metric, nthreads, scores);
}

int64_t oracle_subset_with_ties_f64(const double* dist, int64_t len, int64_t k) {
  return subset_with_ties<double>(dist, len, k);
}
int64_t oracle_subset_with_ties_f32(const float* dist, int64_t len, int64_t k) {
  return subset_with_ties<float>(dist, len, k);
}

void oracle_kmeans_pairwise_f32(const float* A, int64_t m, const float* B, int64_t n, int64_t d, int nthreads,
                                float* out) {
  kmeans_pairwise<float>(A, m, B, n, d, nthreads, out);
}
void oracle_kmeans_pairwise_f64(const double* A, int64_t m, const double* B, int64_t n, int64_t d, int nthreads,
                                double* out) {
  kmeans_pairwise<double>(A, m, B, n, d, nthreads, out);
}

void oracle_dbscan_f32(const float* X, int64_t n, int64_t d, double eps, int64_t min_points, int metric,
                       int variant, int nthreads, int64_t* labels) {
  if (variant == 0)
    dbscan<float>(X, n, d, eps, min_points, metric, labels);
  else
    parallel_dbscan<float>(X, n, d, eps, min_points, metric, nthreads, labels);
}
void oracle_dbscan_f64(const double* X, int64_t n, int64_t d, double eps, int64_t min_points, int metric,
                       int variant, int nthreads, int64_t* labels) {
  if (variant == 0)
    dbscan<double>(X, n, d, eps, min_points, metric, labels);
  else
    parallel_dbscan<double>(X, n, d, eps, min_points, metric, nthreads, labels);
}
void oracle_optics_f32(const float* X, int64_t n, int64_t d, double eps, int64_t min_points, int metric,
                       int64_t* ordering, float* reach, float* core, int64_t* pred, int64_t* labels) {
  optics<float>(X, n, d, eps, min_points, metric, ordering, reach, core, pred, labels);
}
void oracle_optics_f64(const double* X, int64_t n, int64_t d, double eps, int64_t min_points, int metric,
                       int64_t* ordering, double* reach, double* core, int64_t* pred, int64_t* labels) {
  optics<double>(X, n, d, eps, min_points, metric, ordering, reach, core, pred, labels);
}
void oracle_extract_xi_f32(const int64_t* ordering, const float* reach, const int64_t* pred, int64_t n,
                           double xi, int64_t min_points, int64_t* labels) {
  extract_xi<float>(ordering, reach, pred, n, xi, min_points, labels);
}
void oracle_extract_xi_f64(const int64_t* ordering, const double* reach, const int64_t* pred, int64_t n,
                           double xi, int64_t min_points, int64_t* labels) {
  extract_xi<double>(ordering, reach, pred, n, xi, min_points, labels);
}


typedef int (*oracle_skip_fn)(void* ctx, int64_t index);
#define ORACLE_PRIORITY(T, F)                                                                                \
  void* oracle_priority_create_##T(const F* X, int64_t n, int64_t d, const F* q, int metric) {            \
    auto* s = new PrioritySearcher<F>();                                                                   \
    s->X = X;                                                                                              \
    s->n = n;                                                                                              \
    s->d = d;                                                                                              \
    s->q.assign(q, q + d);                                                                                 \
    s->metric = metric;                                                                                    \
    return s;                                                                                              \
  }                                                                                                        \
  void oracle_priority_destroy_##T(void* h) { delete static_cast<PrioritySearcher<F>*>(h); }               \
  void oracle_priority_reset_##T(void* h, double cutoff, double skip) {                                    \
    auto* s = static_cast<PrioritySearcher<F>*>(h);                                                        \
    s->reset_state();                                                                                      \
    s->cutoff = (F)cutoff;                                                                                 \
    s->skip = (F)skip;                                                                                     \
  }                                                                                                        \
  void oracle_priority_decrease_cutoff_##T(void* h, double t) { static_cast<PrioritySearcher<F>*>(h)->cutoff = (F)t; } \
  int oracle_priority_next_##T(void* h, oracle_skip_fn fn, void* ctx, int64_t* idx, F* dist) {             \
    DistPair<F> out{};                                                                                     \
    const bool found = static_cast<PrioritySearcher<F>*>(h)->find_next(                                   \
        [&](int64_t i) { return fn ? fn(ctx, i) != 0 : false; }, out);                                     \
    *idx = found ? out.index : -1;                                                                         \
    *dist = found ? out.distance : (F)INFINITY;                                                            \
    return found ? 1 : 0;                                                                                  \
  }                                                                                                        \
  F oracle_priority_all_lower_bound_##T(void* h) { return static_cast<PrioritySearcher<F>*>(h)->all_lower_bound(); }
ORACLE_PRIORITY(f32, float)
ORACLE_PRIORITY(f64, double)
#undef ORACLE_PRIORITY

}  // extern "C"
