/*
 * fuel_amd — MI355X (gfx950) exact linear-scan kNN / range / pairwise
 * squared-Euclidean engine behind a C ABI.
 *
 * Drop-in boundary for fuel's linear-scan search path.  Every entry point below
 * names the reference interface it replaces (paths relative to the fuel
 * repository).  Conventions shared by all entry points:
 *   - return value: FUEL_AMD_OK (0) or an error code; the message of the last
 *     error on the calling thread is available from fuel_amd_last_error().
 *     Nothing aborts the process.  (fuel: Result<_, String> -> PyRuntimeError,
 *     src/api/parallel.rs:167-171; bad names -> ValueError, python/mod.rs:64-69.)
 *   - data and queries are C-contiguous row-major (rows x d) arrays of the
 *     index dtype, like the numpy arrays fuel takes (src/api/ndarray.rs:96-99).
 *   - results are identical to fuel's CPU linear scan: distances are the f32
 *     (or f64) values of fuel's own AVX2 kernel
 *     (src/distance/squared_euclidean.rs:13-67), sqrt'd for 'euclidean'
 *     (src/distance/euclidean.rs:7-9), and neighbours are ordered by
 *     (distance, index) (src/api/search.rs:23-30).
 *   - host-pointer entry points are synchronous; *_device entry points take
 *     device pointers and a hipStream_t (as void*) and enqueue their work on
 *     it; fuel_amd_knn_device in an MFMA mode also waits for the stream once
 *     per query chunk (see there).
 *   - an index is immutable after creation.  kNN calls (fuel_amd_knn,
 *     fuel_amd_knn_device) on one index from several threads run concurrently,
 *     each on a context and stream of its own (fuel's SearchIndex is shared by
 *     rayon's worker threads); the other entry points on one index are
 *     serialized internally.  fuel_amd_last_stats reports the last kNN call to
 *     finish.
 */
#ifndef FUEL_AMD_H
#define FUEL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FUEL_AMD_OK 0
#define FUEL_AMD_ERR_INVALID 1     /* bad argument (ValueError in fuel's Python layer) */
#define FUEL_AMD_ERR_UNSUPPORTED 2 /* valid request outside this engine's scope */
#define FUEL_AMD_ERR_DEVICE 3      /* HIP runtime error or no usable GPU */
#define FUEL_AMD_ERR_INTERRUPTED 4 /* fuel_amd_request_cancel() was called (src/api/parallel.rs:73) */
#define FUEL_AMD_ERR_NOMEM 5       /* device or host allocation failed */

/* dtypes (fuel: build_linear_scan_f32 / _f64, python/search.rs:1353-1389) */
#define FUEL_AMD_F32 0
#define FUEL_AMD_F64 1

/* metrics (fuel: parse_distance_fn "euclidean"|"l2" and
 * "sqeuclidean"|"squared_euclidean", python/mod.rs:49-50) */
#define FUEL_AMD_EUCLIDEAN 0
#define FUEL_AMD_SQEUCLIDEAN 1

/* self-exclusion for self-queries (query row i is data row i) */
#define FUEL_AMD_EXCLUDE_NONE 0
#define FUEL_AMD_EXCLUDE_FOR_EACH_KNN 1 /* search k+1, drop self, truncate: src/outlier/common.rs:27,41-44 */
#define FUEL_AMD_EXCLUDE_PY_LINEAR 2    /* search k, drop self: python/search.rs:485-488 */

/* candidate-generation precision.  Every mode returns the exact reference
 * answer; the modes differ only in how candidates are generated before the
 * exact rerank and certification. */
#define FUEL_AMD_MODE_AUTO 0
#define FUEL_AMD_MODE_EXACT 1 /* restated distance for every pair */
#define FUEL_AMD_MODE_FP32 2  /* FP32 MFMA ||x||^2-2x.y candidates + exact rerank */
#define FUEL_AMD_MODE_BF16 3  /* BF16 MFMA candidates + exact FP32 rerank + FP64 bound */
#define FUEL_AMD_MODE_BF16R 4 /* as BF16, on residuals against coarse centroids: rows partitioned
                                 into lists at index build; for clustered data, where the plain
                                 form's bound (~ |q| |x|) is wider than the neighbour gaps */
#define FUEL_AMD_MODE_FP8R 5  /* as BF16R with the resident copy in FP8 E4M3 (one byte per
                                 coordinate, a power-of-two scale per row) and the query side
                                 split into two E4M3 parts (hi + lo); d <= 256; default
                                 widen 16, up to 256 candidates (the FP8 bound band is
                                 wider) */

/* pairwise output forms (fuel: compute_pairwise_dense / _condensed) */
#define FUEL_AMD_FORM_SQUARE 0
#define FUEL_AMD_FORM_CONDENSED 1 /* fuel's internal order: pair (i>j) at i(i-1)/2+j */

typedef struct fuel_amd_index fuel_amd_index;

/* Index option flags.
 * FUEL_AMD_FLAG_FALLBACK_ARITH: distances take the rounding of fuel's
 * squared_euclidean_distance_fallback (src/distance/squared_euclidean.rs:99-133),
 * which fuel runs on every host without AVX+FMA (the dispatch at :71-97): 8
 * lanes of fl(t * t) + acc for f32 and f64, a left fold from +0, an unfused
 * tail.  Without the flag the engine matches the AVX+FMA kernels (:13-67), the
 * ones fuel runs on x86 hosts with AVX2 and FMA.  Results are then identical to
 * fuel on such a host; every mode and entry point of the index follows the flag
 * (the batched problems of fuel_amd_batch_* always use the AVX form). */
#define FUEL_AMD_FLAG_FALLBACK_ARITH 1
/* FUEL_AMD_FLAG_NO_SEED: start the plain MFMA scans (FP32 / BF16) at no
 * threshold instead of the seeded one (the KP-th smallest of per-row-group key
 * minima over a ~1/11 row sample, computed on the matrix cores first), and the
 * residual scans (BF16R / FP8R) at their list bound instead of the sampled one
 * (the KP-th smallest exact distance to the first rows of the query's nearest
 * list).  Answers are the same either way; the seeds only save candidate
 * appends. */
#define FUEL_AMD_FLAG_NO_SEED 2

typedef struct fuel_amd_opts {
  int32_t mode;   /* FUEL_AMD_MODE_* */
  int32_t widen;  /* candidate widening factor for the MFMA modes (0: 4, or 16 for FP8R) */
  int32_t device; /* HIP device ordinal (default 0) */
  int32_t flags;  /* FUEL_AMD_FLAG_* (0: none) */
  int64_t reserved[6];
} fuel_amd_opts;

/* Variable-length range results as CSR; allocated by the library, released by
 * fuel_amd_csr_free.  dist has the index dtype. */
typedef struct fuel_amd_csr {
  int64_t* offsets; /* m + 1 */
  int64_t* index;   /* nnz */
  void* dist;       /* nnz */
  int64_t nnz;
  int64_t m;
} fuel_amd_csr;

/* Statistics of the last kNN call on an index (certification / fallback). */
typedef struct fuel_amd_stats {
  int64_t queries;
  int64_t fallback_queries; /* queries whose first-level bound did not certify */
  int32_t mode_used;
  int32_t splits;
  int32_t kp;
  int32_t exact_queries;    /* of those, queries that needed the exact restated scan */
  double ms_total;
} fuel_amd_stats;

/* Version string of the library. */
const char* fuel_amd_version(void);

/* Number of usable GPUs (0 when none: every compute entry point then fails with
 * FUEL_AMD_ERR_DEVICE; there is no CPU fallback). */
int32_t fuel_amd_device_count(void);

/* Message of the last error on this thread ("" if none). */
const char* fuel_amd_last_error(void);

/* Cooperative cancellation (fuel: SHUTDOWN_REQUESTED, src/api/mod.rs:29, polled
 * in par_try_map, src/api/parallel.rs:73).  Running calls poll the flag between
 * launches and return FUEL_AMD_ERR_INTERRUPTED. */
void fuel_amd_request_cancel(void);
void fuel_amd_clear_cancel(void);

/* Build a resident linear-scan index: copies the (n x d) rows to HBM.
 * Replaces build_linear_scan_f32 / build_linear_scan_f64
 * (python/search.rs:1353-1389) and LinearScanSearcher::new
 * (src/search/linear_scan.rs:8-24).  opts may be NULL. */
int32_t fuel_amd_index_create(const void* data, int64_t n, int32_t d, int32_t dtype, int32_t metric,
                              const fuel_amd_opts* opts, fuel_amd_index** out);
/* Same, from rows already resident in device memory (row stride ld elements). */
int32_t fuel_amd_index_create_device(const void* d_data, int64_t n, int32_t d, int64_t ld,
                                     int32_t dtype, int32_t metric, const fuel_amd_opts* opts,
                                     fuel_amd_index** out);
int32_t fuel_amd_index_destroy(fuel_amd_index* index);

/* kNN of m query rows.  out_idx (m x k int64, -1 padded) and out_dist
 * (m x k of the index dtype, +inf padded) are caller-allocated host arrays.
 * Replaces SearchIndex.knn for the LinearScan variants (python/search.rs:560-569,
 * LinearScanSearcher::query_search_knn :524-539, linear_scan_knn_batch
 * :471-491, knn_results_to_arrays :1195-1216) and, with
 * FUEL_AMD_EXCLUDE_FOR_EACH_KNN, the search loop of for_each_knn
 * (src/outlier/common.rs:12-49). */
int32_t fuel_amd_knn(fuel_amd_index* index, const void* query, int64_t m, int32_t k,
                     int32_t exclude_self, int64_t* out_idx, void* out_dist);
/* Device-pointer form (query rows with stride ldq elements), on `stream`.
 * In the exact mode the call only enqueues work and returns.  In an MFMA mode
 * (FP32 / BF16 / BF16R, and AUTO, which picks one of them) the call also waits
 * for `stream` once per query chunk of up to ~100k queries, after that chunk's
 * exact rerank: it reads the number of queries whose certificate failed,
 * because the launches of the next certification level are sized from it.  On
 * return the results of every chunk but the last are complete; the last
 * chunk's results are complete when it certified (the common case) and are
 * otherwise finished by kernels still queued on `stream`.  Synchronize
 * `stream` before reading the outputs from another stream or the host. */
int32_t fuel_amd_knn_device(fuel_amd_index* index, const void* d_query, int64_t m, int64_t ldq,
                            int32_t k, int32_t exclude_self, int64_t* d_out_idx, void* d_out_dist,
                            void* stream);

/* Allocation-free kNN.  No counterpart in fuel: it is for callers that cannot
 * allocate device memory inside the call (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and answers
 * as fuel_amd_knn_device, computed by the exact restated scan whatever the
 * index mode, with k + (exclude_self ? 1 : 0) <= 256.  The candidate slots and
 * list counts live in a fixed 64 MiB device array per GPU that the library
 * holds until the process ends; the call calls no hipMalloc / hipFree and never
 * waits for `stream`.  A batch that does not fit the array even with one DB
 * split (about 16,000 f32 or 8,000 f64 queries) is refused with
 * FUEL_AMD_ERR_INVALID.  Calls on one index are serialized, but every call on
 * a GPU uses that GPU's one array: calls on different indexes of one GPU whose
 * work overlaps on the device (different streams) race; order them on one
 * stream. */
int32_t fuel_amd_knn_device_noalloc(fuel_amd_index* index, const void* d_query, int64_t m, int64_t ldq, int32_t k,
                                    int32_t exclude_self, int64_t* d_out_idx, void* d_out_dist, void* stream);

/* All neighbours within `radius` (inclusive), sorted by (distance, index).
 * Replaces SearchIndex.radius_search for the LinearScan variants
 * (python/search.rs:776-787, linear_scan_range_batch :494-514,
 * range_results_to_py :1218-1235; src/api/search.rs:154-166).  The radius is
 * cast to the index dtype first, like fuel (python/search.rs:785). */
int32_t fuel_amd_range(fuel_amd_index* index, const void* query, int64_t m, double radius,
                       int32_t exclude_self, fuel_amd_csr* out);
void fuel_amd_csr_free(fuel_amd_csr* csr);

/* kNN with fuel's tie semantics: for each query, every point whose key is <=
 * the k-th smallest key (so possibly more than k), sorted by (distance,
 * index), as CSR (library-allocated, fuel_amd_csr_free).  k = 0 gives empty
 * lists; k >= n gives all points.  Replaces KNNHeap-based search_knn
 * (src/api/search.rs:80-151, :169-184), the source of the table of
 * PrecomputedKnnSearcher::new (src/search/precomputed.rs:36-56), which
 * subset_with_ties (:60-67) then cuts per request. */
int32_t fuel_amd_knn_ties(fuel_amd_index* index, const void* query, int64_t m, int32_t k, fuel_amd_csr* out);

/* ---- priority searchers ------------------------------------------------------
 * fuel's LinearScanPrioritySearcher (src/search/linear_scan.rs:70-235, the
 * PrioritySearcher trait of src/api/search.rs:216-229) for m queries at once:
 * slot q is the searcher of query row q (index dtype, d per row).  Each slot
 * keeps the reference's state -- cutoff (initially +inf), skip (initially 0),
 * an initialized flag and a position -- and on its first next() the list of
 * every row with skip <= distance <= cutoff sorted by (distance, index), which
 * the engine computes on the GPU for a group of slots at once.  The answers
 * equal the reference searcher's call for call.  Calls on one searcher are
 * serialized; the index must outlive it. */
typedef struct fuel_amd_priority fuel_amd_priority;
int32_t fuel_amd_priority_create(fuel_amd_index* index, const void* query, int64_t m, fuel_amd_priority** out);
int32_t fuel_amd_priority_destroy(fuel_amd_priority* searcher);
/* reset_with_limits(cutoff, skip) (:200-209); reset() is cutoff = +inf, skip = 0.
 * Limits are cast to the index dtype; 0 <= skip <= cutoff is required. */
int32_t fuel_amd_priority_reset(fuel_amd_priority* searcher, int64_t q, double cutoff, double skip);
/* decrease_cutoff (:189-194); a larger threshold is FUEL_AMD_ERR_INVALID. */
int32_t fuel_amd_priority_decrease_cutoff(fuel_amd_priority* searcher, int64_t q, double threshold);
/* next / next_with_filter (:132-155): the next (index, distance) of slot q, or
 * index -1 (distance +inf) when none is left.  skip_point (may be NULL) is the
 * SearchFilter: a nonzero return skips that row. */
int32_t fuel_amd_priority_next(fuel_amd_priority* searcher, int64_t q, int32_t (*skip_point)(void* ctx, int64_t index),
                               void* ctx, int64_t* out_index, void* out_distance);
/* all_lower_bound (:157-175): 0 before the first next(), else the distance of
 * the next candidate within the limits (+inf when none). */
int32_t fuel_amd_priority_all_lower_bound(fuel_amd_priority* searcher, int64_t q, void* out_distance);

/* DBSCAN over the index rows with the exact range search (eps inclusive, cast
 * to the index dtype like fuel, python/dbscan.rs:45).  variant 0 = dbscan
 * (src/cluster/dbscan.rs:14-83), 1 = parallel_dbscan
 * (src/cluster/parallel_dbscan.rs:16-113); labels (n int64): -1 noise, else
 * cluster ids from 0.  Replaces python/dbscan.rs:35-75 with a linear index. */
int32_t fuel_amd_dbscan(fuel_amd_index* index, double eps, int64_t min_points, int32_t variant,
                        int64_t* out_labels);

/* OPTICS over the index rows (src/cluster/optics.rs:56-133): processing order,
 * reachability and core distance (index dtype; core NaN when not a core
 * point), predecessor (-1 when none) and the DBSCAN-style labels at max_eps
 * (:345-377).  All outputs have n entries, caller-allocated.  Replaces
 * python/dbscan.rs optics_f32/_f64 with a linear index. */
int32_t fuel_amd_optics(fuel_amd_index* index, double max_eps, int64_t min_points, int64_t* out_ordering,
                        void* out_reachability, void* out_core_distance, int64_t* out_predecessor,
                        int64_t* out_labels);

/* k-means distance matrix: out[i * ncols + j] = squared distance of row i of
 * points1 (nrows x d) and row j of points2 (ncols x d), in the rounding order
 * of fuel's math::pairwise_sqdist (src/math/mod.rs:562-579): for d >= 4 one
 * sequential fused multiply-add chain per output (src/math/avx2.rs:888-963,
 * :1036-1082), for d < 4 the rounded squares summed left to right
 * (src/math/scalar.rs:10-16).  Replaces N::vec_pairwise_sqdist in the k-means
 * family (src/cluster/kmeans/lloyd.rs:76, elkan.rs:83, hamerly.rs:89,
 * selkan.rs:73, shamerly.rs:74).  Host arrays, synchronous. */
int32_t fuel_amd_pairwise_sqdist(const void* points1, int64_t nrows, const void* points2, int64_t ncols,
                                 int32_t d, int32_t dtype, void* out);
/* Same on device buffers (row pitches in elements), asynchronous on `stream`,
 * for a k-means loop that keeps points and centres resident. */
int32_t fuel_amd_pairwise_sqdist_device(const void* d_points1, int64_t nrows, int64_t ld1,
                                        const void* d_points2, int64_t ncols, int64_t ld2, int32_t d,
                                        int32_t dtype, void* d_out, int64_t ldo, void* stream);

/* Pairwise distance matrix of n rows (square n x n, or condensed n(n-1)/2 in
 * fuel's internal lower-triangular order), caller-allocated host output.
 * Replaces _compute_pairwise_distances_{f32,f64} and the _condensed_ variants
 * (python/mod.rs:73-123; src/api/square_distance_matrix.rs:12-42;
 * src/api/condensed_distance_matrix.rs:22-54). */
int32_t fuel_amd_pairwise(const void* data, int64_t n, int32_t d, int32_t dtype, int32_t metric,
                          int32_t form, void* out);
/* Same over the rows of an index (its metric and option flags), e.g. a resident
 * dataset or one built with FUEL_AMD_FLAG_FALLBACK_ARITH.  Synchronous; out is
 * a caller-allocated host array of n x n or n(n-1)/2 entries. */
int32_t fuel_amd_index_pairwise(fuel_amd_index* index, int32_t form, void* out);

/* Local outlier factor over the self-kNN of the index rows (f32 or f64 scores,
 * n entries).  Replaces local_outlier_factor over a LinearScanSearcher
 * (src/outlier/lof.rs:14-81 with for_each_knn, src/outlier/common.rs:12-49). */
int32_t fuel_amd_lof(fuel_amd_index* index, int32_t k, void* out_scores);

/* kNN outlier score (distance to the k-th neighbour, self excluded).
 * Replaces k_nearest_neighbors_outlier (src/outlier/knn.rs:14-32). */
int32_t fuel_amd_knn_outlier(fuel_amd_index* index, int32_t k, void* out_scores);

/* Statistics of the last kNN call on this index. */
int32_t fuel_amd_last_stats(const fuel_amd_index* index, fuel_amd_stats* out);
/* Query rows of the last kNN call on this index whose first-level certificate
 * failed (stats.fallback_queries of them; up to 2^20 are kept), in the call's
 * row numbering: *count receives how many are kept, the first min(cap, count)
 * are written to rows.  A diagnostic with no counterpart in fuel (its linear
 * scan has no certificate): it names the queries the certification chain
 * answered, e.g. for checking them against a CPU scan. */
int32_t fuel_amd_last_fallback_rows(const fuel_amd_index* index, int64_t* rows, int64_t cap, int64_t* count);

/* Profiling aid (no counterpart in fuel, which has no GPU): when enabled, the
 * dominant scan kernel of every kNN call on this index is bracketed by HIP
 * events recorded on the call's stream.  fuel_amd_timing_read waits for the
 * recorded events, returns the summed kernel time, the number of timed
 * launches and the kernel's name, and resets the record. */
int32_t fuel_amd_timing_enable(fuel_amd_index* index, int32_t enable);
int32_t fuel_amd_timing_read(fuel_amd_index* index, double* kernel_ms, int64_t* launches, char* name,
                             int32_t name_len);

/* ---- serving ----------------------------------------------------------------
 * A request loop over one resident index (no counterpart in fuel, whose
 * SearchIndex.knn, python/search.rs:560-569, answers one batch per call).
 * Requests are submitted from any thread and answered by worker threads, each
 * with its own stream and query workspaces.  A free worker admits every
 * request waiting at that moment (arrival order, same k) up to max_batch
 * queries and while the query-side device memory of all in-flight batches
 * stays within hbm_budget, so requests arriving while other batches run are
 * coalesced into the next launch and overlap with them on the device
 * (continuous batching).  Each query's answer is exactly what fuel_amd_knn
 * (exclude_self 0) returns for it alone.  Destroy every server of an index
 * before the index. */
typedef struct fuel_amd_server fuel_amd_server;

typedef struct fuel_amd_server_opts {
  int32_t workers;    /* batches in flight at once, one stream each (0: 2) */
  int32_t reserved0;
  int64_t max_batch;  /* queries per launch (0: 4096) */
  int64_t hbm_budget; /* bytes of query-side device memory over all in-flight batches (0: 4 GiB) */
  int64_t reserved[5];
} fuel_amd_server_opts;

typedef struct fuel_amd_serve_stats {
  int64_t requests;  /* completed since the last reset */
  int64_t queries;
  int64_t batches;   /* launches */
  int64_t max_batch_queries;
  int64_t failed;    /* requests answered with an error */
  double p50_ms;     /* request latency, submit -> results ready on the host; p50/p99 over the */
  double p99_ms;     /* last 2^20 requests (a ring), max and mean over every request since the reset */
  double max_ms;
  double mean_ms;
  double busy_s;     /* seconds with at least one batch in flight */
} fuel_amd_serve_stats;

/* The index must outlive the server (destroy the server first); other calls
 * on the index may run while the server does. */
int32_t fuel_amd_server_create(fuel_amd_index* index, const fuel_amd_server_opts* opts, fuel_amd_server** out);
/* Copies m query rows (index dtype, row-major, d per row); never blocks on the
 * GPU.  m must not exceed max_batch. */
int32_t fuel_amd_server_submit(fuel_amd_server* server, const void* query, int64_t m, int32_t k, uint64_t* ticket);
/* Blocks until the request is answered, writes int64 [m,k] indices and
 * [m,k] distances (index dtype; -1 / +inf padding as fuel_amd_knn), and
 * releases the ticket.  latency_ms and batch_queries (queries of the launch
 * that answered it) may be NULL.  One wait per ticket: a second concurrent
 * wait on the same ticket fails with FUEL_AMD_ERR_INVALID. */
int32_t fuel_amd_server_wait(fuel_amd_server* server, uint64_t ticket, int64_t* out_idx, void* out_dist,
                             double* latency_ms, int64_t* batch_queries);
int32_t fuel_amd_server_stats(fuel_amd_server* server, int32_t reset, fuel_amd_serve_stats* out);
/* Answers every submitted request, then stops the workers.  No other call on
 * the server may be in progress or follow. */
int32_t fuel_amd_server_destroy(fuel_amd_server* server);

/* ---- batched problems ---------------------------------------------------------
 * Many independent self-kNN problems of uneven size and dimension, each the
 * search loop of fuel's for_each_knn over its own dataset
 * (src/outlier/common.rs:12-49; with FUEL_AMD_EXCLUDE_NONE, SearchIndex.knn of
 * the data against itself, python/search.rs:560-569).  fuel has no batched
 * form: it runs one dataset per call.  Problems of one dimension share one
 * persistent MFMA scan launch fed by a work queue over all their (query block,
 * DB split) items; the operand and query preparation, that scan and the exact
 * rerank of every problem are captured once into a hipGraph and replayed by
 * each run, after which the few queries whose bound did not certify take the
 * certification chain.  Results are identical to fuel_amd_knn on each problem.
 * f32 only.  The caller keeps every problem's rows and outputs alive and
 * unchanged in shape while the batch exists (their contents may change). */
typedef struct fuel_amd_batch fuel_amd_batch;

typedef struct fuel_amd_problem {
  const void* data; /* device rows: n x d f32, row stride ld elements (ld % 8 == 0, 32-byte aligned) */
  int64_t n;
  int64_t ld;
  int32_t d;
  int32_t k;
  int64_t* out_idx; /* device, n x k (-1 padded) */
  void* out_dist;   /* device, n x k f32 (+inf padded) */
} fuel_amd_problem;

typedef struct fuel_amd_batch_opts {
  int32_t mode;     /* FUEL_AMD_MODE_AUTO: the autotuner picks mode, widening and split size per
                       dimension by timing them on a sample of that dimension's problems, then
                       its four leading choices on all of them;
                       FUEL_AMD_MODE_BF16 / _FP32 / _EXACT: fixed for every problem */
  int32_t widen;    /* candidate widening with a fixed mode (0: 4) */
  int32_t device;
  int32_t no_graph; /* 1: launch directly instead of replaying the captured graph */
  const char* tune_cache; /* NULL, or a text file of the autotuner's choices: read at create,
                             appended to for every group tuned, so that groups of the same
                             shape (dimension, metric, exclusion, CU count, the sampled
                             problems' k and size bucket floor(4 log2 n)) skip the timing in
                             later processes; within one process choices are kept anyway */
  int64_t reserved[3];
} fuel_amd_batch_opts;

int32_t fuel_amd_batch_create(const fuel_amd_problem* problems, int32_t count, int32_t metric, int32_t exclude_self,
                              const fuel_amd_batch_opts* opts, fuel_amd_batch** out);
/* One pass over every problem; synchronous.  Calls on one batch from several
 * threads run one after another. */
int32_t fuel_amd_batch_run(fuel_amd_batch* batch);
/* The plan (per dimension: mode, widening, split size, work items, the problems
 * that start from seeded thresholds, tuning times) and the last run's statistics
 * as JSON text (truncated to len - 1 bytes). */
int32_t fuel_amd_batch_describe(fuel_amd_batch* batch, char* text, int32_t len);
int32_t fuel_amd_batch_destroy(fuel_amd_batch* batch);

#ifdef __cplusplus
}
#endif
#endif /* FUEL_AMD_H */
