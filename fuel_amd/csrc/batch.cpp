// Batched self-kNN over many independent problems of uneven size and
// dimension (each one the search loop of fuel's for_each_knn,
// src/outlier/common.rs:12-49, over its own dataset).
//
// Plan: problems are grouped by dimension.  Every group runs with one
// configuration (MFMA kind, candidate widening, split size), chosen by the
// autotuner or fixed by the caller.  A group's problems keep their own operand
// copies and workspaces (carved from one device pool that the groups reuse in
// turn) and share one persistent MFMA scan launch whose workgroups take
// (problem, query block, DB split) items from a counter, so the small problems
// fill the gaps the large ones leave.  One pass =
//   per group: per problem {bound constants reset, operand copy, query
//   preparation, certification counter reset}, persistent scan, per problem
//   exact rerank + certificate;
// captured once into a hipGraph (fill kernels, never memset nodes) and
// replayed.  Then the certification counters come back in one copy, and the
// queries that did not certify take the chain of the single-index path (same
// kind from the certified threshold, f32, exact) through a shell index that
// borrows the problem's rows.
//
// The autotuner times every configuration (bf16 / f32 MFMA x widening 2, 4, 8
// x split size 1x, 4x the persistent slots) on a sample of each group's
// problems, without the graph, and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

namespace fa {
namespace {

constexpr int KP_LIMIT = 128;  // widest kept list of the MFMA path's first level
constexpr int KP_CHAIN = 256;  // ... and of the pass from the certified threshold
constexpr int kChainLanes = 8;  // certification chains of different problems run concurrently (8 against 4: -0.5 % per pass)
using Clock = std::chrono::steady_clock;

struct Cfg {
  int kind = 0;  // 0 bf16 MFMA, 1 f32 MFMA, 2 bf16 MFMA on the residual operand, -1 exact scan
  int widen = 4;
  int fill = 1;  // work items per persistent workgroup slot, / 4
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct ProbPlan {
  int bucket = -1;  // -1: outside the MFMA groups
  bool exact = false;
  int kind = 0, ks = 0, Kp = 0, KP = 0, KK = 0;
  int64_t n_pad = 0, nqb = 0, nq_pad = 0, rps = 0;
  int nsplit = 0;
  size_t o_xop = 0, o_hx = 0, o_consts = 0, o_qop = 0, o_q32 = 0, o_qn = 0, o_qint = 0, o_gthr = 0, o_slots = 0,
         o_lcnts = 0, o_lconsts = 0, o_qtab = 0, o_qs2 = 0;  // level-0 pool (shared by the groups in turn)
  size_t o_fl = 0, o_ft = 0;  // kept pool: certification failures of every problem
  SeedPlan seed;              // seeded thresholds of the plain operands (none when ntiles = 0)
  size_t o_seed = 0;
};

// Residual layout of one problem (plan time; the graph rebuilds the operand
// from the current rows over it).
struct ResidPart {
  DevBuf cents, rperm, tile_list, lcount, loff;
  int L = 0;
  int64_t n_rpad = 0;
};

struct Bucket {
  int d = 0;
  Cfg cfg;
  int ks = 0, V = 4, grid = 0;
  std::vector<int> members;  // largest first
  int64_t nitems = 0;
  int64_t range_start[SCAN_XCDS + 1] = {};  // item ranges of about equal work, one per XCD
  size_t pool_bytes = 0;
  DevBuf probs_dev, starts_dev;
  // grouped per-problem stages (one launch per group): resets, operand copies
  // (plain operands), query preparation, seed selection, rerank
  DevBuf group_dev;
  const FillEntry* fills = nullptr;
  int nfill = 0, sel_vmax = 4;
  bool thread_rows = false;  // operand copies and query preparation one row per thread
  GroupTab g_db{}, g_q{}, g_seed{}, g_sel{}, g_rr{};
  GroupTab g_resid{}, g_qlist{}, g_theta{}, g_sample{};  // residual problems
};

}  // namespace
}  // namespace fa

struct fuel_amd_batch {
  int device = 0, metric = 0, exclude = 0;
  bool use_graph = true;
  std::vector<fuel_amd_problem> probs;
  std::vector<char> force_exact;  // norms beyond the MFMA scan's sentinels
  std::vector<std::unique_ptr<fa::ResidPart>> parts;  // residual layouts, per problem
  std::vector<fa::ProbPlan> plan;
  std::deque<fa::Bucket> buckets;  // in place: a group owns device buffers
  hipStream_t stream = nullptr;
  fa::DevBuf pool, keep, fail_counts, counter;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Certification chain: shells that borrow one problem's rows at a time, each
  // with its own stream, so that the chains of several problems overlap.
  struct ChainLane {
    fuel_amd_index shell;
    hipStream_t stream = nullptr;
  };
  std::vector<std::unique_ptr<ChainLane>> lanes;
  std::vector<int> h_fail;
  double last_ms = 0.0, last_graph_ms = 0.0;  // whole pass; level 0 up to its certification counts
  std::chrono::steady_clock::time_point t_launch;
  int64_t last_fallback = 0, last_exact = 0, runs = 0;
  std::string tuning;  // JSON array of the autotuner's measurements
  std::mutex mu;       // one run (or describe) at a time: they share the stream and buffers
  ~fuel_amd_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    for (auto& L : lanes) {
      if (L->stream) {
        (void)hipStreamSynchronize(L->stream);
        (void)hipStreamDestroy(L->stream);
      }
      L->shell.stream = nullptr;
    }
  }
};

namespace fa {
namespace {

template <typename T>
T* at(const DevBuf& b, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(b.p) + off);
}

PrepDbArgs prep_db_args(fuel_amd_batch* b, int i);
PrepResidArgs prep_resid_args(fuel_amd_batch* b, int i);
QListArgs qlist_args(fuel_amd_batch* b, int i);
SampleSeedArgs sample_args(fuel_amd_batch* b, int i);
PrepQArgs prep_q_args(fuel_amd_batch* b, int i);
SeedArgs seed_args(fuel_amd_batch* b, int i);
RerankArgs rerank_args(fuel_amd_batch* b, int i, int V);
int32_t build_group_tables(fuel_amd_batch* b, Bucket& B);

int32_t plan_batch(fuel_amd_batch* b, const std::map<int, Cfg>& cfg_by_d) {
  const int cus = device_cu_count(b->device);
  const int np = (int)b->probs.size();
  b->plan.assign(np, ProbPlan{});
  b->buckets.clear();
  std::map<int, std::vector<int>> by_d;
  for (int i = 0; i < np; ++i) {
    const auto& pr = b->probs[i];
    auto& pp = b->plan[i];
    if (pr.n == 0 || pr.k == 0) continue;
    pp.KK = pr.k + (b->exclude != EXCL_NONE ? 1 : 0);
    auto it = cfg_by_d.find(pr.d);
    const Cfg c = it != cfg_by_d.end() ? it->second : Cfg{};
    const int ks = c.kind < 0 ? -1 : mfma_ksteps_bucket(c.kind == 2 ? 0 : c.kind, pr.d);
    if (c.kind < 0 || b->force_exact[i] || pp.KK > KP_LIMIT || ks < 0 || pr.d == 0 ||
        (c.kind == 2 && ks > mfma_resid_max_ksteps())) {
      pp.exact = true;
      continue;
    }
    by_d[pr.d].push_back(i);
  }
  size_t keep_off = 0;
  for (int i = 0; i < np; ++i) {
    const auto& pr = b->probs[i];
    b->plan[i].o_fl = keep_off;
    keep_off += align256(size_t(pr.n) * 8);
    b->plan[i].o_ft = keep_off;
    keep_off += align256(size_t(pr.n) * 4);
  }
  hipError_t e = b->keep.ensure(std::max<size_t>(keep_off, 256));
  if (e == hipSuccess) e = b->fail_counts.ensure(sizeof(int) * std::max(1, np));
  if (e == hipSuccess) e = b->counter.ensure(sizeof(uint32_t) * SCAN_XCDS * SCAN_COUNTER_STRIDE);
  if (e != hipSuccess) return hip_fail(e, "batch workspaces");
  size_t pool_max = 0;
  for (auto& kv : by_d) {
    const int bucket_id = (int)b->buckets.size();
    b->buckets.emplace_back();
    Bucket& bk = b->buckets.back();
    bk.d = kv.first;
    bk.cfg = cfg_by_d.at(kv.first);
    const int kind = bk.cfg.kind;
    const bool resid = kind == 2;
    const int mk = resid ? 0 : kind;  // MFMA operand type
    bk.ks = mfma_ksteps_bucket(mk, bk.d);
    bk.members = kv.second;
    std::sort(bk.members.begin(), bk.members.end(),
              [&](int x, int y) { return b->probs[x].n > b->probs[y].n || (b->probs[x].n == b->probs[y].n && x < y); });
    const int qpb = mfma_queries_per_block(mk, bk.ks);
    const int rows_tile = mfma_rows_per_tile(mk, bk.ks, resid);
    const int64_t slots = int64_t(cus) * mfma_workgroups_per_cu(mk, bk.ks, 4, resid);
    const int Kp = bk.ks * (mk == 0 ? 16 : 8);
    const size_t es = mk == 0 ? 2 : 4;
    double work = 0.0;
    bk.V = 4;
    for (int i : bk.members) {
      auto& pp = b->plan[i];
      const auto& pr = b->probs[i];
      pp.bucket = bucket_id;
      pp.kind = kind;
      pp.ks = bk.ks;
      pp.Kp = Kp;
      pp.KP = std::max(pp.KK, std::min(KP_LIMIT, bk.cfg.widen * pr.k));
      if (pp.KP > 64) bk.V = 8;
      pp.n_pad = std::max<int64_t>(256, (pr.n + 255) & ~int64_t(255));
      if (resid) {
        if (!b->parts[i]) {
          auto part = std::make_unique<ResidPart>();
          int32_t rc = build_partition(static_cast<const float*>(pr.data), pr.ld, pr.n, pr.d, b->stream, part->cents,
                                       part->rperm, part->tile_list, part->lcount, part->loff, part->L,
                                       part->n_rpad);
          if (rc) return rc;
          b->parts[i] = std::move(part);
        }
        pp.n_pad = b->parts[i]->n_rpad;
      }
      pp.nqb = (pr.n + qpb - 1) / qpb;
      pp.nq_pad = pp.nqb * qpb;
      work += double(pp.nqb) * double(pp.n_pad);
    }
    const int NS = 64 * bk.V;
    // split size: about 4 * fill items per persistent workgroup over the group
    const double target = std::max<double>(4.0 * rows_tile, work / (4.0 * bk.cfg.fill * double(slots)));
    size_t off = 0;
    std::vector<int64_t> starts{0};
    for (int i : bk.members) {
      auto& pp = b->plan[i];
      const auto& pr = b->probs[i];
      int64_t ns = std::max<int64_t>(1, std::min<int64_t>(1024, (int64_t)std::ceil(double(pp.n_pad) / target)));
      int64_t rps = (pp.n_pad + ns - 1) / ns;
      rps = (rps + rows_tile - 1) / rows_tile * rows_tile;
      pp.rps = rps;
      pp.nsplit = (int)((pp.n_pad + rps - 1) / rps);
      pp.o_xop = off;
      off += align256(size_t(pp.n_pad) * Kp * es);
      pp.o_hx = off;
      off += align256(size_t(pp.n_pad) * 4);
      pp.o_consts = off;
      off += 256;
      if (resid) {
        const int L = b->parts[i]->L;
        pp.o_lconsts = off;
        off += align256(size_t(L) * 64);
        pp.o_qtab = off;
        off += align256(size_t(L) * pp.nq_pad * 4);
        pp.o_qs2 = off;
        off += align256(size_t(L) * pr.n * 4);
      } else {
        pp.o_qop = off;
        off += align256(size_t(pp.nq_pad) * Kp * es);
      }
      pp.o_q32 = off;
      off += align256(size_t(pr.n) * pr.ld * 4);
      pp.o_qn = off;
      off += align256(size_t(pr.n) * 3 * 8);
      pp.o_qint = off;
      off += align256(size_t(pr.n) * 4);
      pp.o_gthr = off;
      off += align256(size_t(pp.nq_pad) * 4);
      pp.seed = plan_seed(mk, bk.ks, resid, false, 0, pp.n_pad, rows_tile, pp.KP);
      if (pp.seed.ntiles > 0) {
        pp.o_seed = off;
        off += align256(size_t(pp.nq_pad) * pp.seed.parts * 2 * SEED_GROUPS * 4);
      }
      pp.o_slots = off;
      off += align256(size_t(pp.nsplit) * pp.nq_pad * NS * 8);
      pp.o_lcnts = off;
      off += align256(size_t(pp.nsplit) * pp.nq_pad * 4);
      starts.push_back(starts.back() + pp.nqb * pp.nsplit);
    }
    bk.nitems = starts.back();
    {
      // XCD ranges: cut the item sequence where the running work (rows per item)
      // reaches x / SCAN_XCDS of the total; an item's work is its split's rows
      std::vector<double> acc(1, 0.0);
      for (int i : bk.members) acc.push_back(acc.back() + double(b->plan[i].nqb * b->plan[i].nsplit) *
                                                               double(b->plan[i].rps));
      const double total = acc.back();
      bk.range_start[0] = 0;
      size_t pi = 0;  // problem whose items hold the cut
      for (int x = 1; x < SCAN_XCDS; ++x) {
        const double target = total * x / SCAN_XCDS;
        while (pi + 1 < bk.members.size() && acc[pi + 1] < target) ++pi;
        const auto& pp = b->plan[bk.members[pi]];
        const double per = double(pp.rps);
        int64_t within = per > 0 ? (int64_t)((target - acc[pi]) / per) : 0;
        within = std::max<int64_t>(0, std::min<int64_t>(within, starts[pi + 1] - starts[pi]));
        bk.range_start[x] = std::max(bk.range_start[x - 1], std::min(bk.nitems, starts[pi] + within));
      }
      bk.range_start[SCAN_XCDS] = bk.nitems;
    }
    bk.pool_bytes = off;
    bk.grid = (int)slots;
    pool_max = std::max(pool_max, off);
    Bucket& B = bk;
    e = B.starts_dev.ensure(sizeof(int64_t) * starts.size());
    if (e == hipSuccess) e = B.probs_dev.ensure(sizeof(ScanMfmaArgs) * B.members.size());
    if (e != hipSuccess) return hip_fail(e, "batch tables");
    e = hipMemcpyAsync(B.starts_dev.p, starts.data(), sizeof(int64_t) * starts.size(), hipMemcpyHostToDevice,
                       b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "batch tables");
  }
  e = b->pool.ensure(std::max<size_t>(pool_max, 256));
  if (e != hipSuccess) return hip_fail(e, "batch pool");
  // argument tables of the persistent scans
  for (auto& B : b->buckets) {
    std::vector<ScanMfmaArgs> tab;
    for (int i : B.members) {
      const auto& pp = b->plan[i];
      const auto& pr = b->probs[i];
      ScanMfmaArgs sa{};
      sa.Xop = at<void>(b->pool, pp.o_xop);
      sa.hx = at<float>(b->pool, pp.o_hx);
      sa.Kp = pp.Kp;
      sa.n_pad = pp.n_pad;
      sa.Qop = pp.kind == 2 ? nullptr : at<void>(b->pool, pp.o_qop);
      if (pp.kind == 2) {
        const auto& part = *b->parts[i];
        sa.resid = 1;
        sa.tile_list = part.tile_list.as<int>();
        sa.cents = part.cents.as<float>();
        sa.ldc = pr.d;
        sa.Q32 = at<float>(b->pool, pp.o_q32);
        sa.ld32 = pr.ld;
        sa.m = pr.n;
        sa.d = pr.d;
        sa.qtab = at<float>(b->pool, pp.o_qtab);
      }
      sa.nq_pad = pp.nq_pad;
      sa.nqb = (int)pp.nqb;
      sa.rows_per_split = pp.rps;
      sa.KP = pp.KP;
      sa.Csub = (64 * B.V - pp.KP) / 2;
      sa.slots = at<uint64_t>(b->pool, pp.o_slots);
      sa.lcnts = at<int>(b->pool, pp.o_lcnts);
      sa.gthr = at<uint32_t>(b->pool, pp.o_gthr);
      {
        // the problem's work items must stay inside its pool ranges
        const int mk = pp.kind == 2 ? 0 : pp.kind;
        const size_t es = mk == 0 ? 2 : 4;
        ScanExtents ex{};
        ex.xop = size_t(pp.n_pad) * pp.Kp * es;
        ex.hx = size_t(pp.n_pad) * 4;
        ex.qop = pp.kind == 2 ? 0 : size_t(pp.nq_pad) * pp.Kp * es;
        ex.slots = size_t(pp.nsplit) * pp.nq_pad * 64 * B.V * 8;
        ex.lcnts = size_t(pp.nsplit) * pp.nq_pad * 4;
        ex.gthr = size_t(pp.nq_pad) * 4;
        ex.q32 = size_t(pr.n) * pr.ld * 4;
        int L = 0;
        if (pp.kind == 2) {
          const auto& part = *b->parts[i];
          L = part.L;
          ex.tile_list = part.tile_list.bytes;
          ex.cents = part.cents.bytes;
          ex.qtab = size_t(L) * pp.nq_pad * 4;
        }
        std::string why;
        if (scan_check_launch(sa, mk, pp.ks, B.V, pp.nsplit, ex, L, &why))
          return fail(FUEL_AMD_ERR_DEVICE, "internal error: batch scan plan outside its buffers: " + why);
      }
      tab.push_back(sa);
    }
    e = hipMemcpyAsync(B.probs_dev.p, tab.data(), sizeof(ScanMfmaArgs) * tab.size(), hipMemcpyHostToDevice,
                       b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "batch tables");
    const int32_t rc = build_group_tables(b, B);
    if (rc) return rc;
  }
  b->h_fail.assign(np, 0);
  return FUEL_AMD_OK;
}

// Argument tables of a group's grouped stages, in one device buffer: fill
// entries, then per stage the problems' arguments and their block starts.
int32_t build_group_tables(fuel_amd_batch* b, Bucket& B) {
  std::vector<FillEntry> fills;
  std::vector<PrepDbArgs> dbs;
  std::vector<PrepQArgs> qs;
  std::vector<SeedArgs> sels;
  std::vector<RerankArgs> rrs;
  std::vector<PrepResidArgs> rps;
  std::vector<QListArgs> qls;
  std::vector<SampleSeedArgs> sms;
  std::vector<int64_t> db_start{0}, q_start{0}, seed_start{0}, sel_start{0}, rr_start{0};
  std::vector<int64_t> rp_start{0}, ql_start{0}, th_start{0}, sm_start{0};
  int vmax = 1;
  // rows of at most 64 operand columns: one row per thread (the batch's rows
  // are 32-byte aligned with pitches a multiple of 8, and so are its copies)
  B.thread_rows = !B.members.empty() && b->plan[B.members[0]].Kp <= PREP_THREAD_ROWS_MAX_KP;
  const int64_t db_rows = B.thread_rows ? 256 : 64, q_rows = B.thread_rows ? 256 : 4;
  for (int i : B.members) {
    const auto& pp = b->plan[i];
    const auto& pr = b->probs[i];
    const PrepDbArgs pa = prep_db_args(b, i);
    if (pp.kind != 2) {
      fills.push_back({reinterpret_cast<uint32_t*>(pa.consts), 16, 0u, 0u});
      fills.push_back({reinterpret_cast<uint32_t*>(pa.flags), 2, 1u, 0u});
      dbs.push_back(pa);
      db_start.push_back(db_start.back() + (pa.n_pad + db_rows - 1) / db_rows);
    } else {
      fills.push_back({reinterpret_cast<uint32_t*>(pa.consts), 32, 0u, 0u});  // constants, zero flags
      fills.push_back({at<uint32_t>(b->pool, pp.o_lconsts), int64_t(16) * b->parts[i]->L, 0u, 0u});
      const PrepResidArgs ra = prep_resid_args(b, i);
      if (ra.n_rpad % 64 != 0) return fail(FUEL_AMD_ERR_DEVICE, "internal error: residual rows not in whole blocks");
      rps.push_back(ra);
      rp_start.push_back(rp_start.back() + ra.n_rpad / 64);
      const QListArgs la = qlist_args(b, i);
      qls.push_back(la);
      ql_start.push_back(ql_start.back() + (int64_t(la.L) * la.nq_pad + 255) / 256);
      th_start.push_back(th_start.back() + (la.nq_pad + 3) / 4);
      const SampleSeedArgs sa = sample_args(b, i);
      if (sa.KP < 1 || sa.KP > SEED_SAMPLE_ROWS || (sa.ldx & 3) || (sa.ld32 & 3))
        return fail(FUEL_AMD_ERR_DEVICE, "internal error: sampled seed outside its limits");
      sms.push_back(sa);
      sm_start.push_back(sm_start.back() + pr.n);
    }
    fills.push_back({reinterpret_cast<uint32_t*>(b->fail_counts.as<int>() + i), 1, 0u, 0u});
    const PrepQArgs qa = prep_q_args(b, i);
    qs.push_back(qa);
    q_start.push_back(q_start.back() + (qa.m_pad + q_rows - 1) / q_rows);
    if (pp.seed.ntiles > 0) {
      const SeedArgs sd = seed_args(b, i);
      const int total = sd.parts * 2 * SEED_GROUPS;
      if (total > 1024 || total < sd.KP || sd.stride * (sd.ntiles - 1) + mfma_rows_per_tile(pp.kind, pp.ks, false) >
                                               pp.n_pad)
        return fail(FUEL_AMD_ERR_DEVICE, "internal error: seed plan outside its buffers");
      sels.push_back(sd);
      seed_start.push_back(seed_start.back() + int64_t(sd.nqb) * sd.parts);
      sel_start.push_back(sel_start.back() + pr.n);
      vmax = std::max(vmax, (total + 63) / 64);
    }
    rrs.push_back(rerank_args(b, i, B.V));
    rr_start.push_back(rr_start.back() + pr.n);
  }
  fills.push_back({b->counter.as<uint32_t>(), int64_t(SCAN_XCDS) * SCAN_COUNTER_STRIDE, 0u, 0u});
  if (vmax > 16) return fail(FUEL_AMD_ERR_DEVICE, "internal error: seed minima beyond the selection's registers");
  std::vector<char> host;
  auto put = [&](const void* src, size_t bytes) {
    const size_t off = align256(host.size());
    host.resize(off + bytes);
    if (bytes) std::memcpy(host.data() + off, src, bytes);
    return off;
  };
  const size_t o_fill = put(fills.data(), fills.size() * sizeof(FillEntry));
  const size_t o_db = put(dbs.data(), dbs.size() * sizeof(PrepDbArgs));
  const size_t o_dbs = put(db_start.data(), db_start.size() * 8);
  const size_t o_q = put(qs.data(), qs.size() * sizeof(PrepQArgs));
  const size_t o_qs = put(q_start.data(), q_start.size() * 8);
  const size_t o_sel = put(sels.data(), sels.size() * sizeof(SeedArgs));
  const size_t o_seeds = put(seed_start.data(), seed_start.size() * 8);
  const size_t o_sels = put(sel_start.data(), sel_start.size() * 8);
  const size_t o_rp = put(rps.data(), rps.size() * sizeof(PrepResidArgs));
  const size_t o_rps = put(rp_start.data(), rp_start.size() * 8);
  const size_t o_ql = put(qls.data(), qls.size() * sizeof(QListArgs));
  const size_t o_qls = put(ql_start.data(), ql_start.size() * 8);
  const size_t o_ths = put(th_start.data(), th_start.size() * 8);
  const size_t o_sm = put(sms.data(), sms.size() * sizeof(SampleSeedArgs));
  const size_t o_sms = put(sm_start.data(), sm_start.size() * 8);
  const size_t o_rr = put(rrs.data(), rrs.size() * sizeof(RerankArgs));
  const size_t o_rrs = put(rr_start.data(), rr_start.size() * 8);
  hipError_t e = B.group_dev.ensure(std::max<size_t>(host.size(), 256));
  if (e == hipSuccess)
    e = hipMemcpyAsync(B.group_dev.p, host.data(), host.size(), hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "batch group tables");
  auto tab = [&](size_t o_tab, size_t o_start, size_t np, const std::vector<int64_t>& start) {
    GroupTab g{};
    g.tab = at<void>(B.group_dev, o_tab);
    g.start = at<int64_t>(B.group_dev, o_start);
    g.np = (int)np;
    g.nblocks = start.back();
    return g;
  };
  B.fills = at<FillEntry>(B.group_dev, o_fill);
  B.nfill = (int)fills.size();
  B.g_db = tab(o_db, o_dbs, dbs.size(), db_start);
  B.g_q = tab(o_q, o_qs, qs.size(), q_start);
  B.g_seed = tab(o_sel, o_seeds, sels.size(), seed_start);
  B.g_sel = tab(o_sel, o_sels, sels.size(), sel_start);
  B.g_resid = tab(o_rp, o_rps, rps.size(), rp_start);
  B.g_qlist = tab(o_ql, o_qls, qls.size(), ql_start);
  B.g_theta = tab(o_ql, o_ths, qls.size(), th_start);
  B.g_sample = tab(o_sm, o_sms, sms.size(), sm_start);
  B.g_rr = tab(o_rr, o_rrs, rrs.size(), rr_start);
  B.sel_vmax = vmax;
  return FUEL_AMD_OK;
}

PrepDbArgs prep_db_args(fuel_amd_batch* b, int i) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  PrepDbArgs pa{};
  pa.X = static_cast<const float*>(pr.data);
  pa.ldx = pr.ld;
  pa.n = pr.n;
  pa.n_pad = pp.n_pad;
  pa.d = pr.d;
  pa.Kp = pp.Kp;
  pa.bf16 = pp.kind == 0;
  pa.Xop = at<void>(b->pool, pp.o_xop);
  pa.hx = at<float>(b->pool, pp.o_hx);
  pa.consts = at<unsigned long long>(b->pool, pp.o_consts);
  pa.flags = at<int>(b->pool, pp.o_consts + 64);
  return pa;
}

// Operand copy of problem i (and its bound constants, reset first).
PrepResidArgs prep_resid_args(fuel_amd_batch* b, int i) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  const auto& part = *b->parts[i];
  PrepResidArgs r{};
  r.X = static_cast<const float*>(pr.data);
  r.ldx = pr.ld;
  r.d = pr.d;
  r.Kp = pp.Kp;
  r.n_rpad = part.n_rpad;
  r.rperm = part.rperm.as<uint32_t>();
  r.tile_list = part.tile_list.as<int>();
  r.cents = part.cents.as<float>();
  r.ldc = pr.d;
  r.Xop = at<uint16_t>(b->pool, pp.o_xop);
  r.hx = at<float>(b->pool, pp.o_hx);
  r.lconsts = at<unsigned long long>(b->pool, pp.o_lconsts);
  r.gconsts = at<unsigned long long>(b->pool, pp.o_consts);
  return r;
}

// (reset: the bound constants are reset here first; the grouped pass resets
// them in its fill launch)
hipError_t enqueue_operand(fuel_amd_batch* b, int i, hipStream_t st, bool reset) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  const PrepDbArgs pa = prep_db_args(b, i);
  if (pp.kind != 2) {
    hipError_t e = hipSuccess;
    if (reset) e = launch_fill_u32(reinterpret_cast<uint32_t*>(pa.consts), 16, 0u, st);
    if (reset && e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(pa.flags), 2, 1u, st);
    if (e == hipSuccess) e = launch_prep_db(pa, st);
    return e;
  }
  const auto& part = *b->parts[i];
  hipError_t e = hipSuccess;
  if (reset) e = launch_fill_u32(reinterpret_cast<uint32_t*>(pa.consts), 32, 0u, st);  // constants, zero flags
  if (reset && e == hipSuccess) e = launch_fill_u32(at<uint32_t>(b->pool, pp.o_lconsts), 16 * part.L, 0u, st);
  const PrepResidArgs r = prep_resid_args(b, i);
  if (e == hipSuccess) e = launch_prep_resid(r, st);
  return e;
}

PrepQArgs prep_q_args(fuel_amd_batch* b, int i) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  PrepQArgs qa{};
  qa.Q = static_cast<const float*>(pr.data);
  qa.ldq = pr.ld;
  qa.m = pr.n;
  qa.m_pad = pp.nq_pad;
  qa.d = pr.d;
  qa.Kp = pp.Kp;
  qa.bf16 = pp.kind != 1;
  qa.Qop = pp.kind == 2 ? nullptr : at<void>(b->pool, pp.o_qop);
  qa.Q32 = at<float>(b->pool, pp.o_q32);
  qa.ld32 = pr.ld;
  qa.qn2 = at<double>(b->pool, pp.o_qn);
  qa.eq2 = qa.qn2 + pr.n;
  qa.qt2 = qa.eq2 + pr.n;
  qa.qint = at<int>(b->pool, pp.o_qint);
  qa.gthr = at<uint32_t>(b->pool, pp.o_gthr);
  return qa;
}

SeedArgs seed_args(fuel_amd_batch* b, int i) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  SeedArgs sd{};
  sd.Xop = at<void>(b->pool, pp.o_xop);
  sd.hx = at<float>(b->pool, pp.o_hx);
  sd.Kp = pp.Kp;
  sd.Qop = at<void>(b->pool, pp.o_qop);
  sd.nq_pad = pp.nq_pad;
  sd.nqb = (int)pp.nqb;
  sd.stride = pp.seed.stride;
  sd.ntiles = pp.seed.ntiles;
  sd.parts = pp.seed.parts;
  sd.mins = at<float>(b->pool, pp.o_seed);
  sd.m = pr.n;
  sd.KP = pp.KP;
  sd.gthr = at<uint32_t>(b->pool, pp.o_gthr);
  return sd;
}

RerankArgs rerank_args(fuel_amd_batch* b, int i, int V) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  RerankArgs ra{};
  ra.slots = at<uint64_t>(b->pool, pp.o_slots);
  ra.lcnts = at<int>(b->pool, pp.o_lcnts);
  ra.nsplit = pp.nsplit;
  ra.nq = pr.n;
  ra.nq_pad = pp.nq_pad;
  ra.slot_ns = 64 * V;
  ra.KP = pp.KP;
  ra.X = static_cast<const float*>(pr.data);
  ra.ldx = pr.ld;
  ra.Q32 = at<float>(b->pool, pp.o_q32);
  ra.ld32 = pr.ld;
  ra.d = pr.d;
  ra.metric = b->metric;
  ra.arith = ARITH_AVX;
  ra.bf16 = pp.kind != 1;
  ra.resid = pp.kind == 2 ? 1 : 0;
  ra.rperm = pp.kind == 2 ? b->parts[i]->rperm.as<uint32_t>() : nullptr;
  ra.qn2 = at<double>(b->pool, pp.o_qn);
  ra.eq2 = ra.qn2 + pr.n;
  ra.qt2 = ra.eq2 + pr.n;
  ra.qint = at<int>(b->pool, pp.o_qint);
  ra.consts = at<unsigned long long>(b->pool, pp.o_consts);
  ra.flags = at<int>(b->pool, pp.o_consts + 64);
  ra.KK = pp.KK;
  ra.k = pr.k;
  ra.mode = b->exclude;
  ra.qbase = 0;
  ra.rows = nullptr;
  ra.out_idx = pr.out_idx;
  ra.out_dist = static_cast<float*>(pr.out_dist);
  ra.fail_count = b->fail_counts.as<int>() + i;
  ra.fail_list = at<int64_t>(b->keep, pp.o_fl);
  ra.gthr = at<uint32_t>(b->pool, pp.o_gthr);
  ra.fail_theta = at<float>(b->keep, pp.o_ft);
  return ra;
}

// Residual problem i: query-to-centroid norms (per problem, after the query
// preparation), then through the grouped launches the per-(query, list) terms
// and proven start, and the sampled threshold.
KmeansPairwiseArgs<float> centroid_args(fuel_amd_batch* b, int i) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  const auto& part = *b->parts[i];
  KmeansPairwiseArgs<float> pw{};
  pw.P1 = at<float>(b->pool, pp.o_q32);
  pw.ld1 = pr.ld;
  pw.m = pr.n;
  pw.P2 = part.cents.as<float>();
  pw.ld2 = pr.d;
  pw.n = part.L;
  pw.d = pr.d;
  pw.out = at<float>(b->pool, pp.o_qs2);
  pw.ldo = part.L;
  return pw;
}

QListArgs qlist_args(fuel_amd_batch* b, int i) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  const auto& part = *b->parts[i];
  QListArgs la{};
  la.s2 = at<float>(b->pool, pp.o_qs2);
  la.m = pr.n;
  la.nq_pad = pp.nq_pad;
  la.d = pr.d;
  la.L = part.L;
  la.lconsts = at<unsigned long long>(b->pool, pp.o_lconsts);
  la.qtab = at<float>(b->pool, pp.o_qtab);
  la.lcount = part.lcount.as<uint32_t>();
  la.KP = pp.KP;
  la.gthr = at<uint32_t>(b->pool, pp.o_gthr);
  return la;
}

// the sampled threshold: exact distances to the nearest list's first rows
SampleSeedArgs sample_args(fuel_amd_batch* b, int i) {
  const auto& pp = b->plan[i];
  const auto& pr = b->probs[i];
  const auto& part = *b->parts[i];
  SampleSeedArgs ss{};
  ss.s2 = at<float>(b->pool, pp.o_qs2);
  ss.m = pr.n;
  ss.L = part.L;
  ss.KP = pp.KP;
  ss.lcount = part.lcount.as<uint32_t>();
  ss.loff = part.loff.as<int64_t>();
  ss.rperm = part.rperm.as<uint32_t>();
  ss.X = static_cast<const float*>(pr.data);
  ss.ldx = pr.ld;
  ss.Q32 = at<float>(b->pool, pp.o_q32);
  ss.ld32 = pr.ld;
  ss.d = pr.d;
  ss.gthr = at<uint32_t>(b->pool, pp.o_gthr);
  return ss;
}

// Level 0 of one group on stream st (no host synchronisation: capturable).
// Stage by stage over the group's problems: resets (bound constants, flags,
// residual list constants, certification counters and the scan's work
// counters: one fill launch), operand copies and query preparation (one
// launch each; residual operands per problem), seed passes (per problem) and
// their selection (one launch), the persistent scan, the reranks (one launch).
hipError_t enqueue_bucket(fuel_amd_batch* b, const Bucket& B, hipStream_t st) {
  const int mk = B.cfg.kind == 2 ? 0 : B.cfg.kind;
  hipError_t e = launch_fill_multi(B.fills, B.nfill, st);
  if (e == hipSuccess) e = launch_prep_db_group(B.g_db, B.thread_rows, st);
  if (e == hipSuccess) e = launch_prep_resid_group(B.g_resid, st);
  if (e == hipSuccess) e = launch_prep_queries_group(B.g_q, B.thread_rows, st);
  if (e == hipSuccess) e = launch_seed_mfma_group(B.g_seed, mk, B.ks, st);
  if (e == hipSuccess) e = launch_seed_select_group(B.g_sel, B.sel_vmax, st);
  for (int i : B.members)
    if (e == hipSuccess && b->plan[i].kind == 2) e = launch_kmeans_pairwise<float>(centroid_args(b, i), st);
  if (e == hipSuccess) e = launch_prep_qlist_group(B.g_qlist, B.g_theta, st);
  if (e == hipSuccess) e = launch_seed_sample_group(B.g_sample, st);
  if (e != hipSuccess) return e;
  ScanBatchArgs sb{};
  sb.probs = B.probs_dev.as<ScanMfmaArgs>();
  sb.item_start = B.starts_dev.as<int64_t>();
  sb.nprob = (int)B.members.size();
  sb.nitems = B.nitems;
  sb.counter = b->counter.as<unsigned int>();
  for (int x = 0; x <= SCAN_XCDS; ++x) sb.range_start[x] = B.range_start[x];
  if (e == hipSuccess)
    e = launch_scan_mfma_batch(sb, B.cfg.kind == 2 ? 0 : B.cfg.kind, B.ks, B.V, B.grid, B.cfg.kind == 2, st);
  if (e == hipSuccess) e = launch_rerank_group(B.g_rr, B.V, ARITH_AVX, st);
  return e;
}

hipError_t enqueue_all(fuel_amd_batch* b, hipStream_t st) {
  for (const auto& B : b->buckets) {
    hipError_t e = enqueue_bucket(b, B, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void point_shell(fuel_amd_batch* b, fuel_amd_index& s, hipStream_t st, int i, int mode) {
  const auto& pr = b->probs[i];
  s.device = b->device;
  s.dtype = FUEL_AMD_F32;
  s.metric = b->metric;
  s.n = pr.n;
  s.d = pr.d;
  s.ld = pr.ld;
  s.opts.mode = mode;
  s.opts.widen = 4;
  s.opts.device = b->device;
  s.X.borrow(pr.data, size_t(pr.n) * pr.ld * 4);
  s.mop[0].ready.store(false);
  s.mop[1].ready.store(false);
  s.stream = st;
}

// Norms beyond the MFMA scan's finite sentinels: such problems take the exact
// scan (found once, at planning, like the single-index path's operand check).
int32_t drop_unusable(fuel_amd_batch* b, bool& changed) {
  changed = false;
  std::vector<unsigned long long> x2(b->probs.size(), 0);
  for (const auto& B : b->buckets) {
    for (int i : B.members) {
      const PrepDbArgs pa = prep_db_args(b, i);
      hipError_t e = enqueue_operand(b, i, b->stream, true);
      if (e == hipSuccess)
        e = hipMemcpyAsync(&x2[i], pa.consts, sizeof(unsigned long long), hipMemcpyDeviceToHost, b->stream);
      if (e != hipSuccess) return hip_fail(e, "batch operand check");
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "batch operand check");
  }
  for (const auto& B : b->buckets)
    for (int i : B.members) {
      double v;
      std::memcpy(&v, &x2[i], sizeof(v));
      if (!(v < MFMA_NORM2_LIMIT)) {
        b->force_exact[i] = 1;
        changed = true;
      }
    }
  return FUEL_AMD_OK;
}

int32_t finish(fuel_amd_batch* b) {
  const int np = (int)b->probs.size();
  hipError_t e = hipSuccess;
  if (!b->buckets.empty()) {
    e = hipMemcpyAsync(b->h_fail.data(), b->fail_counts.p, sizeof(int) * np, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "certification counts");
  }
  b->last_graph_ms = std::chrono::duration<double, std::milli>(Clock::now() - b->t_launch).count();
  if (np > 0 && cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
  std::vector<int> work;  // problems with queries left for the chain, or on the exact scan
  for (int i = 0; i < np; ++i)
    if ((b->plan[i].bucket >= 0 && b->h_fail[i] > 0) || (b->plan[i].bucket < 0 && b->plan[i].exact)) work.push_back(i);
  const int nl = std::min<int>((int)work.size(), kChainLanes);
  while ((int)b->lanes.size() < nl) {
    auto L = std::make_unique<fuel_amd_batch::ChainLane>();
    e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "chain stream");
    b->lanes.push_back(std::move(L));
  }
  std::atomic<int> next{0};
  std::atomic<int64_t> fallback{0}, exact{0};
  std::vector<int32_t> lane_rc(nl, FUEL_AMD_OK);
  std::vector<std::string> lane_err(nl);
  // one lane: problems taken in turn from `work`, each through its whole chain
  // on the lane's stream (the problems' outputs and fail lists are disjoint)
  auto lane_main = [&](int l) {
    (void)hipSetDevice(b->device);
    auto& shell = b->lanes[l]->shell;
    hipStream_t st = b->lanes[l]->stream;
    int32_t rc = FUEL_AMD_OK;
    for (;;) {
      const int w = next.fetch_add(1);
      if (w >= (int)work.size()) break;
      if (cancelled()) {
        rc = fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
        break;
      }
      const int i = work[w];
      const auto& pp = b->plan[i];
      const auto& pr = b->probs[i];
      if (pp.bucket >= 0) {
        fallback += b->h_fail[i];
        // bf16 / f32: the same kind from the certified threshold, then the chain;
        // residual: the f32 form (its threshold is in other key units), then exact
        point_shell(b, shell, st, i, pp.kind == 0 ? FUEL_AMD_MODE_BF16 : FUEL_AMD_MODE_FP32);
        const int64_t before = shell.ctx.stats.exact_queries;
        const bool resid = pp.kind == 2;
        if (resid && !mfma_supported(&shell, 1, pr.k, b->exclude)) {
          rc = knn_exact_rows(&shell, &shell.ctx, static_cast<const float*>(pr.data), pr.ld,
                              at<int64_t>(b->keep, pp.o_fl), b->h_fail[i], 0, pr.k, b->exclude, pr.out_idx,
                              static_cast<float*>(pr.out_dist), st);
          shell.ctx.stats.exact_queries += b->h_fail[i];
        } else {
          rc = knn_mfma(&shell, &shell.ctx, static_cast<const float*>(pr.data), b->h_fail[i], pr.ld, pr.k,
                        b->exclude, resid ? 1 : pp.kind, pr.out_idx, static_cast<float*>(pr.out_dist), st,
                        at<int64_t>(b->keep, pp.o_fl), 0, resid ? 2 : 1, resid ? KP_LIMIT : KP_CHAIN,
                        resid ? nullptr : at<float>(b->keep, pp.o_ft));
        }
        exact += shell.ctx.stats.exact_queries - before;
      } else {
        point_shell(b, shell, st, i, FUEL_AMD_MODE_EXACT);
        rc = knn_run(&shell, &shell.ctx, pr.data, pr.n, pr.ld, pr.k, b->exclude, pr.out_idx, pr.out_dist, st);
        exact += pr.n;
      }
      if (rc) break;
    }
    const hipError_t se = hipStreamSynchronize(st);
    if (rc == FUEL_AMD_OK && se != hipSuccess) rc = hip_fail(se, "batch fallback");
    lane_rc[l] = rc;
    if (rc) lane_err[l] = fuel_amd_last_error();  // the message is per thread
  };
  std::vector<std::thread> threads;
  for (int l = 1; l < nl; ++l) threads.emplace_back(lane_main, l);
  if (nl > 0) lane_main(0);
  for (auto& t : threads) t.join();
  for (int l = 0; l < nl; ++l)
    if (lane_rc[l]) return fail(lane_rc[l], lane_err[l]);
  b->last_fallback = fallback.load();
  b->last_exact = exact.load();
  return FUEL_AMD_OK;
}

int32_t run_once(fuel_amd_batch* b) {
  const auto t0 = Clock::now();
  b->t_launch = t0;
  hipError_t e;
  if (b->use_graph && b->exec)
    e = hipGraphLaunch(b->exec, b->stream);
  else
    e = enqueue_all(b, b->stream);
  if (e != hipSuccess) return hip_fail(e, "batch launch");
  int32_t rc = finish(b);
  b->last_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  b->runs += 1;
  return rc;
}

int32_t make_batch(const fuel_amd_problem* probs, const std::vector<int>& which, int metric, int exclude,
                   int device, std::map<int, Cfg> cfg_by_d, bool use_graph, std::unique_ptr<fuel_amd_batch>& out) {
  auto b = std::make_unique<fuel_amd_batch>();
  b->device = device;
  b->metric = metric;
  b->exclude = exclude;
  b->use_graph = use_graph;
  for (int i : which) b->probs.push_back(probs[i]);
  b->force_exact.assign(b->probs.size(), 0);
  b->parts.resize(b->probs.size());
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "batch stream");
  int32_t rc = plan_batch(b.get(), cfg_by_d);
  if (rc) return rc;
  bool changed = false;
  rc = drop_unusable(b.get(), changed);
  if (rc) return rc;
  if (changed) {
    rc = plan_batch(b.get(), cfg_by_d);  // the flagged problems leave their groups
    if (rc) return rc;
  }
  if (use_graph && !b->buckets.empty()) {
    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "graph capture");
    hipError_t le = enqueue_all(b.get(), b->stream);
    e = hipStreamEndCapture(b->stream, &b->graph);
    if (le != hipSuccess) return hip_fail(le, "graph capture");
    if (e != hipSuccess) return hip_fail(e, "graph capture");
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "graph instantiate");
  }
  out = std::move(b);
  return FUEL_AMD_OK;
}

const char* kind_name(int kind) {
  return kind == 0 ? "bf16" : (kind == 1 ? "fp32" : (kind == 2 ? "bf16r" : "exact"));
}

// Plan cache of the autotuner.  A group's key is its dimension, metric,
// exclusion, the device's CU count and the (size bucket, k) of every sampled
// problem, the size bucket being floor(4 log2 n): groups of like shape share a
// choice.  Entries live for the process and, when the caller names a file, in
// that file (one "key<TAB>kind widen fill" line each, appended as groups are
// tuned), so that later processes skip the timing.
std::mutex g_tune_mu;
std::map<std::string, Cfg> g_tune_cache;
std::string g_tune_loaded;  // the file already read into the cache

std::string tune_key(int d, int metric, int exclude, int device, const fuel_amd_problem* probs,
                     const std::vector<int>& sample) {
  std::string key = "d" + std::to_string(d) + "/m" + std::to_string(metric) + "/x" + std::to_string(exclude) + "/cu" +
                    std::to_string(device_cu_count(device)) + "/";
  for (int i : sample) {
    const int bucket = (int)std::floor(4.0 * std::log2((double)std::max<int64_t>(1, probs[i].n)));
    key += std::to_string(bucket) + ":" + std::to_string(probs[i].k) + ",";
  }
  return key;
}

void tune_cache_load(const char* path) {
  if (!path || !*path || g_tune_loaded == path) return;
  g_tune_loaded = path;
  FILE* f = std::fopen(path, "r");
  if (!f) return;  // no file yet: it is written as groups are tuned
  char line[4096];
  while (std::fgets(line, sizeof(line), f)) {
    char* tab = std::strchr(line, '\t');
    if (!tab) continue;
    *tab = 0;
    Cfg c;
    if (std::sscanf(tab + 1, "%d %d %d", &c.kind, &c.widen, &c.fill) == 3 && c.kind >= -1 && c.kind <= 2 &&
        (c.widen == 2 || c.widen == 4 || c.widen == 8) && (c.fill == 1 || c.fill == 4))
      g_tune_cache[line] = c;
  }
  std::fclose(f);
}

void tune_cache_store(const char* path, const std::string& key, const Cfg& c) {
  g_tune_cache[key] = c;
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;  // an unwritable cache file only costs the next process its timing
  std::fprintf(f, "%s\t%d %d %d\n", key.c_str(), c.kind, c.widen, c.fill);
  std::fclose(f);
}

// Times every configuration on a sample of each dimension's problems (or
// takes the cached choice for a group of the same shape).
int32_t autotune(const fuel_amd_problem* probs, int count, int metric, int exclude, int device,
                 const char* cache_path, std::map<int, Cfg>& best, std::string& log) {
  std::map<int, std::vector<int>> by_d;
  for (int i = 0; i < count; ++i)
    if (probs[i].n > 0 && probs[i].k > 0) by_d[probs[i].d].push_back(i);
  std::lock_guard<std::mutex> cache_lock(g_tune_mu);
  tune_cache_load(cache_path);
  log = "[";
  bool first = true;
  for (auto& kv : by_d) {
    const int d = kv.first;
    auto members = kv.second;
    std::sort(members.begin(), members.end(), [&](int x, int y) { return probs[x].n > probs[y].n; });
    std::vector<int> sample;
    const size_t step = std::max<size_t>(1, members.size() / 12);
    for (size_t j = 0; j < members.size(); j += step) sample.push_back(members[j]);
    const std::string key = tune_key(d, metric, exclude, device, probs, sample);
    auto hit = g_tune_cache.find(key);
    if (hit != g_tune_cache.end()) {
      best[d] = hit->second;
      char buf[200];
      std::snprintf(buf, sizeof(buf), "%s{\"d\": %d, \"mode\": \"%s\", \"widen\": %d, \"fill\": %d, \"cached\": true}",
                    first ? "" : ", ", d, kind_name(hit->second.kind), hit->second.widen, hit->second.fill);
      log += buf;
      first = false;
      continue;
    }
    struct Trial {
      Cfg c;
      double ms;
      int64_t fallback;
    };
    std::vector<Trial> trials;
    for (int kind : {0, 2, 1}) {
      const int ks = mfma_ksteps_bucket(kind == 2 ? 0 : kind, d);
      if (ks < 0 || (kind == 2 && ks > mfma_resid_max_ksteps())) continue;
      for (int widen : {2, 4, 8}) {
        for (int fill : {1, 4}) {
          Cfg c{kind, widen, fill};
          std::unique_ptr<fuel_amd_batch> sb;
          int32_t rc = make_batch(probs, sample, metric, exclude, device, {{d, c}}, false, sb);
          if (rc) return rc;
          rc = run_once(sb.get());  // warm-up: first-touch allocations of the fallback path
          if (rc) return rc;
          // the faster of two timed passes: one pass of a dozen problems is short
          // enough for clock and scheduling noise to reorder close configurations
          double best_pass = 1e300;
          for (int rep = 0; rep < 2; ++rep) {
            rc = run_once(sb.get());
            if (rc) return rc;
            best_pass = std::min(best_pass, sb->last_ms);
          }
          // the time as logged (3 decimals), so that the logged table shows the choice
          char msbuf[32];
          std::snprintf(msbuf, sizeof(msbuf), "%.3f", best_pass);
          const double ms = std::strtod(msbuf, nullptr);
          char buf[320];
          std::snprintf(buf, sizeof(buf),
                        "%s{\"d\": %d, \"mode\": \"%s\", \"widen\": %d, \"fill\": %d, \"sample\": %zu, "
                        "\"ms\": %.3f, \"fallback\": %lld}",
                        first ? "" : ", ", d, kind_name(kind), widen, fill, sample.size(), ms,
                        (long long)sb->last_fallback);
          log += buf;
          first = false;
          trials.push_back(Trial{c, ms, sb->last_fallback});
        }
      }
    }
    // the fastest; among those within 5 % of it, the one with the fewest
    // certification fallbacks (each failing problem costs a host-synchronised
    // chain, and the whole group has many more problems than the sample)
    double fastest = 1e300;
    for (const auto& t : trials) fastest = std::min(fastest, t.ms);
    Cfg best_cfg{-1, 4, 1};
    double best_ms = 1e300;
    int64_t best_fb = INT64_MAX;
    for (const auto& t : trials) {
      if (t.ms > 1.05 * fastest) continue;
      if (t.fallback < best_fb || (t.fallback == best_fb && t.ms < best_ms)) {
        best_fb = t.fallback;
        best_ms = t.ms;
        best_cfg = t.c;
      }
    }
    // A dozen problems do not fill the GPU at small dimensions, so the sample
    // ranks the configurations only roughly: that choice and the three fastest
    // on the sample are timed again on the whole group (the faster of two
    // passes, eager), and the fastest whole pass is kept.
    if (members.size() > sample.size()) {
      std::vector<Trial> order = trials;
      std::stable_sort(order.begin(), order.end(), [](const Trial& x, const Trial& y) { return x.ms < y.ms; });
      std::vector<Cfg> finalists{best_cfg};
      for (const auto& t : order) {
        if (finalists.size() >= 4) break;
        bool seen = false;
        for (const auto& f : finalists)
          seen = seen || (f.kind == t.c.kind && f.widen == t.c.widen && f.fill == t.c.fill);
        if (!seen) finalists.push_back(t.c);
      }
      double best_group = 1e300;
      for (const Cfg& c : finalists) {
        std::unique_ptr<fuel_amd_batch> gb;
        int32_t rc = make_batch(probs, members, metric, exclude, device, {{d, c}}, false, gb);
        if (rc) return rc;
        rc = run_once(gb.get());
        if (rc) return rc;
        double pass = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
          rc = run_once(gb.get());
          if (rc) return rc;
          pass = std::min(pass, gb->last_ms);
        }
        char buf[320];
        std::snprintf(buf, sizeof(buf),
                      ", {\"d\": %d, \"mode\": \"%s\", \"widen\": %d, \"fill\": %d, \"group\": %zu, "
                      "\"ms\": %.3f, \"fallback\": %lld}",
                      d, kind_name(c.kind), c.widen, c.fill, members.size(), pass, (long long)gb->last_fallback);
        log += buf;
        if (pass < best_group) {
          best_group = pass;
          best_cfg = c;
        }
      }
    }
    best[d] = best_cfg;
    tune_cache_store(cache_path, key, best_cfg);
  }
  log += "]";
  return FUEL_AMD_OK;
}

}  // namespace
}  // namespace fa

using namespace fa;

extern "C" {

int32_t fuel_amd_batch_create(const fuel_amd_problem* problems, int32_t count, int32_t metric, int32_t exclude_self,
                              const fuel_amd_batch_opts* opts, fuel_amd_batch** out) {
  if (!out) return fail(FUEL_AMD_ERR_INVALID, "out must not be NULL");
  *out = nullptr;
  if (count < 0 || (count > 0 && !problems)) return fail(FUEL_AMD_ERR_INVALID, "bad problem list");
  if (metric != FUEL_AMD_EUCLIDEAN && metric != FUEL_AMD_SQEUCLIDEAN)
    return fail(FUEL_AMD_ERR_UNSUPPORTED, "only the 'euclidean' and 'sqeuclidean' distances run on this engine");
  if (exclude_self < 0 || exclude_self > 2) return fail(FUEL_AMD_ERR_INVALID, "bad exclude_self");
  fuel_amd_batch_opts o{};
  if (opts) o = *opts;
  if (o.mode < FUEL_AMD_MODE_AUTO || o.mode > FUEL_AMD_MODE_BF16R) return fail(FUEL_AMD_ERR_INVALID, "bad mode");
  for (int i = 0; i < count; ++i) {
    const auto& p = problems[i];
    if (p.n < 0 || p.d < 0 || p.k < 0) return fail(FUEL_AMD_ERR_INVALID, "problem sizes must be non-negative");
    if (p.n > 0xFFFFFFFFll - 1) return fail(FUEL_AMD_ERR_UNSUPPORTED, "more than 2^32-2 rows");
    if (p.n > 0 && p.k > 0) {
      if (!p.data || !p.out_idx || !p.out_dist) return fail(FUEL_AMD_ERR_INVALID, "NULL problem buffer");
      if (p.ld < std::max(p.d, 1) || p.ld % 8 != 0) return fail(FUEL_AMD_ERR_INVALID, "ld must be >= d and a multiple of 8");
      if (reinterpret_cast<uintptr_t>(p.data) % 32 != 0) return fail(FUEL_AMD_ERR_INVALID, "rows must be 32-byte aligned");
    }
  }
  int32_t rc = check_device(o.device);
  if (rc) return rc;
  std::map<int, Cfg> cfg;
  std::string log = "[]";
  if (o.mode == FUEL_AMD_MODE_AUTO) {
    rc = autotune(problems, count, metric, exclude_self, o.device, o.tune_cache, cfg, log);
    if (rc) return rc;
  } else {
    const int kind = o.mode == FUEL_AMD_MODE_BF16 ? 0
                     : o.mode == FUEL_AMD_MODE_FP32 ? 1
                     : o.mode == FUEL_AMD_MODE_BF16R ? 2 : -1;
    for (int i = 0; i < count; ++i) {
      const int d = problems[i].d;
      // no residual variant beyond the widest one: the plain bf16 form
      const int k2 = (kind == 2 && mfma_ksteps_bucket(0, d) > mfma_resid_max_ksteps()) ? 0 : kind;
      cfg[d] = Cfg{k2, o.widen > 0 ? o.widen : 4, 1};
    }
  }
  std::vector<int> all(count);
  for (int i = 0; i < count; ++i) all[i] = i;
  std::unique_ptr<fuel_amd_batch> b;
  rc = make_batch(problems, all, metric, exclude_self, o.device, cfg, o.no_graph == 0, b);
  if (rc) return rc;
  b->tuning = log;
  *out = b.release();
  return FUEL_AMD_OK;
}

int32_t fuel_amd_batch_run(fuel_amd_batch* b) {
  if (!b) return fail(FUEL_AMD_ERR_INVALID, "batch must not be NULL");
  std::lock_guard<std::mutex> lock(b->mu);
  int32_t rc = check_device(b->device);
  if (rc) return rc;
  return run_once(b);
}

int32_t fuel_amd_batch_describe(fuel_amd_batch* b, char* text, int32_t len) {
  if (!b || !text || len <= 0) return fail(FUEL_AMD_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lock(b->mu);
  std::string s = "{\"graph\": ";
  s += (b->use_graph && b->exec) ? "true" : "false";
  s += ", \"groups\": [";
  for (size_t g = 0; g < b->buckets.size(); ++g) {
    const auto& B = b->buckets[g];
    int64_t rows = 0;
    int seeded = 0;  // problems whose level 0 starts from a seeded threshold
    for (int i : B.members) {
      rows += b->probs[i].n;
      seeded += (b->plan[i].seed.ntiles > 0 || b->plan[i].kind == 2) ? 1 : 0;
    }
    char buf[360];
    std::snprintf(buf, sizeof(buf),
                  "%s{\"d\": %d, \"problems\": %zu, \"rows\": %lld, \"mode\": \"%s\", \"widen\": %d, \"fill\": %d, "
                  "\"k_steps\": %d, \"work_items\": %lld, \"workgroups\": %d, \"seeded\": %d}",
                  g ? ", " : "", B.d, B.members.size(), (long long)rows, kind_name(B.cfg.kind), B.cfg.widen,
                  B.cfg.fill, B.ks, (long long)B.nitems, B.grid, seeded);
    s += buf;
  }
  int64_t nexact = 0;
  for (const auto& pp : b->plan) nexact += pp.exact ? 1 : 0;
  char tail[320];
  std::snprintf(tail, sizeof(tail),
                "], \"exact_problems\": %lld, \"runs\": %lld, \"last_ms\": %.3f, \"last_level0_ms\": %.3f, \"last_fallback_queries\": %lld, "
                "\"last_exact_queries\": %lld, \"tuning\": ",
                (long long)nexact, (long long)b->runs, b->last_ms, b->last_graph_ms, (long long)b->last_fallback,
                (long long)b->last_exact);
  s += tail;
  s += b->tuning;
  s += "}";
  std::strncpy(text, s.c_str(), (size_t)len - 1);
  text[len - 1] = 0;
  return s.size() < (size_t)len ? FUEL_AMD_OK : fail(FUEL_AMD_ERR_INVALID, "text buffer too small");
}

int32_t fuel_amd_batch_destroy(fuel_amd_batch* b) {
  if (!b) return FUEL_AMD_OK;
  (void)hipSetDevice(b->device);
  delete b;
  return FUEL_AMD_OK;
}

}  // extern "C"
