// Host orchestration of the mixed-precision kNN path:
//   prep (once per index and kind) -> per query batch: prep queries -> MFMA
//   candidate scan with fused selection -> exact rerank + certification ->
//   the queries whose bound did not certify go one level down the chain:
//   level 0  the requested kind (bf16 or f32), KP from the widening factor;
//   level 1  the same kind, started at the certified threshold theta the
//            rerank computed from the query's current KK-th exact key (every
//            row that can still enter the answer lies below it, so the scan
//            keeps only those and its answer certifies with theta);
//   level 2  f32 MFMA (after bf16), again from theta;
//   then the exact restated scan, which always answers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

namespace fa {

namespace {

constexpr size_t MFMA_SLOT_BUDGET = size_t(2) << 30;  // bytes of candidate slots per batch
constexpr int KP_MAX = 128;

int kp_for(const fuel_amd_index* ix, int k, int exclude) {
  const int KK = k + (exclude != EXCL_NONE ? 1 : 0);
  const int w = ix->opts.widen > 0 ? ix->opts.widen : 4;
  return std::max(KK, std::min(KP_MAX, w * k));
}

int32_t ensure_prepared(fuel_amd_index* ix, int kind, hipStream_t st) {
  auto& op = ix->mop[kind];
  if (op.ready.load(std::memory_order_acquire)) return FUEL_AMD_OK;
  std::lock_guard<std::mutex> lock(ix->prep_mu);
  if (op.ready.load(std::memory_order_relaxed)) return FUEL_AMD_OK;
  const int ks = mfma_ksteps_bucket(kind, ix->d);
  if (ks < 0) return fail(FUEL_AMD_ERR_UNSUPPORTED, "dimension too large for the MFMA candidate scan");
  const int kp = ks * (kind == 0 ? 16 : 8);
  const int64_t n_pad = std::max<int64_t>(256, (ix->n + 255) & ~int64_t(255));
  const size_t es = kind == 0 ? 2 : 4;
  hipError_t e = op.Xop.ensure(size_t(n_pad) * kp * es);
  if (e == hipSuccess) e = op.hx.ensure(sizeof(float) * n_pad);
  if (e == hipSuccess) e = op.consts.ensure(128);
  if (e != hipSuccess) return hip_fail(e, "MFMA operand copy");
  // 8 u64 bound constants (zero) followed by the two flags (one)
  int hf[2] = {1, 1};
  char init[128];
  std::memset(init, 0, sizeof(init));
  std::memcpy(init + 64, hf, sizeof(hf));
  e = hipMemcpyAsync(op.consts.p, init, 128, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "bound constants");
  PrepDbArgs pa{};
  pa.X = ix->X.as<float>();
  pa.ldx = ix->ld;
  pa.n = ix->n;
  pa.n_pad = n_pad;
  pa.d = ix->d;
  pa.Kp = kp;
  pa.bf16 = kind == 0;
  pa.Xop = op.Xop.p;
  pa.hx = op.hx.as<float>();
  pa.consts = op.consts.as<unsigned long long>();
  pa.flags = reinterpret_cast<int*>(op.consts.as<char>() + 64);
  e = launch_prep_db(pa, st);
  unsigned long long x2bits = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&x2bits, op.consts.p, sizeof(x2bits), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "MFMA operand preparation");
  double x2;
  std::memcpy(&x2, &x2bits, sizeof(x2));
  op.usable = x2 < MFMA_NORM2_LIMIT;
  op.ks = ks;
  op.kp = kp;
  ix->n_pad = n_pad;
  op.ready.store(true, std::memory_order_release);
  return FUEL_AMD_OK;
}

// Number of DB splits: fill the resident workgroup slots with as little tail as
// possible (the splits of a query share its threshold, so extra splits cost
// only the per-split warm-up and the rerank's merge of one more list).  Small
// batches (serving) need many splits to fill the chip: one query block of 64
// queries against 10M rows runs as ~500 splits of ~20k rows.
int mfma_splits(int64_t nqb, int64_t n_pad, int cus, int64_t rows_tile, int wg_per_cu) {
  const int64_t slots = int64_t(cus) * wg_per_cu;
  const int64_t max_s = std::max<int64_t>(1, std::min<int64_t>(1024, n_pad / (rows_tile * 4)));
  int best = 1;
  double best_eff = -1.0;
  for (int64_t s = 1; s <= max_s; ++s) {
    const int64_t units = nqb * s;
    const int64_t rounds = (units + slots - 1) / slots;
    const double eff = double(units) / double(rounds * slots);
    if (eff > best_eff + 0.03) {
      best_eff = eff;
      best = (int)s;
    }
    if (eff > 0.93) break;
  }
  return best;
}

}  // namespace

bool mfma_supported(const fuel_amd_index* ix, int kind, int k, int exclude) {
  if (ix->dtype != FUEL_AMD_F32) return false;
  if (mfma_ksteps_bucket(kind, ix->d) < 0) return false;
  const int KK = k + (exclude != EXCL_NONE ? 1 : 0);
  return KK <= KP_MAX;
}

int32_t prepare_mode(fuel_amd_index* ix, hipStream_t st) {
  int mode = ix->opts.mode;
  if (mode == FUEL_AMD_MODE_AUTO) mode = FUEL_AMD_MODE_BF16;
  if (mode == FUEL_AMD_MODE_EXACT || ix->dtype != FUEL_AMD_F32) return FUEL_AMD_OK;
  const int kind = mode == FUEL_AMD_MODE_BF16 ? 0 : 1;
  if (mfma_ksteps_bucket(kind, ix->d) < 0) return FUEL_AMD_OK;
  int32_t rc = ensure_prepared(ix, kind, st);
  // the certification chain continues bf16 with f32
  if (rc == FUEL_AMD_OK && kind == 0 && mfma_ksteps_bucket(1, ix->d) >= 0) rc = ensure_prepared(ix, 1, st);
  return rc;
}

int32_t knn_mfma(fuel_amd_index* ix, QueryCtx* ctx, const float* Q, int64_t m, int64_t ldq, int k, int exclude,
                 int kind, int64_t* out_idx, float* out_dist, hipStream_t st, const int64_t* d_rows, int64_t qbase,
                 int level, int kp, const float* d_theta) {
  int32_t rc = ensure_prepared(ix, kind, st);
  if (rc) return rc;
  auto& op = ix->mop[kind];
  if (!op.usable) {
    // norms beyond the scan's finite sentinels: every query takes the exact scan
    ctx->stats.exact_queries += m;
    return knn_exact_rows(ix, ctx, Q, ldq, d_rows, m, qbase, k, exclude, out_idx, out_dist, st);
  }
  auto& ws = ctx->lv[level];
  const int KK = k + (exclude != EXCL_NONE ? 1 : 0);
  const int KP = kp > 0 ? std::max(kp, k + (exclude != EXCL_NONE ? 1 : 0)) : kp_for(ix, k, exclude);
  const int V = KP <= 64 ? 4 : 8;
  const int NS = 64 * V;
  const int csub = (NS - KP) / 2;
  const int ks = op.ks;
  const int Kp = op.kp;
  const int qpb = mfma_queries_per_block(kind, ks);
  const int rows_tile = mfma_rows_per_tile(kind, ks);
  const int wg_per_cu = mfma_workgroups_per_cu(kind, ks);
  const int cus = device_cu_count(ix->device);
  const size_t es = kind == 0 ? 2 : 4;
  const int64_t ld32 = ix->ld;
  const bool timed = level == 0 && ctx->timed;

  int64_t done = 0;
  while (done < m) {
    if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
    int64_t chunk = m - done;
    int64_t nqb = (chunk + qpb - 1) / qpb;
    int nsplit = mfma_splits(nqb, ix->n_pad, cus, rows_tile, wg_per_cu);
    const size_t per_q = size_t(nsplit) * NS * 8;
    const int64_t max_chunk = std::max<int64_t>(qpb, int64_t(MFMA_SLOT_BUDGET / per_q) / qpb * qpb);
    if (chunk > max_chunk) {
      chunk = max_chunk;
      nqb = chunk / qpb;
    }
    const int64_t nq_pad = nqb * qpb;
    int64_t rps = (ix->n_pad + nsplit - 1) / nsplit;
    rps = (rps + rows_tile - 1) / rows_tile * rows_tile;
    nsplit = (int)((ix->n_pad + rps - 1) / rps);

    hipError_t e = ws.Qop.ensure(size_t(nq_pad) * Kp * es);
    if (e == hipSuccess) e = ws.Q32.ensure(size_t(chunk) * ld32 * 4);
    if (e == hipSuccess) e = ws.qnorm.ensure(sizeof(double) * 3 * size_t(chunk));
    if (e == hipSuccess) e = ws.qint.ensure(sizeof(int) * size_t(chunk));
    if (e == hipSuccess) e = ws.slots.ensure(per_q * nq_pad);
    if (e == hipSuccess) e = ws.lcnts.ensure(sizeof(int) * size_t(nsplit) * nq_pad);
    if (e == hipSuccess) e = ws.fail_count.ensure(64);
    if (e == hipSuccess) e = ws.fail_list.ensure(sizeof(int64_t) * size_t(chunk));
    if (e == hipSuccess) e = ws.fail_theta.ensure(sizeof(float) * size_t(chunk));
    if (e == hipSuccess) e = ws.gthr.ensure(sizeof(uint32_t) * size_t(nq_pad));
    if (e != hipSuccess) return hip_fail(e, "MFMA workspaces");

    // rows of this chunk: an explicit list (fallback levels) or done + i
    const int64_t* rows = d_rows ? d_rows + done : nullptr;
    const float* Qb = d_rows ? Q : Q + done * ldq;

    PrepQArgs qa{};
    qa.Q = Qb;
    qa.ldq = ldq;
    qa.m = chunk;
    qa.m_pad = nq_pad;
    qa.d = ix->d;
    qa.Kp = Kp;
    qa.bf16 = kind == 0;
    qa.Qop = ws.Qop.p;
    qa.Q32 = ws.Q32.as<float>();
    qa.ld32 = ld32;
    qa.qn2 = ws.qnorm.as<double>();
    qa.eq2 = qa.qn2 + chunk;
    qa.qt2 = qa.eq2 + chunk;
    qa.qint = ws.qint.as<int>();
    qa.gthr = ws.gthr.as<uint32_t>();
    qa.rows = rows;
    qa.theta = d_theta ? d_theta + done : nullptr;
    e = launch_prep_queries(qa, st);
    if (e == hipSuccess) e = hipMemsetAsync(ws.fail_count.p, 0, sizeof(int), st);
    if (e != hipSuccess) return hip_fail(e, "query preparation");

    ScanMfmaArgs sa{};
    sa.Xop = op.Xop.p;
    sa.hx = op.hx.as<float>();
    sa.Kp = Kp;
    sa.n_pad = ix->n_pad;
    sa.Qop = ws.Qop.p;
    sa.nq_pad = nq_pad;
    sa.nqb = (int)nqb;
    sa.rows_per_split = rps;
    sa.KP = KP;
    sa.Csub = csub;
    sa.slots = ws.slots.as<uint64_t>();
    sa.lcnts = ws.lcnts.as<int>();
    sa.gthr = ws.gthr.as<uint32_t>();
    if (timed) timing_mark(ix, st, true, kind == 0 ? "scan_mfma_kernel<bf16>" : "scan_mfma_kernel<f32>");
    e = launch_scan_mfma(sa, kind, ks, V, nsplit, st);
    if (timed) timing_mark(ix, st, false, nullptr);
    if (e != hipSuccess) return hip_fail(e, "MFMA scan launch");

    RerankArgs ra{};
    ra.slots = sa.slots;
    ra.lcnts = sa.lcnts;
    ra.nsplit = nsplit;
    ra.nq = chunk;
    ra.nq_pad = nq_pad;
    ra.slot_ns = NS;
    ra.KP = KP;
    ra.X = ix->X.as<float>();
    ra.ldx = ix->ld;
    ra.Q32 = qa.Q32;
    ra.ld32 = ld32;
    ra.d = ix->d;
    ra.metric = ix->metric;
    ra.bf16 = kind == 0;
    ra.qn2 = qa.qn2;
    ra.eq2 = qa.eq2;
    ra.qt2 = qa.qt2;
    ra.qint = qa.qint;
    ra.consts = op.consts.as<unsigned long long>();
    ra.flags = reinterpret_cast<const int*>(op.consts.as<char>() + 64);
    ra.KK = KK;
    ra.k = k;
    ra.mode = exclude;
    ra.qbase = d_rows ? qbase : qbase + done;
    ra.rows = rows;
    ra.out_idx = d_rows ? out_idx : out_idx + done * k;
    ra.out_dist = d_rows ? out_dist : out_dist + done * k;
    ra.fail_count = ws.fail_count.as<int>();
    ra.fail_list = ws.fail_list.as<int64_t>();
    ra.gthr = sa.gthr;
    ra.fail_theta = ws.fail_theta.as<float>();
    e = launch_rerank(ra, V, st);
    if (e != hipSuccess) return hip_fail(e, "rerank launch");

    int nfail = 0;
    e = hipMemcpyAsync(&nfail, ws.fail_count.p, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "certification count");
    if (nfail > 0) {
      // the fail list holds query rows relative to Qb and to this chunk's outputs
      int next_kind = -1;
      if (level == 0)
        next_kind = kind;
      else if (level == 1 && kind == 0 && mfma_supported(ix, 1, k, exclude))
        next_kind = 1;
      if (next_kind >= 0) {
        rc = knn_mfma(ix, ctx, Qb, nfail, ldq, k, exclude, next_kind, ra.out_idx, ra.out_dist, st, ra.fail_list,
                      ra.qbase, level + 1, KP_MAX, ra.fail_theta);
      } else {
        ctx->stats.exact_queries += nfail;  // queries answered by the exact scan
        rc = knn_exact_rows(ix, ctx, Qb, ldq, ra.fail_list, nfail, ra.qbase, k, exclude, ra.out_idx, ra.out_dist,
                            st);
      }
      if (rc) return rc;
    }
    if (level == 0) {
      ctx->stats.fallback_queries += nfail;
      ctx->stats.splits = nsplit;
      ctx->stats.kp = KP;
    }
    done += chunk;
  }
  return FUEL_AMD_OK;
}

}  // namespace fa
