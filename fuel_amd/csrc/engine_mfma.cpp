// Host orchestration of the mixed-precision kNN path:
//   prep (once per index and kind) -> per query batch: prep queries -> MFMA
//   candidate scan with fused selection -> exact rerank + certification ->
//   the queries whose bound did not certify go one level down the chain:
//   level 0  the requested kind (bf16, f32, or bf16 / E4M3 on the residual
//            operand), KP from the widening factor;
//   level 1  the same kind, started at the certified threshold theta the
//            rerank computed from the query's current KK-th exact key (every
//            row that can still enter the answer lies below it, so the scan
//            keeps only those and its answer certifies with theta);
//   level 2  f32 MFMA (after bf16), again from theta;
//   then the exact restated scan, which always answers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

namespace fa {

namespace {

constexpr size_t MFMA_SLOT_BUDGET = size_t(4) << 30;  // bytes of candidate slots per batch
constexpr size_t MFMA_QFRAG_BUDGET = size_t(1) << 30;  // bytes of E4M3 query fragments per batch
constexpr int kMaxLists = 256;                          // resid_lists never exceeds it
// Widest kept list: 128 (slots of 512 hold it with two 192-entry append
// segments), 256 for the E4M3 copy, whose bound band is wider (two 128-entry
// segments, still whole tiles of appends).
constexpr int kp_max(int kind) { return kind == 3 ? 256 : 128; }
// A pass from the certified threshold keeps up to 256 for every kind: only the
// rows of the query's bound band append, and a wide band is what sent the query
// there.  (A level without a threshold, the f32 pass after a residual one,
// keeps its kind's first-level limit.)
constexpr int KP_CHAIN = 256;
constexpr size_t kFailRowsKept = size_t(1) << 20;  // rows of failed queries recorded per call

// Candidates kept per query: widen * k (default 4; 16 for the E4M3 copy, whose
// bound band is about ten times the bf16 one), at least KK, at most kp_max.
int kp_for(const fuel_amd_index* ix, int kind, int k, int exclude) {
  const int KK = k + (exclude != EXCL_NONE ? 1 : 0);
  const int w = ix->opts.widen > 0 ? ix->opts.widen : (kind == 3 ? 16 : 4);
  return std::max(KK, std::min(kp_max(kind), w * k));
}

// Lists of the residual operand: about 256 rows per list, at most 256.  Many
// more lists than clusters: a list that straddles two clusters has a wide
// per-list bound, and evenly spaced starting rows miss few clusters when
// they outnumber them several times (with 512-row lists a 40k-row,
// 50-cluster set sent a quarter of its queries down the chain).
int resid_lists(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(256, n / 256)); }

}  // namespace

// Padded layout of the residual operand for lists of cnt[j] rows: list j
// occupies the 64-row blocks [pad[j] / 64, pad[j + 1] / 64); the operand has
// n_rpad rows (a multiple of 256, the value returned) and tl[b] is the list of
// block b (the last list for the blocks past the end), so every entry of tl
// lies in [0, L).
int64_t resid_layout(const std::vector<uint32_t>& cnt, std::vector<int64_t>& pad, std::vector<int>& tl) {
  const int L = (int)cnt.size();
  pad.assign(L + 1, 0);
  for (int j = 0; j < L; ++j) pad[j + 1] = pad[j] + ((int64_t(cnt[j]) + 63) & ~int64_t(63));
  const int64_t n_rpad = std::max<int64_t>(256, (pad[L] + 255) & ~int64_t(255));
  tl.assign(n_rpad / 64, 0);
  for (int j = 0; j < L; ++j)
    for (int64_t b = pad[j] / 64; b < pad[j + 1] / 64; ++b) tl[b] = j;
  for (int64_t b = pad[L] / 64; b < n_rpad / 64; ++b) tl[b] = L - 1;
  return n_rpad;
}

// Partition (Lloyd steps from evenly spaced rows) and the sorted padded layout
// of the residual operand.
int32_t build_partition(const float* X, int64_t ld, int64_t n, int d, hipStream_t st, DevBuf& cents,
                        DevBuf& rperm, DevBuf& tile_list, DevBuf& lcount, DevBuf& loff, int& L_out,
                        int64_t& n_rpad_out) {
  const int L = resid_lists(n);
  DevBuf list, counts, keys_out, vals_in, vals_out, tmp, offs, poffs;
  hipError_t e = list.ensure(sizeof(uint32_t) * std::max<int64_t>(n, 1));
  if (e == hipSuccess) e = counts.ensure(sizeof(uint32_t) * L);
  if (e == hipSuccess) e = keys_out.ensure(sizeof(uint32_t) * std::max<int64_t>(n, 1));
  if (e == hipSuccess) e = vals_in.ensure(sizeof(uint32_t) * std::max<int64_t>(n, 1));
  if (e == hipSuccess) e = vals_out.ensure(sizeof(uint32_t) * std::max<int64_t>(n, 1));
  if (e == hipSuccess) e = offs.ensure(sizeof(int64_t) * (L + 1));
  if (e == hipSuccess) e = poffs.ensure(sizeof(int64_t) * (L + 1));
  if (e == hipSuccess) e = cents.ensure(sizeof(float) * size_t(L) * d);
  if (e != hipSuccess) return hip_fail(e, "partition workspaces");
  int bits = 1;
  while ((1 << bits) < L) ++bits;
  size_t tmp_bytes = 0;
  e = sort_pairs_u32(list.as<uint32_t>(), keys_out.as<uint32_t>(), vals_in.as<uint32_t>(), vals_out.as<uint32_t>(),
                     n, bits, nullptr, tmp_bytes, st);
  if (e == hipSuccess) e = tmp.ensure(std::max<size_t>(tmp_bytes, 256));
  // initial centroids: rows floor(i n / L)
  for (int i = 0; i < L && e == hipSuccess; ++i)
    e = hipMemcpyAsync(cents.as<float>() + size_t(i) * d, X + (int64_t(i) * n / L) * ld, sizeof(float) * d,
                       hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "partition init");
  std::vector<uint32_t> hcnt(L);
  std::vector<int64_t> hoff(L + 1);
  constexpr int kLloydSteps = 4;
  for (int it = 0; it <= kLloydSteps; ++it) {
    AssignArgs aa{};
    aa.X = X;
    aa.ldx = ld;
    aa.n = n;
    aa.d = d;
    aa.cents = cents.as<float>();
    aa.L = L;
    aa.list = list.as<uint32_t>();
    aa.counts = counts.as<uint32_t>();
    e = launch_fill_u32(counts.as<uint32_t>(), L, 0u, st);
    if (e == hipSuccess) e = launch_assign(aa, st);
    if (e == hipSuccess) e = launch_iota_u32(vals_in.as<uint32_t>(), n, st);
    if (e == hipSuccess)
      e = sort_pairs_u32(list.as<uint32_t>(), keys_out.as<uint32_t>(), vals_in.as<uint32_t>(),
                         vals_out.as<uint32_t>(), n, bits, tmp.p, tmp_bytes, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hcnt.data(), counts.p, sizeof(uint32_t) * L, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "partition");
    hoff[0] = 0;
    for (int j = 0; j < L; ++j) hoff[j + 1] = hoff[j] + hcnt[j];
    e = hipMemcpyAsync(offs.p, hoff.data(), sizeof(int64_t) * (L + 1), hipMemcpyHostToDevice, st);
    if (it < kLloydSteps && e == hipSuccess)
      e = launch_list_means(X, ld, d, vals_out.as<uint32_t>(), offs.as<int64_t>(), counts.as<uint32_t>(), L,
                            cents.as<float>(), st);
    if (e != hipSuccess) return hip_fail(e, "partition");
  }
  std::vector<int64_t> hpad;
  std::vector<int> htl;
  const int64_t n_rpad = resid_layout(hcnt, hpad, htl);
  e = rperm.ensure(sizeof(uint32_t) * n_rpad);
  if (e == hipSuccess) e = tile_list.ensure(sizeof(int) * htl.size());
  if (e == hipSuccess) e = lcount.ensure(sizeof(uint32_t) * L);
  if (e == hipSuccess) e = loff.ensure(sizeof(int64_t) * (L + 1));
  if (e == hipSuccess) e = hipMemcpyAsync(lcount.p, hcnt.data(), sizeof(uint32_t) * L, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "residual layout");
  e = hipMemcpyAsync(poffs.p, hpad.data(), sizeof(int64_t) * (L + 1), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(loff.p, hpad.data(), sizeof(int64_t) * (L + 1), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(tile_list.p, htl.data(), sizeof(int) * htl.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_fill_u32(rperm.as<uint32_t>(), n_rpad, RPERM_NONE, st);
  if (e == hipSuccess)
    e = launch_scatter_rperm(vals_out.as<uint32_t>(), keys_out.as<uint32_t>(), n, offs.as<int64_t>(),
                             poffs.as<int64_t>(), rperm.as<uint32_t>(), st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // before the workspaces above are released
  if (e != hipSuccess) return hip_fail(e, "residual layout");
  L_out = L;
  n_rpad_out = n_rpad;
  return FUEL_AMD_OK;
}

namespace {

// The residual operand over the partition: operand rows (bf16, or E4M3 codes
// with per-row scales), per-list constants.
int32_t prepare_resid(fuel_amd_index* ix, fuel_amd_index::MfmaOperand& op, bool fp8, hipStream_t st) {
  const int64_t n = ix->n;
  const int d = ix->d;
  const int ks = mfma_ksteps_bucket(fp8 ? 2 : 0, d);
  if (ks < 0 || ks > mfma_resid_max_ksteps())
    return fail(FUEL_AMD_ERR_UNSUPPORTED, "dimension too large for the residual operand");
  const int Kp = ks * (fp8 ? 32 : 16);
  const float* X = ix->X.as<float>();
  int L = 0;
  int64_t n_rpad = 0;
  int32_t rc =
      build_partition(X, ix->ld, n, d, st, op.cents, op.rperm, op.tile_list, op.lcount, op.loff, L, n_rpad);
  if (rc) return rc;
  hipError_t e = op.lconsts.ensure(sizeof(unsigned long long) * 8 * L);
  if (e == hipSuccess) e = op.Xop.ensure((fp8 ? 1 : sizeof(uint16_t)) * size_t(n_rpad) * Kp);
  if (e == hipSuccess && fp8) e = op.xsc.ensure(size_t(n_rpad));
  if (e == hipSuccess) e = op.hx.ensure(sizeof(float) * n_rpad);
  if (e == hipSuccess) e = op.consts.ensure(128);
  if (e != hipSuccess) return hip_fail(e, "residual operand");
  e = launch_fill_u32(reinterpret_cast<uint32_t*>(op.lconsts.p), 16 * L, 0u, st);
  if (e == hipSuccess) e = launch_fill_u32(op.consts.as<uint32_t>(), 32, 0u, st);  // constants and flags
  PrepResidArgs pr{};
  pr.X = X;
  pr.ldx = ix->ld;
  pr.d = d;
  pr.Kp = Kp;
  pr.n_rpad = n_rpad;
  pr.rperm = op.rperm.as<uint32_t>();
  pr.tile_list = op.tile_list.as<int>();
  pr.cents = op.cents.as<float>();
  pr.ldc = d;
  pr.Xop = op.Xop.p;
  pr.xsc = fp8 ? op.xsc.as<uint8_t>() : nullptr;
  pr.fp8 = fp8 ? 1 : 0;
  pr.hx = op.hx.as<float>();
  pr.lconsts = op.lconsts.as<unsigned long long>();
  pr.gconsts = op.consts.as<unsigned long long>();
  if (e == hipSuccess) e = launch_prep_resid(pr, st);
  unsigned long long x2bits = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&x2bits, op.consts.p, sizeof(x2bits), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "residual operand");
  double x2;
  std::memcpy(&x2, &x2bits, sizeof(x2));
  op.usable = x2 < MFMA_NORM2_LIMIT;
  op.ks = ks;
  op.kp = Kp;
  op.n_pad = n_rpad;
  op.L = L;
  return FUEL_AMD_OK;
}

int32_t ensure_prepared(fuel_amd_index* ix, int kind, hipStream_t st) {
  auto& op = ix->mop[kind];
  if (op.ready.load(std::memory_order_acquire)) return FUEL_AMD_OK;
  std::lock_guard<std::mutex> lock(ix->prep_mu);
  if (op.ready.load(std::memory_order_relaxed)) return FUEL_AMD_OK;
  if (kind == 2 || kind == 3) {
    int32_t rc = prepare_resid(ix, op, kind == 3, st);
    if (rc) return rc;
    op.ready.store(true, std::memory_order_release);
    return FUEL_AMD_OK;
  }
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
  op.n_pad = n_pad;
  op.ready.store(true, std::memory_order_release);
  return FUEL_AMD_OK;
}

// Number of DB splits: fill the resident workgroup slots with as little tail as
// possible (the splits of a query share its threshold, so extra splits cost
// only the per-split warm-up and the rerank's merge of one more list).  Small
// batches (serving) need many splits to fill the chip: one query block of 64
// queries against 10M rows runs as ~500 splits of ~20k rows.  Within one round
// of workgroups a fill above 93 % is taken; over several rounds every further
// 3 % of fill is taken, up to max_split (the splits whose candidate slots fit
// the budget for the whole batch: chunking the batch would cost more than the
// tail).  100,000 queries x 1M rows at d = 128 (391 query blocks, 1024 slots):
// 13 splits fill 99.3 % of five rounds where 5 filled 95.5 % of two, and the
// scan ran 2-4 % faster (10 splits: 1-2 %; 8, 76 % of four rounds: slower).
int mfma_splits(int64_t nqb, int64_t n_pad, int cus, int64_t rows_tile, int wg_per_cu, int64_t max_split) {
  const int64_t slots = int64_t(cus) * wg_per_cu;
  const int64_t max_s =
      std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(1024, n_pad / (rows_tile * 4)), max_split));
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
    if (eff > 0.93 && rounds == 1) break;
  }
  return best;
}

}  // namespace

// Sample of the seed pass (ntiles = 0: no seed).  Every group of every part
// must receive at least one row block (parts * SEED_GROUPS / R tiles), the
// groups number about 4 KP (at least KP), the sample is a fraction of the
// operand's tiles that grows with KP / n (at most 4096 tiles), and operands of
// fewer than two samples' worth of tiles start cold.
SeedPlan plan_seed(int mk, int ks, bool resid, bool has_theta, int flags, int64_t n_pad, int rows_tile, int KP) {
  SeedPlan p;
  if (resid || has_theta || (flags & FUEL_AMD_FLAG_NO_SEED)) return p;
  // the f32 form at more than 4 K steps is bound by its matrix-core rate, not by
  // the candidate appends the seed saves (1M x 128: scan -4.6 ms, seed +13 ms)
  if (mk == 1 && ks > 4) return p;
  const int R = rows_tile / 32;
  // about 4 KP group minima, so that the KP-th of them sits near the sample's
  // KP-th key rather than near its largest group minimum
  p.parts = std::min(32, std::max(4, (4 * KP + 2 * SEED_GROUPS - 1) / (2 * SEED_GROUPS)));
  const int64_t tiles = n_pad / rows_tile;
  const int64_t need = int64_t(p.parts) * SEED_GROUPS / R;
  if (tiles < 2 * need) {
    p.parts = 0;
    return p;
  }
  // Sample fraction: a query still appends about KP / f rows after the seed and
  // the seed reads f n rows, so f grows as sqrt(KP / n): 1/11.3 at KP = 40 and
  // 10^6 rows (measured against 1/16 and 1/22.6 on one box: 26.4 against 26.9
  // and 26.6 ms per step), at most 1/4.
  const double f = std::min(0.25, std::sqrt(50000.0 * double(KP) / double(n_pad)) / 16.0);
  p.ntiles = std::min<int64_t>(4096, std::max<int64_t>(need, (int64_t)(f * double(tiles))));
  p.ntiles = std::min(p.ntiles, tiles);
  p.stride = (tiles / p.ntiles) * rows_tile;
  return p;
}

MfmaChunkPlan plan_mfma_chunk(int mk, bool resid, int ks, int64_t m_total, int64_t remaining, int KP, int64_t n_pad,
                              int cus) {
  MfmaChunkPlan p{};
  p.V = KP <= 64 ? 4 : 8;
  p.NS = 64 * p.V;
  p.csub = (p.NS - KP) / 2;
  // small residual batches (serving): one-wave workgroups of 64 query slots
  p.waves = (m_total <= 192 && mfma_one_wave_supported(mk, ks, p.V, resid)) ? 1 : 4;
  p.qpb = mfma_queries_per_block(mk, ks, p.waves);
  p.rows_tile = mfma_rows_per_tile(mk, ks, resid);
  p.wg_per_cu = mfma_workgroups_per_cu(mk, ks, p.waves, resid);
  p.chunk = remaining;
  p.nqb = (p.chunk + p.qpb - 1) / p.qpb;
  const size_t per_split = size_t(p.NS) * 8 * size_t(p.nqb * p.qpb);  // slot bytes of one split
  int nsplit = mfma_splits(p.nqb, n_pad, cus, p.rows_tile, p.wg_per_cu,
                           std::max<int64_t>(1, int64_t(MFMA_SLOT_BUDGET / per_split)));
  const size_t per_q = size_t(nsplit) * p.NS * 8;
  const int64_t max_chunk = std::max<int64_t>(p.qpb, int64_t(MFMA_SLOT_BUDGET / per_q) / p.qpb * p.qpb);
  if (p.chunk > max_chunk) {
    p.chunk = max_chunk;
    p.nqb = p.chunk / p.qpb;
  }
  if (mk == 2) {
    // the E4M3 scan reads prepared B fragments: 2 x Kp bytes per (list, query slot)
    const size_t per_qf = size_t(2) * kMaxLists * size_t(ks) * 32;
    const int64_t max_qf = std::max<int64_t>(p.qpb, int64_t(MFMA_QFRAG_BUDGET / per_qf) / p.qpb * p.qpb);
    if (p.chunk > max_qf) {
      p.chunk = max_qf;
      p.nqb = p.chunk / p.qpb;
    }
  }
  p.nq_pad = p.nqb * p.qpb;
  p.rps = (n_pad + nsplit - 1) / nsplit;
  p.rps = (p.rps + p.rows_tile - 1) / p.rows_tile * p.rows_tile;
  p.nsplit = (int)((n_pad + p.rps - 1) / p.rps);  // <= the count the slots are sized for
  p.slots_bytes = per_q * size_t(p.nq_pad);
  return p;
}

int64_t mfma_level0_bytes(fuel_amd_index* ix, int kind, int64_t m, int k, int exclude) {
  if (kind < 0 || kind > 3 || m <= 0) return 0;
  auto& op = ix->mop[kind];
  if (!op.ready.load(std::memory_order_acquire)) return 0;
  const bool resid = kind >= 2;
  const int mk = mfma_operand_type(kind);
  const size_t es = mk == 0 ? 2 : (mk == 1 ? 4 : 1);
  const int KP = kp_for(ix, kind, k, exclude);
  const MfmaChunkPlan pl = plan_mfma_chunk(mk, resid, op.ks, m, m, KP, op.n_pad, device_cu_count(ix->device));
  const size_t nq = size_t(pl.nq_pad), ch = size_t(pl.chunk);
  // slots, split counts, thresholds, f32 queries, three norms, integer flags,
  // failure list and thresholds, failure count
  size_t b = pl.slots_bytes + 4 * size_t(pl.nsplit) * nq + 4 * nq + ch * size_t(ix->ld) * 4 + 24 * ch + 4 * ch +
             12 * ch + 64;
  if (resid) {
    b += 4 * size_t(op.L) * nq + 4 * size_t(op.L) * ch;  // per-list terms and ||q - c||^2
    if (kind == 3) b += 2 * size_t(op.L) * nq * size_t(op.kp) + 4 * size_t(op.L) * nq;  // E4M3 fragments
  } else {
    b += nq * size_t(op.kp) * es;  // query operand
    const SeedPlan sp = plan_seed(mk, op.ks, false, false, ix->opts.flags, op.n_pad, pl.rows_tile, KP);
    if (sp.ntiles > 0) b += 4 * nq * size_t(sp.parts) * 2 * SEED_GROUPS;
  }
  return (int64_t)b;
}

bool mfma_supported(const fuel_amd_index* ix, int kind, int k, int exclude) {
  if (ix->dtype != FUEL_AMD_F32) return false;
  const int ks = mfma_ksteps_bucket(mfma_operand_type(kind), ix->d);
  if (ks < 0 || (kind >= 2 && ks > mfma_resid_max_ksteps())) return false;
  const int KK = k + (exclude != EXCL_NONE ? 1 : 0);
  return KK <= kp_max(kind);
}

// Plain bf16 certifies when the neighbour gaps are wide next to ~2^-8 |q| |x|
// (uniform data); on clustered data with large norms most queries fail it,
// and the residual form certifies instead.  The probe runs the first level of
// the plain bf16 form on PROBE_ROWS evenly spaced index rows as queries; if
// more than a tenth of them fail and the residual form has a variant for this
// width, AUTO means the residual form.  The probe only picks the route: every
// mode returns the exact answer.
int32_t resolve_auto_mode(fuel_amd_index* ix, QueryCtx* ctx, int k, hipStream_t st, int* mode) {
  int m = ix->auto_mode.load(std::memory_order_acquire);
  if (m >= 0) {
    *mode = m;
    return FUEL_AMD_OK;
  }
  std::lock_guard<std::mutex> lock(ix->auto_mu);
  m = ix->auto_mode.load(std::memory_order_relaxed);
  if (m >= 0) {
    *mode = m;
    return FUEL_AMD_OK;
  }
  constexpr int64_t PROBE_ROWS = 256;
  constexpr int64_t PROBE_MIN_N = 4096;
  m = FUEL_AMD_MODE_BF16;
  if (ix->dtype != FUEL_AMD_F32 || mfma_ksteps_bucket(0, ix->d) < 0) {
    m = FUEL_AMD_MODE_EXACT;
  } else if (ix->n >= PROBE_MIN_N && k >= 1 && mfma_supported(ix, 0, k, EXCL_NONE) &&
             mfma_supported(ix, 2, k, EXCL_NONE)) {
    const int64_t S = PROBE_ROWS;
    DevBuf q, oi, od;
    hipError_t e = q.ensure(sizeof(float) * size_t(S) * ix->ld);
    if (e == hipSuccess) e = oi.ensure(sizeof(int64_t) * size_t(S) * k);
    if (e == hipSuccess) e = od.ensure(sizeof(float) * size_t(S) * k);
    // rows i * n / S, gathered with one strided copy when n is a multiple of S,
    // else row by row (256 small copies, once per index)
    const float* X = ix->X.as<float>();
    if (e == hipSuccess && ix->n % S == 0)
      e = hipMemcpy2DAsync(q.p, sizeof(float) * ix->ld, X, sizeof(float) * ix->ld * (ix->n / S),
                           sizeof(float) * ix->ld, S, hipMemcpyDeviceToDevice, st);
    for (int64_t i = 0; e == hipSuccess && ix->n % S != 0 && i < S; ++i)
      e = hipMemcpyAsync(q.as<float>() + i * ix->ld, X + (i * ix->n / S) * ix->ld, sizeof(float) * ix->ld,
                         hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "mode probe");
    const fuel_amd_stats saved = ctx->stats;
    ctx->stats = fuel_amd_stats{};
    int32_t rc = knn_mfma(ix, ctx, q.as<float>(), S, ix->ld, k, EXCL_NONE, 0, oi.as<int64_t>(), od.as<float>(), st,
                          nullptr, 0, 0, 0, nullptr, true);
    const int64_t nfail = ctx->stats.fallback_queries + ctx->stats.exact_queries;
    ctx->stats = saved;
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // before the probe buffers are released
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "mode probe");
    ix->auto_probe_rows = S;
    ix->auto_probe_fail = nfail;
    if (nfail * 10 > S) {
      m = FUEL_AMD_MODE_BF16R;
      // the residual chain continues with f32, never with the plain bf16 copy
      std::lock_guard<std::mutex> plock(ix->prep_mu);
      auto& op = ix->mop[0];
      op.ready.store(false, std::memory_order_release);
      op.Xop.release();
      op.hx.release();
      op.consts.release();
    }
  }
  ix->auto_mode.store(m, std::memory_order_release);
  *mode = m;
  return FUEL_AMD_OK;
}

int32_t prepare_mode(fuel_amd_index* ix, hipStream_t st) {
  int mode = ix->opts.mode;
  if (mode == FUEL_AMD_MODE_AUTO) {
    int32_t rc = resolve_auto_mode(ix, &ix->ctx, 10, st, &mode);
    if (rc) return rc;
  }
  if (mode == FUEL_AMD_MODE_EXACT || ix->dtype != FUEL_AMD_F32) return FUEL_AMD_OK;
  int kind = mode == FUEL_AMD_MODE_BF16 ? 0 : (mode == FUEL_AMD_MODE_BF16R ? 2 : (mode == FUEL_AMD_MODE_FP8R ? 3 : 1));
  if (kind == 3 && !mfma_supported(ix, 3, 1, 0)) kind = 2;  // as knn_run
  if (kind == 2 && !mfma_supported(ix, 2, 1, 0)) kind = 0;
  if (mfma_ksteps_bucket(mfma_operand_type(kind), ix->d) < 0) return FUEL_AMD_OK;
  int32_t rc = ensure_prepared(ix, kind, st);
  // the certification chain continues bf16 with f32
  if (rc == FUEL_AMD_OK && kind != 1 && mfma_ksteps_bucket(1, ix->d) >= 0) rc = ensure_prepared(ix, 1, st);
  return rc;
}

int32_t knn_mfma(fuel_amd_index* ix, QueryCtx* ctx, const float* Q, int64_t m, int64_t ldq, int k, int exclude,
                 int kind, int64_t* out_idx, float* out_dist, hipStream_t st, const int64_t* d_rows, int64_t qbase,
                 int level, int kp, const float* d_theta, bool probe) {
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
  const int KP = kp > 0 ? std::max(kp, k + (exclude != EXCL_NONE ? 1 : 0)) : kp_for(ix, kind, k, exclude);
  const int ks = op.ks;
  const int Kp = op.kp;
  const bool resid = kind >= 2;
  const int mk = mfma_operand_type(kind);
  const int cus = device_cu_count(ix->device);
  const size_t es = mk == 0 ? 2 : (mk == 1 ? 4 : 1);
  const int64_t n_pad = op.n_pad;
  const int64_t ld32 = ix->ld;
  const bool timed = level == 0 && ctx->timed;

  int64_t done = 0;
  while (done < m) {
    if (cancelled()) return fail(FUEL_AMD_ERR_INTERRUPTED, "interrupted");
    const MfmaChunkPlan pl = plan_mfma_chunk(mk, resid, ks, m, m - done, KP, n_pad, cus);
    const int V = pl.V, NS = pl.NS, csub = pl.csub, waves = pl.waves, nsplit = pl.nsplit;
    const int64_t chunk = pl.chunk, nqb = pl.nqb, nq_pad = pl.nq_pad, rps = pl.rps;

    hipError_t e = resid ? ws.qtab.ensure(sizeof(float) * size_t(op.L) * nq_pad) : ws.Qop.ensure(size_t(nq_pad) * Kp * es);
    if (resid && e == hipSuccess) e = ws.qs2.ensure(sizeof(float) * size_t(op.L) * chunk);
    if (kind == 3 && e == hipSuccess) e = ws.qfrag.ensure(size_t(2) * op.L * size_t(nq_pad) * Kp);
    if (kind == 3 && e == hipSuccess) e = ws.qsc.ensure(sizeof(int) * size_t(op.L) * nq_pad);
    if (e == hipSuccess) e = ws.Q32.ensure(size_t(chunk) * ld32 * 4);
    if (e == hipSuccess) e = ws.qnorm.ensure(sizeof(double) * 3 * size_t(chunk));
    if (e == hipSuccess) e = ws.qint.ensure(sizeof(int) * size_t(chunk));
    if (e == hipSuccess) e = ws.slots.ensure(pl.slots_bytes);
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
    qa.bf16 = mk == 0;
    qa.Qop = resid ? nullptr : ws.Qop.p;
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
    if (resid && e == hipSuccess) {
      // ||q - c_j||^2 of every (query, list), then the C-input terms
      KmeansPairwiseArgs<float> pw{};
      pw.P1 = qa.Q32;
      pw.ld1 = ld32;
      pw.m = chunk;
      pw.P2 = op.cents.as<float>();
      pw.ld2 = ix->d;
      pw.n = op.L;
      pw.d = ix->d;
      pw.out = ws.qs2.as<float>();
      pw.ldo = op.L;
      e = launch_kmeans_pairwise<float>(pw, st);
      QListArgs la{};
      la.s2 = ws.qs2.as<float>();
      la.m = chunk;
      la.nq_pad = nq_pad;
      la.d = ix->d;
      la.L = op.L;
      la.fp8 = kind == 3 ? 1 : 0;
      la.lconsts = op.lconsts.as<unsigned long long>();
      la.qtab = ws.qtab.as<float>();
      la.lcount = op.lcount.as<uint32_t>();
      la.KP = KP;
      la.gthr = ws.gthr.as<uint32_t>();
      if (e == hipSuccess) e = launch_prep_qlist(la, st);
      if (kind == 3 && e == hipSuccess) {
        QFragArgs fa{};
        fa.Q32 = qa.Q32;
        fa.ld32 = ld32;
        fa.m = chunk;
        fa.nq_pad = nq_pad;
        fa.d = ix->d;
        fa.cents = op.cents.as<float>();
        fa.ldc = ix->d;
        fa.L = op.L;
        fa.Kp = Kp;
        fa.qfrag = ws.qfrag.as<uint8_t>();
        fa.qsc = ws.qsc.as<int>();
        e = launch_prep_qfrag_fp8(fa, st);
      }
      if (!d_theta && !(ix->opts.flags & FUEL_AMD_FLAG_NO_SEED) && e == hipSuccess) {
        // a first level: exact distances to a sample of each query's nearest list
        SampleSeedArgs sd{};
        sd.s2 = ws.qs2.as<float>();
        sd.m = chunk;
        sd.L = op.L;
        sd.KP = KP;
        sd.lcount = op.lcount.as<uint32_t>();
        sd.loff = op.loff.as<int64_t>();
        sd.rperm = op.rperm.as<uint32_t>();
        sd.X = ix->X.as<float>();
        sd.ldx = ix->ld;
        sd.Q32 = qa.Q32;
        sd.ld32 = ld32;
        sd.d = ix->d;
        sd.gthr = ws.gthr.as<uint32_t>();
        e = launch_seed_sample(sd, st);
      }
    }
    if (e == hipSuccess) e = launch_fill_u32(ws.fail_count.as<uint32_t>(), 1, 0u, st);
    if (e != hipSuccess) return hip_fail(e, "query preparation");

    // Seeded thresholds for a cold start of the plain operands: the KP-th
    // smallest group minimum over an evenly strided sample of the tiles
    // (plan_seed) lowers each query's shared threshold before the scan,
    // so its splits append few rows that a tighter threshold would drop.
    {
      const SeedPlan sp = plan_seed(mk, ks, resid, d_theta != nullptr, ix->opts.flags, n_pad, pl.rows_tile, KP);
      if (sp.ntiles > 0) {
        e = ws.seed.ensure(sizeof(float) * size_t(nq_pad) * sp.parts * 2 * SEED_GROUPS);
        if (e != hipSuccess) return hip_fail(e, "seed workspace");
        SeedArgs sd{};
        sd.Xop = op.Xop.p;
        sd.hx = op.hx.as<float>();
        sd.Kp = Kp;
        sd.Qop = ws.Qop.p;
        sd.nq_pad = nq_pad;
        sd.nqb = (int)nqb;
        sd.stride = sp.stride;
        sd.ntiles = sp.ntiles;
        sd.parts = sp.parts;
        sd.mins = ws.seed.as<float>();
        sd.m = chunk;
        sd.KP = KP;
        sd.gthr = ws.gthr.as<uint32_t>();
        // the seed pass reads the same operand rows and query fragments as the scan
        if (size_t(sp.stride) * (sp.ntiles - 1) + size_t(pl.rows_tile) > size_t(n_pad) ||
            size_t(nq_pad) * Kp * es > ws.Qop.bytes)
          return fail(FUEL_AMD_ERR_DEVICE, "internal error: seed plan outside its buffers");
        e = launch_seed_mfma(sd, mk, ks, st);
        if (e != hipSuccess) return hip_fail(e, "seed launch");
      }
    }

    ScanMfmaArgs sa{};
    sa.Xop = op.Xop.p;
    sa.hx = op.hx.as<float>();
    sa.Kp = Kp;
    sa.n_pad = n_pad;
    sa.Qop = resid ? nullptr : ws.Qop.p;
    sa.nq_pad = nq_pad;
    sa.nqb = (int)nqb;
    sa.rows_per_split = rps;
    sa.KP = KP;
    sa.Csub = csub;
    sa.slots = ws.slots.as<uint64_t>();
    sa.lcnts = ws.lcnts.as<int>();
    sa.gthr = ws.gthr.as<uint32_t>();
    sa.waves = waves;
    if (resid) {
      sa.resid = 1;
      sa.tile_list = op.tile_list.as<int>();
      sa.cents = op.cents.as<float>();
      sa.ldc = ix->d;
      sa.Q32 = qa.Q32;
      sa.ld32 = ld32;
      sa.m = chunk;
      sa.d = ix->d;
      sa.qtab = ws.qtab.as<float>();
      sa.xsc = kind == 3 ? op.xsc.as<uint8_t>() : nullptr;
      sa.qfrag = kind == 3 ? ws.qfrag.as<uint8_t>() : nullptr;
      sa.qsc = kind == 3 ? ws.qsc.as<int>() : nullptr;
      sa.L = op.L;
    }
    {
      // refuse a launch whose addresses would leave the buffers it was given
      ScanExtents ex{};
      ex.xop = op.Xop.bytes;
      ex.hx = op.hx.bytes;
      ex.qop = resid ? 0 : ws.Qop.bytes;
      ex.slots = ws.slots.bytes;
      ex.lcnts = ws.lcnts.bytes;
      ex.gthr = ws.gthr.bytes;
      ex.tile_list = resid ? op.tile_list.bytes : 0;
      ex.cents = resid ? op.cents.bytes : 0;
      ex.q32 = ws.Q32.bytes;
      ex.qtab = resid ? ws.qtab.bytes : 0;
      ex.xsc = kind == 3 ? op.xsc.bytes : 0;
      ex.qfrag = kind == 3 ? ws.qfrag.bytes : 0;
      ex.qsc = kind == 3 ? ws.qsc.bytes : 0;
      std::string why;
      if (scan_check_launch(sa, mk, ks, V, nsplit, ex, resid ? op.L : 0, &why))
        return fail(FUEL_AMD_ERR_DEVICE, "internal error: MFMA scan plan outside its buffers: " + why);
    }
    if (timed)
      timing_mark(ix, st, true,
                  kind == 0   ? "scan_mfma_kernel<bf16>"
                  : kind == 1 ? "scan_mfma_kernel<f32>"
                  : kind == 2 ? "scan_mfma_kernel<bf16r>"
                              : "scan_mfma_kernel<fp8r>");
    e = launch_scan_mfma(sa, mk, ks, V, nsplit, st);
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
    ra.arith = ix->arith;
    ra.bf16 = mk == 0;
    ra.resid = resid ? 1 : 0;
    ra.rperm = resid ? op.rperm.as<uint32_t>() : nullptr;
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
    if (nfail > 0 && !probe) {
      // the fail list holds query rows relative to Qb and to this chunk's outputs
      int next_kind = -1;
      if (level == 0)
        next_kind = kind;
      else if (level == 1 && kind != 1 && mfma_supported(ix, 1, k, exclude))
        next_kind = 1;
      if (next_kind >= 0) {
        // theta is in the key units of this kind's scan: the residual keys
        // (||q - x||^2 / 2) are not those of the f32 form
        const float* theta = (next_kind == kind || !resid) ? ra.fail_theta : nullptr;
        rc = knn_mfma(ix, ctx, Qb, nfail, ldq, k, exclude, next_kind, ra.out_idx, ra.out_dist, st, ra.fail_list,
                      ra.qbase, level + 1, theta ? KP_CHAIN : kp_max(next_kind), theta);
      } else {
        ctx->stats.exact_queries += nfail;  // queries answered by the exact scan
        rc = knn_exact_rows(ix, ctx, Qb, ldq, ra.fail_list, nfail, ra.qbase, k, exclude, ra.out_idx, ra.out_dist,
                            st);
      }
      if (rc) return rc;
    }
    if (level == 0 && !probe && nfail > 0 && ctx->fail_rows.size() < kFailRowsKept) {
      // record the rows of the failed queries (in the caller's row numbering)
      const size_t take = std::min<size_t>(size_t(nfail), kFailRowsKept - ctx->fail_rows.size());
      std::vector<int64_t> h(take);
      e = hipMemcpyAsync(h.data(), ra.fail_list, sizeof(int64_t) * take, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hip_fail(e, "certification list");
      for (int64_t r : h) ctx->fail_rows.push_back(d_rows ? r : done + r);
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

// Every launch plan plan_mfma_chunk makes for a matrix of shapes and sizes,
// checked by scan_check_launch against the buffers knn_mfma allocates for it
// (host only: nothing is allocated or launched).  perturb > 0 plants one
// mistake in every plan (a buffer one element short of what the kernel reads,
// one split too many, an append segment one entry too long, ...), so that the
// check can be seen to catch it.  Returns the number of plans flagged.
int64_t check_scan_plans(int perturb, int64_t* launches, std::string* first) {
  int64_t bad = 0, count = 0;
  const int KPs[] = {10, 11, 16, 20, 40, 64, 65, 128, 160, 256};
  const int64_t ms[] = {1, 63, 64, 65, 192, 193, 257, 4097, 100000};
  const int64_t ns[] = {300, 65536 + 13, 1000000, 10000000};
  const int cus_list[] = {256, 1};
  for (int variant = 0; variant < 4; ++variant) {
    const int mk = variant == 1 ? 1 : (variant == 3 ? 2 : 0);  // variant 3: the E4M3 residual operand
    const bool resid = variant >= 2;
    for (int dd = 1; dd <= (mk == 0 ? 1024 : 256) + 1; ++dd) {
      // the widest and the narrowest dimension of every K step bucket
      const int ks = mfma_ksteps_bucket(mk, dd);
      if (ks < 0 || (resid && ks > mfma_resid_max_ksteps())) continue;
      if (mfma_ksteps_bucket(mk, dd - 1) == ks && mfma_ksteps_bucket(mk, dd + 1) == ks) continue;
      const int d = dd;
      const int Kp = ks * (mk == 0 ? 16 : (mk == 1 ? 8 : 32));
      const size_t es = mk == 0 ? 2 : (mk == 1 ? 4 : 1);
      const int64_t ld32 = (d + 7) & ~int64_t(7);
      for (int64_t n : ns) {
        int L = 0;
        int64_t n_pad = std::max<int64_t>(256, (n + 255) & ~int64_t(255));
        size_t tile_list_bytes = 0;
        if (resid) {
          L = (int)std::max<int64_t>(1, std::min<int64_t>(256, n / 256));
          std::vector<uint32_t> cnt(L, uint32_t(n / L));
          for (int64_t j = 0; j < n % L; ++j) cnt[j] += 1;
          std::vector<int64_t> pad;
          std::vector<int> tl;
          n_pad = resid_layout(cnt, pad, tl);
          tile_list_bytes = tl.size() * sizeof(int);
        }
        for (int KP : KPs)
          for (int64_t m : ms)
            for (int cus : cus_list) {
              if (m * n > int64_t(1) << 40 && cus == 1) continue;
              int64_t done = 0;
              while (done < m) {
                const MfmaChunkPlan pl = plan_mfma_chunk(mk, resid, ks, m, m - done, KP, n_pad, cus);
                ScanMfmaArgs sa{};
                sa.Kp = Kp;
                sa.n_pad = n_pad;
                sa.nq_pad = pl.nq_pad;
                sa.nqb = (int)pl.nqb;
                sa.rows_per_split = pl.rps;
                sa.KP = KP;
                sa.Csub = pl.csub;
                sa.waves = pl.waves;
                ScanExtents ex{};
                ex.xop = size_t(n_pad) * Kp * es;
                ex.hx = size_t(n_pad) * 4;
                ex.qop = resid ? 0 : size_t(pl.nq_pad) * Kp * es;
                ex.slots = pl.slots_bytes;
                ex.lcnts = size_t(pl.nsplit) * pl.nq_pad * 4;
                ex.gthr = size_t(pl.nq_pad) * 4;
                ex.q32 = size_t(pl.chunk) * ld32 * 4;
                if (resid) {
                  sa.resid = 1;
                  sa.ldc = d;
                  sa.ld32 = ld32;
                  sa.m = pl.chunk;
                  sa.d = d;
                  ex.tile_list = tile_list_bytes;
                  ex.cents = size_t(L) * d * 4;
                  ex.qtab = size_t(L) * pl.nq_pad * 4;
                }
                if (mk == 2) {
                  sa.L = L;
                  ex.xsc = size_t(n_pad);
                  ex.qfrag = size_t(2) * L * pl.nq_pad * Kp;
                  ex.qsc = size_t(L) * pl.nq_pad * 4;
                }
                int nsplit = pl.nsplit;
                const size_t need_slots = size_t(pl.nsplit) * pl.nq_pad * pl.NS * 8;
                switch (perturb) {
                  case 1: ex.slots = need_slots - 8; break;
                  case 2: ex.lcnts -= 4; break;
                  case 3: ex.gthr -= 4; break;
                  case 4: nsplit += 1; break;
                  case 5: sa.Csub += 1; break;
                  case 6: ex.xop -= 16; break;
                  case 7: sa.rows_per_split += 1; break;
                  case 8: sa.waves = sa.waves == 1 ? 4 : 1; break;
                  case 9: ex.hx -= 4; break;
                  case 10:
                    if (resid) ex.qtab -= 4; else ex.qop -= 16;
                    break;
                  case 11:
                    if (resid) ex.q32 = size_t(pl.chunk - 1) * ld32 * 4 + size_t(d) * 4 - 4; else ex.gthr -= 4;
                    break;
                  case 12:
                    if (resid) ex.cents -= 4; else ex.lcnts -= 4;
                    break;
                  case 13:  // E4M3: row scales, fragments or their scales short; others as case 1
                    if (mk == 2) {
                      const int which = int(count % 3);
                      if (which == 0) ex.xsc -= 1; else if (which == 1) ex.qfrag -= 16; else ex.qsc -= 4;
                    } else {
                      ex.slots = need_slots - 8;
                    }
                    break;
                  default: break;
                }
                std::string why;
                const int64_t v = scan_check_launch(sa, mk, ks, pl.V, nsplit, ex, L, &why) ? 1 : 0;
                if (v && bad == 0 && first) {
                  char buf[160];
                  std::snprintf(buf, sizeof(buf), "d %d, n %lld, m %lld (chunk at %lld), KP %d, %d CUs: ", d,
                                (long long)n, (long long)m, (long long)done, KP, cus);
                  *first = buf + why;
                }
                bad += v;
                ++count;
                done += pl.chunk;
              }
            }
      }
    }
  }
  if (launches) *launches = count;
  return bad;
}

}  // namespace fa

// Test hook (not part of the ABI in include/fuel_amd.h): runs check_scan_plans
// and returns its violation count; `text` receives the first violation.
extern "C" int64_t fuel_amd_internal_check_scan_plans(int32_t perturb, int64_t* launches, char* text, int32_t len) {
  std::string first;
  const int64_t bad = fa::check_scan_plans(perturb, launches, &first);
  if (text && len > 0) {
    std::snprintf(text, size_t(len), "%s", first.c_str());
  }
  return bad;
}
