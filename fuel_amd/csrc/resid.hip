// Residual (centroid-relative) operand of the bf16 candidate scan.
//
// On clustered data ||x|| and ||q|| are large next to the neighbour distances,
// and the bf16 bound of the plain form, ~2^-9 ||q|| ||x||, swamps the gaps
// between them.  Rows are partitioned around L centroids (a few Lloyd steps),
// sorted by list, and each list padded to whole 64-row blocks.  For a row x of
// list j with r = x - c_j,
//     ||q - x||^2 / 2 = ||q - c_j||^2 / 2 - (q - c_j).r + ||r||^2 / 2,
// so the MFMA contracts bf16(q - c_j) with bf16(r): the rounding error scales
// with ||q - c_j|| ||r||, the distances to the centroid, instead of ||q|| ||x||.
// ||r||^2 / 2 rides in the per-row C input as before and ||q - c_j||^2 / 2,
// minus a proven bound on every other error, in a per-(list, query) term, so
// the scan's keys are lower bounds of ||q - x||^2 / 2 (up to the row tag).
// The partition only affects speed: every answer is still certified against
// the exact restated distances.
#include "fp8.h"
#include "kernels.h"

namespace fa {

namespace {

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

constexpr int ASSIGN_CC = 16;     // centroids per LDS chunk
constexpr int ASSIGN_MAXD = 1024;  // largest dimension of the bf16 scan

}  // namespace

// One thread per row; centroids staged through LDS 16 at a time (zero-padded
// to a multiple of 16 coordinates, as are the row's coordinates).
__global__ __launch_bounds__(256) void assign_kernel(AssignArgs a) {
  __shared__ float cs[ASSIGN_CC * ASSIGN_MAXD];
  const int64_t row = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int dp = (a.d + 15) & ~15;
  float best = INFINITY;
  uint32_t bi = 0;
  for (int c0 = 0; c0 < a.L; c0 += ASSIGN_CC) {
    const int cc = a.L - c0 < ASSIGN_CC ? a.L - c0 : ASSIGN_CC;
    __syncthreads();
    for (int e = threadIdx.x; e < ASSIGN_CC * dp; e += 256) {
      const int c = e / dp, k = e - c * dp;
      cs[e] = (c < cc && k < a.d) ? a.cents[size_t(c0 + c) * a.d + k] : 0.0f;
    }
    __syncthreads();
    if (row < a.n) {
      float acc[ASSIGN_CC];
#pragma unroll
      for (int c = 0; c < ASSIGN_CC; ++c) acc[c] = 0.0f;
      const float* x = a.X + row * a.ldx;
      for (int k0 = 0; k0 < dp; k0 += 16) {
        float xv[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) xv[e] = k0 + e < a.d ? x[k0 + e] : 0.0f;
#pragma unroll
        for (int c = 0; c < ASSIGN_CC; ++c)
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const float t = xv[e] - cs[c * dp + k0 + e];
            acc[c] = fmaf(t, t, acc[c]);
          }
      }
      for (int c = 0; c < cc; ++c)
        if (acc[c] < best) {
          best = acc[c];
          bi = uint32_t(c0 + c);
        }
    }
  }
  if (row < a.n) {
    a.list[row] = bi;
    atomicAdd(&a.counts[bi], 1u);
  }
}

hipError_t launch_assign(const AssignArgs& a, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
  if (a.d > ASSIGN_MAXD || a.L < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(assign_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// One workgroup per list; thread t sums coordinates t, t + 256, ... over the
// list's rows in sorted order (f64).
__global__ __launch_bounds__(256) void list_means_kernel(const float* X, int64_t ldx, int d,
                                                         const uint32_t* sorted_rows, const int64_t* offsets,
                                                         const uint32_t* counts, float* cents) {
  const int j = blockIdx.x;
  const uint32_t cnt = counts[j];
  if (cnt == 0) return;
  const int64_t off = offsets[j];
  for (int k = threadIdx.x; k < d; k += 256) {
    double s = 0.0;
    for (uint32_t i = 0; i < cnt; ++i) s += (double)X[int64_t(sorted_rows[off + i]) * ldx + k];
    cents[size_t(j) * d + k] = (float)(s / (double)cnt);
  }
}

hipError_t launch_list_means(const float* X, int64_t ldx, int d, const uint32_t* sorted_rows,
                             const int64_t* offsets, const uint32_t* counts, int L, float* cents,
                             hipStream_t st) {
  if (L <= 0 || d <= 0) return hipSuccess;
  hipLaunchKernelGGL(list_means_kernel, dim3((unsigned)L), dim3(256), 0, st, X, ldx, d, sorted_rows, offsets, counts,
                     cents);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void scatter_rperm_kernel(const uint32_t* sorted_rows, const uint32_t* sorted_list,
                                                            int64_t n, const int64_t* offsets,
                                                            const int64_t* pad_offsets, uint32_t* rperm) {
  const int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (s >= n) return;
  const uint32_t j = sorted_list[s];
  rperm[pad_offsets[j] + (s - offsets[j])] = sorted_rows[s];
}

hipError_t launch_scatter_rperm(const uint32_t* sorted_rows, const uint32_t* sorted_list, int64_t n,
                                const int64_t* offsets, const int64_t* pad_offsets, uint32_t* rperm,
                                hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_rperm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sorted_rows,
                     sorted_list, n, offsets, pad_offsets, rperm);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void iota_u32_kernel(uint32_t* p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

hipError_t launch_iota_u32(uint32_t* p, int64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(iota_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m));
  return v;
}

// One workgroup per 64-row block (every block lies in one list): each wave
// prepares 16 positions, the per-list maxima are reduced in the workgroup and
// published with one atomic per constant (a per-row atomic on the few
// addresses of a list serialises).  FP8: position p holds the E4M3 codes of
// fl32(x - c_j) 2^-e_p, e_p = fp8_scale_exp of the row's largest |coordinate|,
// and xsc[p] the E8M0 code of 2^e_p; r~ is the dequantized row.
template <bool FP8>
__device__ __forceinline__ void prep_resid_body(const PrepResidArgs& a, const int64_t blk) {
  __shared__ double red[4][6];
  const int wave = threadIdx.x >> 6;
  const int lane = __lane_id();
  const int j = a.tile_list[blk];
  const float* cj = a.cents + size_t(j) * a.ldc;
  const double slack = 1.0 + (a.d + 6) * 0x1p-52;
  double m[6] = {0, 0, 0, 0, 0, 0};  // ||r||^2, ||r~-r||^2, ||r~||^2, hxn error, fl32(||r||^2/2), ||x||^2
  for (int i = 0; i < 16; ++i) {
    const int64_t pos = blk * 64 + wave * 16 + i;
    uint16_t* dst = static_cast<uint16_t*>(a.Xop) + size_t(pos) * a.Kp;
    uint8_t* dst8 = static_cast<uint8_t*>(a.Xop) + size_t(pos) * a.Kp;
    const uint32_t row = a.rperm[pos];
    if (row == RPERM_NONE) {
      for (int c = lane; c < a.Kp; c += 64) {
        if constexpr (FP8)
          dst8[c] = 0;
        else
          dst[c] = 0;
      }
      if (lane == 0) {
        a.hx[pos] = MFMA_PAD_HXN;
        if constexpr (FP8) a.xsc[pos] = (uint8_t)fp8_scale_code(0);
      }
      continue;
    }
    const float* x = a.X + int64_t(row) * a.ldx;
    int e = 0;
    if constexpr (FP8) {
      float mx = 0.0f;
      for (int c = lane; c < a.d; c += 64) mx = fmaxf(mx, fabsf((float)((double)x[c] - (double)cj[c])));
      e = fp8_scale_exp(wave_max(mx));
      if (lane == 0) a.xsc[pos] = (uint8_t)fp8_scale_code(e);
    }
    double s2 = 0.0, e2 = 0.0, t2 = 0.0, x2 = 0.0;
    for (int c = lane; c < a.Kp; c += 64) {
      const double xv = c < a.d ? (double)x[c] : 0.0;
      x2 += xv * xv;
      const double r = c < a.d ? xv - (double)cj[c] : 0.0;
      double rt;
      if constexpr (FP8) {
        const uint32_t code = fp8_e4m3_encode((float)r, e);
        dst8[c] = (uint8_t)code;
        rt = ldexp((double)fp8_e4m3_value(code), e);
      } else {
        const uint16_t h = bf16_rne((float)r);
        dst[c] = h;
        rt = (double)bf16_val(h);
      }
      s2 += r * r;
      e2 += (rt - r) * (rt - r);
      t2 += rt * rt;
    }
    s2 = wave_sum(s2);
    e2 = wave_sum(e2);
    t2 = wave_sum(t2);
    x2 = wave_sum(x2);
    const double v = s2 * 0.5;
    const float hf = (float)v;
    if (lane == 0) a.hx[pos] = -hf;
    m[0] = fmax(m[0], s2 * slack);
    m[1] = fmax(m[1], e2 * slack);
    m[2] = fmax(m[2], t2 * slack);
    m[3] = fmax(m[3], fabs((double)hf - v) + v * (a.d + 6) * 0x1p-52);
    m[4] = fmax(m[4], (double)hf);
    m[5] = fmax(m[5], x2 * slack);
  }
  if (lane == 0)
    for (int t = 0; t < 6; ++t) red[wave][t] = m[t];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int t = threadIdx.x;
    const double v = fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t]));
    if (v > 0.0) {
      if (t < 5)
        atomic_max_nonneg(&a.lconsts[size_t(j) * 8 + t], v);
      else
        atomic_max_nonneg(&a.gconsts[0], v);
    }
  }
}

template <bool FP8>
__global__ __launch_bounds__(256) void prep_resid_kernel(PrepResidArgs a) {
  prep_resid_body<FP8>(a, blockIdx.x);
}
__global__ __launch_bounds__(256) void prep_resid_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const PrepResidArgs a = static_cast<const PrepResidArgs*>(g.tab)[p];
  prep_resid_body<false>(a, int64_t(blockIdx.x) - g.start[p]);
}

hipError_t launch_prep_resid_group(const GroupTab& g, hipStream_t st) {
  if (g.np <= 0 || g.nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(prep_resid_group_kernel, dim3((unsigned)g.nblocks), dim3(256), 0, st, g);
  return hipGetLastError();
}

hipError_t launch_prep_resid(const PrepResidArgs& a, hipStream_t st) {
  if (a.n_rpad == 0) return hipSuccess;
  if (a.n_rpad % 64 != 0) return hipErrorInvalidValue;  // lists are whole 64-row blocks
  if (a.fp8 && (!a.xsc || a.Kp % 64 != 0)) return hipErrorInvalidValue;  // whole 64-byte MFMA steps
  if (a.fp8)
    hipLaunchKernelGGL(prep_resid_kernel<true>, dim3((unsigned)(a.n_rpad / 64)), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(prep_resid_kernel<false>, dim3((unsigned)(a.n_rpad / 64)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// One thread per (list, query slot).  s^ = the f32 value of ||q - c||^2 (a
// sum of d rounded squares of rounded differences, fused or not): |s^ - s| <=
// g s + (d + 1) 2^-126 with g = (d + 3) u / (1 - (d + 3) u).  With N = ||q - c||
// (upper bound), the B operand b~ = bf16(fl32(q - c)) has ||b~ - (q - c)|| <=
// EA = N (2^-8 + 2^-23) + sqrt(d) 2^-126 (bf16 keeps 8 significant bits, so
// round-to-nearest errs by up to 2^-8 of a coordinate, and the f32 difference
// by 2^-24; the last term covers differences below the smallest normal, where
// bf16 has fewer bits and the matrix core may flush them) and ||b~|| <= N + EA;
// per list EX = max ||r~ - r|| (measured), plus sqrt(d) 2^-126 for the same
// flush of a row coordinate,
// RN = max ||r||, RT = max ||r~||, EH = max |fl32(||r||^2/2) - ||r||^2/2|,
// H = max fl32(||r||^2/2).
//   operand rounding   N EX + EA RN + EA EX
//   hxn rounding       EH
//   accumulation       4 (d + 2) u (|b~| RT + H + Q2 + e)   (C input included)
//   C-input addition   u (H + Q2 + e)
//   underflow          2 (d + 2) 2^-126
// e solves e >= base + (4 (d + 2) u + u) e; the term is -Q2lo + e rounded up,
// Q2lo <= ||q - c||^2 / 2 <= Q2.
// FP8 operand: b~ = hi 2^e + lo 2^(e-4) with hi = e4m3(dq 2^-e), lo =
// e4m3((dq - hi 2^e) 2^-(e-4)), dq = fl32(q - c) and 2^e <= max |dq| / 224
// (or 2^-100).  Per coordinate |dq - hi 2^e| <= 2^-4 |dq| + 2^(e-10) and the lo
// rounding adds 2^-4 of that plus 2^(e-14), so |dq - b~| <= 2^-8 |dq| + 2^(e-13):
//   EA = N (2^-8 + 2^-23 + sqrt(d) 2^-13 / 224) + sqrt(d) 2^-112
// (2^-23 covers the f32 subtraction; the last term a clamped scale and flushed
// subnormal differences).  The MFMA sums 2d products against hi and lo, with
// sum |a||b| <= RT (||hi|| + ||lo||) <= RT (N + EA)(1 + 2^-3).
__device__ __forceinline__ void prep_qlist_body(const QListArgs& a, const int64_t blk) {
  const int64_t w = blk * 256 + threadIdx.x;
  if (w >= int64_t(a.L) * a.nq_pad) return;
  const int j = (int)(w / a.nq_pad);
  const int64_t q = w - int64_t(j) * a.nq_pad;
  const double sh = (double)a.s2[(q < a.m ? q : a.m - 1) * a.L + j];
  const double u = 0x1p-24;
  const int d = a.d;
  const double g = (d + 3) * u / (1.0 - (d + 3) * u) * (1 + 0x1p-20);
  const double tiny = (d + 1) * 0x1p-126;
  const double s_hi = (sh * (1 + g) + tiny) * (1 + 0x1p-50);
  const double s_lo = fmax(0.0, (sh - g * sh - tiny) * (1 - 0x1p-50));
  const unsigned long long* lc = a.lconsts + size_t(j) * 8;
  const double RN = sqrt(__longlong_as_double((long long)lc[0])) * (1 + 0x1p-50);
  const double sd = sqrt((double)d) * (1 + 0x1p-50);
  const double EX = sqrt(__longlong_as_double((long long)lc[1])) * (1 + 0x1p-50) + (a.fp8 ? 0.0 : sd * 0x1p-126);
  const double RT = sqrt(__longlong_as_double((long long)lc[2])) * (1 + 0x1p-50);
  const double EH = __longlong_as_double((long long)lc[3]);
  const double H = __longlong_as_double((long long)lc[4]);
  const double N = sqrt(s_hi) * (1 + 0x1p-50);
  const double EA = a.fp8 ? (N * (0x1p-8 + 0x1p-23 + sd * 0x1p-13 / 224.0) + sd * 0x1p-112) * (1 + 0x1p-40)
                          : (N * (0x1p-8 + 0x1p-23) + sd * 0x1p-126) * (1 + 0x1p-40);
  const double B = a.fp8 ? (N + EA) * (1 + 0x1p-3) * (1 + 0x1p-40) : (N + EA) * (1 + 0x1p-50);
  const double Q2 = 0.5 * s_hi;
  const double Q2lo = 0.5 * s_lo;
  const int terms = a.fp8 ? 2 * d : d;
  const double gacc = 4.0 * (terms + 2) * u;
  const double base = N * EX + EA * RN + EA * EX + EH + gacc * (B * RT + H + Q2) + u * (H + Q2) +
                      2.0 * (terms + 2) * 0x1p-126;
  const double e = base / (1.0 - gacc - u) * (1 + 0x1p-40) + 0x1p-140;
  const double t = -Q2lo + e;
  float tf = (float)t;
  if ((double)tf < t) tf = nextafterf(tf, INFINITY);
  a.qtab[w] = tf;
}

// One wave per query slot, lanes over the lists (see QListArgs).
__device__ __forceinline__ void seed_theta_body(const QListArgs& a, const int64_t blk) {
  const int64_t q = blk * 4 + (threadIdx.x >> 6);
  if (q >= a.nq_pad) return;
  const int lane = __lane_id();
  const float* s2 = a.s2 + (q < a.m ? q : a.m - 1) * a.L;
  const double g = (a.d + 3) * 0x1p-24 / (1.0 - (a.d + 3) * 0x1p-24) * (1 + 0x1p-20);
  const double tiny = (a.d + 1) * 0x1p-126;
  double best = INFINITY;
  for (int j = lane; j < a.L; j += 64) {
    if (a.lcount[j] < (uint32_t)a.KP) continue;
    const double sh = (double)s2[j];
    const double N = sqrt((sh * (1 + g) + tiny) * (1 + 0x1p-50)) * (1 + 0x1p-50);
    const double RN = sqrt(__longlong_as_double((long long)a.lconsts[size_t(j) * 8])) * (1 + 0x1p-50);
    const double v = 0.5 * (N + RN) * (N + RN) * (1 + 0x1p-16) + 0x1p-140;
    best = fmin(best, v);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) best = fmin(best, __shfl_xor(best, m));
  if (lane == 0 && best < (double)MFMA_KEY_LIMIT) {
    float t = (float)best;
    if ((double)t < best) t = nextafterf(t, INFINITY);
    const uint32_t o = ord_f32(t);
    if (o < a.gthr[q]) a.gthr[q] = o;
  }
}

__global__ __launch_bounds__(256) void prep_qlist_kernel(QListArgs a) { prep_qlist_body(a, blockIdx.x); }
__global__ __launch_bounds__(256) void seed_theta_kernel(QListArgs a) { seed_theta_body(a, blockIdx.x); }
__global__ __launch_bounds__(256) void prep_qlist_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const QListArgs a = static_cast<const QListArgs*>(g.tab)[p];
  prep_qlist_body(a, int64_t(blockIdx.x) - g.start[p]);
}
__global__ __launch_bounds__(256) void seed_theta_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const QListArgs a = static_cast<const QListArgs*>(g.tab)[p];
  seed_theta_body(a, int64_t(blockIdx.x) - g.start[p]);
}

__global__ __launch_bounds__(256) void prep_qfrag_fp8_kernel(QFragArgs a) {
  const int64_t w = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (w >= int64_t(a.L) * a.nq_pad) return;
  const int lane = __lane_id();
  const int j = (int)(w / a.nq_pad);
  const int64_t q = w - int64_t(j) * a.nq_pad;
  const float* qv = a.Q32 + (q < a.m ? q : a.m - 1) * a.ld32;
  const float* cj = a.cents + size_t(j) * a.ldc;
  float mx = 0.0f;
  for (int c = lane; c < a.d; c += 64) mx = fmaxf(mx, fabsf(qv[c] - cj[c]));
  const int e = fp8_scale_exp(wave_max(mx));
  uint8_t* hi = a.qfrag + size_t(w) * a.Kp;
  uint8_t* lo = hi + size_t(a.L) * a.nq_pad * a.Kp;
  for (int c = lane; c < a.Kp; c += 64) {
    const float v = c < a.d ? qv[c] - cj[c] : 0.0f;
    const uint32_t ch = fp8_e4m3_encode(v, e);
    const float res = v - ldexpf(fp8_e4m3_value(ch), e);  // exact: a multiple of v's unit, |res| <= |v|
    hi[c] = (uint8_t)ch;
    lo[c] = (uint8_t)fp8_e4m3_encode(res, e - 4);
  }
  if (lane == 0) a.qsc[w] = (int)fp8_scale_code(e);
}

hipError_t launch_prep_qfrag_fp8(const QFragArgs& a, hipStream_t st) {
  const int64_t cnt = int64_t(a.L) * a.nq_pad;
  if (cnt == 0 || a.m == 0) return hipSuccess;
  if (a.Kp % 64 != 0 || a.d > a.Kp) return hipErrorInvalidValue;
  hipLaunchKernelGGL(prep_qfrag_fp8_kernel, dim3((unsigned)((cnt + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// One workgroup of four waves per query: every wave finds the nearest list
// (the same answer), wave w bounds rows w * 128 + lane and + 64 of its sample
// (float4 loads, f64 sums in four chains), the bounds meet in LDS and wave 0
// selects the KP-th smallest.
__device__ __forceinline__ void seed_sample_body(const SampleSeedArgs& a, const int64_t q) {
  constexpr int RPL = SEED_SAMPLE_ROWS / 256;  // sample rows per thread
  constexpr int RW = SEED_SAMPLE_ROWS / 64;    // sample rows per lane of wave 0's select
  __shared__ uint32_t keys[SEED_SAMPLE_ROWS];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const float* s2 = a.s2 + q * a.L;
  float best = INFINITY;
  int bj = a.L;  // none
  for (int j = lane; j < a.L; j += 64)
    if (a.lcount[j] >= (uint32_t)a.KP && s2[j] < best) {
      best = s2[j];
      bj = j;
    }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const float ov = __shfl_xor(best, m);
    const int oj = __shfl_xor(bj, m);
    if (ov < best || (ov == best && oj < bj)) {
      best = ov;
      bj = oj;
    }
  }
  if (bj >= a.L) return;  // uniform over the workgroup: no list holds KP rows
  const uint32_t cnt = a.lcount[bj];
  const int S = cnt < (uint32_t)SEED_SAMPLE_ROWS ? (int)cnt : SEED_SAMPLE_ROWS;
  const int64_t p0 = a.loff[bj];  // positions p0 .. p0 + cnt - 1 are the list's rows
  const float* qv = a.Q32 + q * a.ld32;
  const int d4 = a.d & ~3;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int p = wave * (64 * RPL) + r * 64 + lane;
    uint32_t o = 0xFFFFFFFFu;
    if (p < S) {
      const float* x = a.X + int64_t(a.rperm[p0 + p]) * a.ldx;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (int c = 0; c < d4; c += 4) {  // rows and queries are 16-byte aligned (pitch a multiple of 8)
        const float4 xv = *reinterpret_cast<const float4*>(x + c);
        const float4 qq = *reinterpret_cast<const float4*>(qv + c);
        const double t0 = (double)qq.x - (double)xv.x, t1 = (double)qq.y - (double)xv.y;
        const double t2 = (double)qq.z - (double)xv.z, t3 = (double)qq.w - (double)xv.w;
        s[0] = fma(t0, t0, s[0]);
        s[1] = fma(t1, t1, s[1]);
        s[2] = fma(t2, t2, s[2]);
        s[3] = fma(t3, t3, s[3]);
      }
      for (int c = d4; c < a.d; ++c) {
        const double t = (double)qv[c] - (double)x[c];
        s[0] = fma(t, t, s[0]);
      }
      // an upper bound of the row's scan key (f64 sums in any order err by far
      // less than the 2^-16 slack)
      const double v = 0.5 * ((s[0] + s[1]) + (s[2] + s[3])) * (1 + 0x1p-16) + 0x1p-140;
      float f = (float)v;
      if ((double)f < v) f = nextafterf(f, INFINITY);
      o = ord_f32(f);
    }
    keys[wave * (64 * RPL) + r * 64 + lane] = o;
  }
  __syncthreads();
  if (wave != 0) return;
  uint32_t key[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) key[r] = keys[r * 64 + lane];
  uint32_t p = 0;  // the KP-th smallest bound: smallest t with #(key <= t) >= KP
  for (int bit = 31; bit >= 0; --bit) {
    const uint32_t probe = p | ((1u << bit) - 1u);
    int c = 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) c += __popcll(__ballot(key[r] <= probe));
    if (c < a.KP) p |= (1u << bit);
  }
  const float t = unord_f32(p);
  if (lane == 0 && t < MFMA_KEY_LIMIT) {
    const uint32_t o = ord_f32(nextafterf(t, INFINITY));
    if (o < a.gthr[q]) a.gthr[q] = o;
  }
}

__global__ __launch_bounds__(256) void seed_sample_kernel(SampleSeedArgs a) { seed_sample_body(a, blockIdx.x); }
__global__ __launch_bounds__(256) void seed_sample_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const SampleSeedArgs a = static_cast<const SampleSeedArgs*>(g.tab)[p];
  seed_sample_body(a, int64_t(blockIdx.x) - g.start[p]);
}

hipError_t launch_seed_sample_group(const GroupTab& g, hipStream_t st) {
  if (g.np <= 0 || g.nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(seed_sample_group_kernel, dim3((unsigned)g.nblocks), dim3(256), 0, st, g);
  return hipGetLastError();
}

hipError_t launch_prep_qlist_group(const GroupTab& terms, const GroupTab& theta, hipStream_t st) {
  if (terms.np > 0 && terms.nblocks > 0)
    hipLaunchKernelGGL(prep_qlist_group_kernel, dim3((unsigned)terms.nblocks), dim3(256), 0, st, terms);
  if (theta.np > 0 && theta.nblocks > 0)
    hipLaunchKernelGGL(seed_theta_group_kernel, dim3((unsigned)theta.nblocks), dim3(256), 0, st, theta);
  return hipGetLastError();
}

hipError_t launch_seed_sample(const SampleSeedArgs& a, hipStream_t st) {
  if (a.m <= 0 || a.L <= 0) return hipSuccess;
  if (a.KP < 1 || a.KP > SEED_SAMPLE_ROWS) return hipErrorInvalidValue;
  if ((a.ldx & 3) || (a.ld32 & 3)) return hipErrorInvalidValue;  // float4 rows
  hipLaunchKernelGGL(seed_sample_kernel, dim3((unsigned)a.m), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_prep_qlist(const QListArgs& a, hipStream_t st) {
  const int64_t cnt = int64_t(a.L) * a.nq_pad;
  if (cnt == 0 || a.m == 0) return hipSuccess;
  hipLaunchKernelGGL(prep_qlist_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, a);
  if (a.gthr && a.lcount)
    hipLaunchKernelGGL(seed_theta_kernel, dim3((unsigned)((a.nq_pad + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace fa
