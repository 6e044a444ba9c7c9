// Preparation of the MFMA operands and of the quantities the error bound needs.
//
// DB rows (once per index): the bf16 operand copy x~ (round to nearest even),
// hx = fl32(||x||^2 / 2) (+inf for padding rows, which then never qualify), and
// index-wide maxima of ||x||^2, ||x~ - x||^2 and ||x~||^2 in f64, plus flags
// "all coordinates are integers of magnitude <= 256" and "all >= 0".
// Queries (per call): the operand copy q~, an aligned f32 copy for the exact
// rerank, and ||q||^2, ||q~ - q||^2, ||q~||^2, the integer flag.
// Sums of squares of f32 values are accumulated in f64; each product is exact
// in f64 and the d-term sum has relative error below d * 2^-53, which the
// bound's slack covers.
#include "kernels.h"

namespace fa {

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

// Row statistics of one f32 row (one wave); writes the operand row.
struct RowStats {
  double s2, e2, t2;
  bool isint, nonneg;
};

__device__ __forceinline__ RowStats prep_row(const float* __restrict__ src, int d, int Kp, int bf16, void* dst) {
  const int lane = __lane_id();
  double s2 = 0.0, e2 = 0.0, t2 = 0.0;
  bool isint = true, nonneg = true;
  for (int c = lane; c < Kp; c += 64) {
    const float v = c < d ? src[c] : 0.0f;
    float vt = v;
    if (bf16) {
      const uint16_t h = bf16_rne(v);
      reinterpret_cast<uint16_t*>(dst)[c] = h;
      vt = bf16_val(h);
    } else {
      reinterpret_cast<float*>(dst)[c] = v;
    }
    const double dv = (double)v, dt = (double)vt;
    s2 += dv * dv;
    e2 += (dt - dv) * (dt - dv);
    t2 += dt * dt;
    isint = isint && (v == rintf(v)) && fabsf(v) <= 256.0f;
    nonneg = nonneg && v >= 0.0f;
  }
  RowStats r;
  r.s2 = wave_sum(s2);
  r.e2 = wave_sum(e2);
  r.t2 = wave_sum(t2);
  r.isint = __all(isint);
  r.nonneg = __all(nonneg);
  return r;
}

__global__ __launch_bounds__(256) void prep_db_kernel(PrepDbArgs a) {
  const int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (row >= a.n_pad) return;
  const int lane = __lane_id();
  const size_t es = a.bf16 ? 2 : 4;
  char* dst = reinterpret_cast<char*>(a.Xop) + size_t(row) * a.Kp * es;
  if (row >= a.n) {
    for (int c = lane; c < a.Kp; c += 64) {
      if (a.bf16)
        reinterpret_cast<uint16_t*>(dst)[c] = 0;
      else
        reinterpret_cast<float*>(dst)[c] = 0.0f;
    }
    if (lane == 0) {
      a.hx[row] = INFINITY;
      // padding rows: -inf in the norm column makes their MFMA result -inf, so
      // their key (+inf) never passes a threshold
      if (a.bf16) reinterpret_cast<uint16_t*>(dst)[a.d] = 0xFF80;
    }
    return;
  }
  const RowStats r = prep_row(a.X + row * a.ldx, a.d, a.Kp, a.bf16, dst);
  if (lane == 0) {
    const double v = r.s2 * 0.5;
    a.hx[row] = (float)v;
    const double slack = 1.0 + (a.d + 4) * 0x1p-52;
    atomic_max_nonneg(&a.consts[0], r.s2 * slack);
    atomic_max_nonneg(&a.consts[1], r.e2 * slack);
    atomic_max_nonneg(&a.consts[2], r.t2 * slack);
    if (a.bf16) {
      // -||x||^2/2 as three bf16 pieces (exact for values of <= 24 significant
      // bits); the queries hold 1 in these columns.
      uint16_t* h = reinterpret_cast<uint16_t*>(dst) + a.d;
      const uint16_t h1 = bf16_rne((float)-v);
      const double r1 = -v - (double)bf16_val(h1);
      const uint16_t h2 = bf16_rne((float)r1);
      const double r2 = r1 - (double)bf16_val(h2);
      const uint16_t h3 = bf16_rne((float)r2);
      h[0] = h1;
      h[1] = h2;
      h[2] = h3;
      const double hs = (double)bf16_val(h1) + (double)bf16_val(h2) + (double)bf16_val(h3);
      // |split - ||x||^2/2| including the f64 error of the sum of squares
      const double err = fabs(hs + v) + (v > 0 ? v : 0) * (a.d + 4) * 0x1p-53;
      atomic_max_nonneg(&a.consts[3], err);
      atomic_max_nonneg(&a.consts[4], fabs((double)bf16_val(h1)) + fabs((double)bf16_val(h2)) +
                                          fabs((double)bf16_val(h3)));
    }
    if (!r.isint) atomicAnd(&a.flags[0], 0);
    if (!r.nonneg) atomicAnd(&a.flags[1], 0);
  }
}

__global__ __launch_bounds__(256) void prep_query_kernel(PrepQArgs a) {
  const int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (row >= a.m_pad) return;
  const int lane = __lane_id();
  const size_t es = a.bf16 ? 2 : 4;
  char* dst = reinterpret_cast<char*>(a.Qop) + size_t(row) * a.Kp * es;
  if (lane == 0 && a.gthr) a.gthr[row] = 0xFF800000u;  // order-preserving bits of +inf
  if (row >= a.m) {
    for (int c = lane; c < a.Kp; c += 64) {
      if (a.bf16)
        reinterpret_cast<uint16_t*>(dst)[c] = 0;
      else
        reinterpret_cast<float*>(dst)[c] = 0.0f;
    }
    return;
  }
  const float* src = a.Q + (a.rows ? a.rows[row] : row) * a.ldq;
  for (int c = lane; c < a.ld32; c += 64) a.Q32[row * a.ld32 + c] = c < a.d ? src[c] : 0.0f;
  const RowStats r = prep_row(src, a.d, a.Kp, a.bf16, dst);
  if (lane == 0 && a.bf16) {  // 1 in the three norm columns
    uint16_t* h = reinterpret_cast<uint16_t*>(dst) + a.d;
    h[0] = 0x3F80;
    h[1] = 0x3F80;
    h[2] = 0x3F80;
  }
  if (lane == 0) {  // raw f64 sums; the bound applies directed slack
    a.qn2[row] = r.s2;
    a.eq2[row] = r.e2;
    a.qt2[row] = r.t2;
    a.qint[row] = r.isint ? 1 : 0;
  }
}

hipError_t launch_prep_db(const PrepDbArgs& a, hipStream_t st) {
  if (a.n_pad == 0) return hipSuccess;
  hipLaunchKernelGGL(prep_db_kernel, dim3((unsigned)((a.n_pad + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_prep_queries(const PrepQArgs& a, hipStream_t st) {
  if (a.m_pad == 0) return hipSuccess;
  hipLaunchKernelGGL(prep_query_kernel, dim3((unsigned)((a.m_pad + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace fa
