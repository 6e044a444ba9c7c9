// Preparation of the MFMA operands and of the quantities the error bound needs.
//
// DB rows (once per index and operand kind): the operand copy x~ (bf16 by round
// to nearest even, or the f32 values), hxn = -fl32(||x||^2 / 2) (MFMA_PAD_HXN
// for padding rows, whose result then never beats a threshold),
// and index-wide maxima of ||x||^2, ||x~ - x||^2, ||x~||^2, |fl32(||x||^2/2) -
// ||x||^2/2| and |fl32(||x||^2/2)| in f64, plus the flags "all coordinates are
// integers of magnitude <= 256" and "all >= 0".
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
      if (dst) reinterpret_cast<uint16_t*>(dst)[c] = h;
      vt = bf16_val(h);
    } else if (dst) {
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

// One row per thread, for operand rows of at most 64 columns (there one wave
// per row leaves most lanes idle): the row's Kp values in steps of 4 (source
// rows 16-byte aligned, pitches multiples of 4), the sums in f64 in column
// order (any order: the bound's slack covers the f64 error of a d-term sum).
__device__ __forceinline__ RowStats prep_row_thread(const float* __restrict__ src, int d, int Kp, int bf16,
                                                    void* dst) {
  double s2 = 0.0, e2 = 0.0, t2 = 0.0;
  bool isint = true, nonneg = true;
  for (int c0 = 0; c0 < Kp; c0 += 4) {
    float v[4];
    if (c0 + 4 <= d) {
      const float4 f = *reinterpret_cast<const float4*>(src + c0);
      v[0] = f.x;
      v[1] = f.y;
      v[2] = f.z;
      v[3] = f.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = c0 + e < d ? src[c0 + e] : 0.0f;
    }
    float vt[4];
    if (bf16) {
      uint16_t h[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        h[e] = bf16_rne(v[e]);
        vt[e] = bf16_val(h[e]);
      }
      if (dst)
        *reinterpret_cast<uint2*>(static_cast<uint16_t*>(dst) + c0) =
            make_uint2(uint32_t(h[0]) | (uint32_t(h[1]) << 16), uint32_t(h[2]) | (uint32_t(h[3]) << 16));
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) vt[e] = v[e];
      if (dst) *reinterpret_cast<float4*>(static_cast<float*>(dst) + c0) = make_float4(v[0], v[1], v[2], v[3]);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double dv = (double)v[e], dt = (double)vt[e];
      s2 += dv * dv;
      e2 += (dt - dv) * (dt - dv);
      t2 += dt * dt;
      isint = isint && (v[e] == rintf(v[e])) && fabsf(v[e]) <= 256.0f;
      nonneg = nonneg && v[e] >= 0.0f;
    }
  }
  RowStats r;
  r.s2 = s2;
  r.e2 = e2;
  r.t2 = t2;
  r.isint = isint;
  r.nonneg = nonneg;
  return r;
}

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  return v;
}

// The thread-per-row form of prep_db_body: workgroup `blk` prepares rows
// 256 blk + threadIdx.x.
__device__ __forceinline__ void prep_db_thread_body(const PrepDbArgs& a, const int64_t blk) {
  __shared__ double red[4][5];
  __shared__ int redf[4][2];
  const int wave = threadIdx.x >> 6;
  const int lane = __lane_id();
  const size_t es = a.bf16 ? 2 : 4;
  const double slack = 1.0 + (a.d + 4) * 0x1p-52;
  double m[5] = {0, 0, 0, 0, 0};
  bool allint = true, allnonneg = true;
  const int64_t row = blk * 256 + threadIdx.x;
  if (row < a.n_pad) {
    char* dst = reinterpret_cast<char*>(a.Xop) + size_t(row) * a.Kp * es;
    if (row >= a.n) {
      for (int c = 0; c < a.Kp * (int)es; c += 16) *reinterpret_cast<uint4*>(dst + c) = make_uint4(0, 0, 0, 0);
      a.hx[row] = MFMA_PAD_HXN;
    } else {
      const RowStats r = prep_row_thread(a.X + row * a.ldx, a.d, a.Kp, a.bf16, dst);
      const double v = r.s2 * 0.5;
      const float h = (float)v;
      a.hx[row] = -h;
      m[0] = r.s2 * slack;
      m[1] = r.e2 * slack;
      m[2] = r.t2 * slack;
      m[3] = fabs((double)h - v) + (r.isint ? 0.0 : v * (a.d + 4) * 0x1p-53);
      m[4] = (double)h;
      allint = r.isint;
      allnonneg = r.nonneg;
    }
  }
#pragma unroll
  for (int t = 0; t < 5; ++t) m[t] = wave_max_d(m[t]);
  allint = __all(allint);
  allnonneg = __all(allnonneg);
  if (lane == 0) {
    for (int t = 0; t < 5; ++t) red[wave][t] = m[t];
    redf[wave][0] = allint ? 1 : 0;
    redf[wave][1] = allnonneg ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int t = threadIdx.x;
    const double v = fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t]));
    if (v > 0.0) atomic_max_nonneg(&a.consts[t], v);
  } else if (threadIdx.x < 7) {
    const int f = threadIdx.x - 5;
    if (!(redf[0][f] && redf[1][f] && redf[2][f] && redf[3][f])) atomicAnd(&a.flags[f], 0);
  }
}

// The thread-per-row form of prep_query_body: rows 256 blk + threadIdx.x.
__device__ __forceinline__ void prep_query_thread_body(const PrepQArgs& a, const int64_t blk) {
  const int64_t row = blk * 256 + threadIdx.x;
  if (row >= a.m_pad) return;
  const size_t es = a.bf16 ? 2 : 4;
  char* dst = a.Qop ? reinterpret_cast<char*>(a.Qop) + size_t(row) * a.Kp * es : nullptr;
  if (a.gthr)
    a.gthr[row] = row >= a.m ? ord_f32(-MFMA_KEY_LIMIT) : (a.theta ? ord_f32(a.theta[row]) : ord_f32(INFINITY));
  if (row >= a.m) {
    if (dst)
      for (int c = 0; c < a.Kp * (int)es; c += 16) *reinterpret_cast<uint4*>(dst + c) = make_uint4(0, 0, 0, 0);
    return;
  }
  const float* src = a.Q + (a.rows ? a.rows[row] : row) * a.ldq;
  float* q32 = a.Q32 + row * a.ld32;
  for (int c0 = 0; c0 < a.ld32; c0 += 4) {
    float4 f;
    if (c0 + 4 <= a.d) {
      f = *reinterpret_cast<const float4*>(src + c0);
    } else {
      f.x = c0 < a.d ? src[c0] : 0.0f;
      f.y = c0 + 1 < a.d ? src[c0 + 1] : 0.0f;
      f.z = c0 + 2 < a.d ? src[c0 + 2] : 0.0f;
      f.w = c0 + 3 < a.d ? src[c0 + 3] : 0.0f;
    }
    *reinterpret_cast<float4*>(q32 + c0) = f;
  }
  const RowStats r = prep_row_thread(src, a.d, a.Kp, a.bf16, dst);
  a.qn2[row] = r.s2;
  a.eq2[row] = r.e2;
  a.qt2[row] = r.t2;
  a.qint[row] = r.isint ? 1 : 0;
}

// One workgroup per 64 rows (16 per wave; `blk` is the workgroup's block of
// the problem); the index-wide maxima and flags are reduced in the workgroup
// and published with one atomic each (per-row atomics on the same few
// addresses serialise).
__device__ __forceinline__ void prep_db_body(const PrepDbArgs& a, const int64_t blk) {
  __shared__ double red[4][5];
  __shared__ int redf[4][2];
  const int wave = threadIdx.x >> 6;
  const int lane = __lane_id();
  const size_t es = a.bf16 ? 2 : 4;
  const double slack = 1.0 + (a.d + 4) * 0x1p-52;
  double m[5] = {0, 0, 0, 0, 0};
  bool allint = true, allnonneg = true;
  for (int i = 0; i < 16; ++i) {
    const int64_t row = blk * 64 + wave * 16 + i;
    if (row >= a.n_pad) break;
    char* dst = reinterpret_cast<char*>(a.Xop) + size_t(row) * a.Kp * es;
    if (row >= a.n) {
      for (int c = lane; c < a.Kp; c += 64) {
        if (a.bf16)
          reinterpret_cast<uint16_t*>(dst)[c] = 0;
        else
          reinterpret_cast<float*>(dst)[c] = 0.0f;
      }
      if (lane == 0) a.hx[row] = MFMA_PAD_HXN;
      continue;
    }
    const RowStats r = prep_row(a.X + row * a.ldx, a.d, a.Kp, a.bf16, dst);
    const double v = r.s2 * 0.5;
    const float h = (float)v;
    if (lane == 0) a.hx[row] = -h;
    m[0] = fmax(m[0], r.s2 * slack);
    m[1] = fmax(m[1], r.e2 * slack);
    m[2] = fmax(m[2], r.t2 * slack);
    // |fl32(||x||^2/2) - ||x||^2/2|, including the f64 error of the sum of
    // squares (none for integer rows: every partial sum is an integer < 2^53)
    m[3] = fmax(m[3], fabs((double)h - v) + (r.isint ? 0.0 : v * (a.d + 4) * 0x1p-53));
    m[4] = fmax(m[4], (double)h);
    allint = allint && r.isint;
    allnonneg = allnonneg && r.nonneg;
  }
  if (lane == 0) {
    for (int t = 0; t < 5; ++t) red[wave][t] = m[t];
    redf[wave][0] = allint ? 1 : 0;
    redf[wave][1] = allnonneg ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int t = threadIdx.x;
    const double v = fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t]));
    if (v > 0.0) atomic_max_nonneg(&a.consts[t], v);
  } else if (threadIdx.x < 7) {
    const int f = threadIdx.x - 5;
    if (!(redf[0][f] && redf[1][f] && redf[2][f] && redf[3][f])) atomicAnd(&a.flags[f], 0);
  }
}

__device__ __forceinline__ void prep_query_body(const PrepQArgs& a, const int64_t blk) {
  const int64_t row = blk * 4 + (threadIdx.x >> 6);
  if (row >= a.m_pad) return;
  const int lane = __lane_id();
  const size_t es = a.bf16 ? 2 : 4;
  // no operand copy for the residual operand (the scan builds it per list)
  char* dst = a.Qop ? reinterpret_cast<char*>(a.Qop) + size_t(row) * a.Kp * es : nullptr;
  // +inf, or the caller's threshold for this slot; a padding slot starts below
  // every key, so that it keeps no candidates (its results are never read)
  if (lane == 0 && a.gthr)
    a.gthr[row] = row >= a.m ? ord_f32(-MFMA_KEY_LIMIT) : (a.theta ? ord_f32(a.theta[row]) : ord_f32(INFINITY));
  if (row >= a.m) {
    if (!dst) return;
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
  if (lane == 0) {  // raw f64 sums; the bound applies directed slack
    a.qn2[row] = r.s2;
    a.eq2[row] = r.e2;
    a.qt2[row] = r.t2;
    a.qint[row] = r.isint ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void prep_db_kernel(PrepDbArgs a) { prep_db_body(a, blockIdx.x); }
__global__ __launch_bounds__(256) void prep_query_kernel(PrepQArgs a) { prep_query_body(a, blockIdx.x); }

// Grouped forms: one grid over the problems of a table (group_problem).
__global__ __launch_bounds__(256) void prep_db_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const PrepDbArgs a = static_cast<const PrepDbArgs*>(g.tab)[p];
  prep_db_body(a, int64_t(blockIdx.x) - g.start[p]);
}
__global__ __launch_bounds__(256) void prep_query_group_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const PrepQArgs a = static_cast<const PrepQArgs*>(g.tab)[p];
  prep_query_body(a, int64_t(blockIdx.x) - g.start[p]);
}

__global__ __launch_bounds__(256) void prep_db_group_thread_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const PrepDbArgs a = static_cast<const PrepDbArgs*>(g.tab)[p];
  prep_db_thread_body(a, int64_t(blockIdx.x) - g.start[p]);
}
__global__ __launch_bounds__(256) void prep_query_group_thread_kernel(GroupTab g) {
  const int p = group_problem(g, blockIdx.x);
  const PrepQArgs a = static_cast<const PrepQArgs*>(g.tab)[p];
  prep_query_thread_body(a, int64_t(blockIdx.x) - g.start[p]);
}

// One workgroup per table entry.
__global__ __launch_bounds__(256) void fill_multi_kernel(const FillEntry* tab) {
  const FillEntry f = tab[blockIdx.x];
  for (int64_t i = threadIdx.x; i < f.count; i += 256) f.p[i] = f.v;
}

hipError_t launch_fill_multi(const FillEntry* tab, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_multi_kernel, dim3((unsigned)n), dim3(256), 0, st, tab);
  return hipGetLastError();
}

hipError_t launch_prep_db_group(const GroupTab& g, bool thread_rows, hipStream_t st) {
  if (g.np <= 0 || g.nblocks <= 0) return hipSuccess;
  if (thread_rows)
    hipLaunchKernelGGL(prep_db_group_thread_kernel, dim3((unsigned)g.nblocks), dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL(prep_db_group_kernel, dim3((unsigned)g.nblocks), dim3(256), 0, st, g);
  return hipGetLastError();
}

hipError_t launch_prep_queries_group(const GroupTab& g, bool thread_rows, hipStream_t st) {
  if (g.np <= 0 || g.nblocks <= 0) return hipSuccess;
  if (thread_rows)
    hipLaunchKernelGGL(prep_query_group_thread_kernel, dim3((unsigned)g.nblocks), dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL(prep_query_group_kernel, dim3((unsigned)g.nblocks), dim3(256), 0, st, g);
  return hipGetLastError();
}

hipError_t launch_prep_db(const PrepDbArgs& a, hipStream_t st) {
  if (a.n_pad == 0) return hipSuccess;
  hipLaunchKernelGGL(prep_db_kernel, dim3((unsigned)((a.n_pad + 63) / 64)), dim3(256), 0, st, a);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, int64_t count, uint32_t v) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < count) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, int64_t count, uint32_t v, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count, v);
  return hipGetLastError();
}

hipError_t launch_prep_queries(const PrepQArgs& a, hipStream_t st) {
  if (a.m_pad == 0) return hipSuccess;
  hipLaunchKernelGGL(prep_query_kernel, dim3((unsigned)((a.m_pad + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace fa
