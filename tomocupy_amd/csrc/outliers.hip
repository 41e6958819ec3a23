// Zinger removal (tomocupy ProcFunctions.remove_outliers,
// src/tomocupy/processing/proc_functions.py:96-107) for gfx950:
//   f = median_filter(data, [w, 1, w])      (angle, slice, detector window)
//   data = where(data > f and data - f > threshold, f, data)
// on uint16 counts or float32 (binned) counts, bit-exact.  The median follows scipy/cupyx
// ndimage.median_filter: mode 'reflect' (half-sample symmetric), window
// offsets -(w/2) .. w-1-w/2 along both axes, rank (w*w)/2 of the sorted
// window (the upper middle for even w).
#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "common.h"
#include "preproc.h"

using namespace tomo;

namespace {

constexpr int kOutRows = 16;  // output rows (angles) per workgroup
constexpr int kOutCols = 64;  // output columns (detector) per workgroup

__device__ __forceinline__ int refl_idx(int m, int N) {
  // half-sample symmetric reflection into [0, N) (scipy 'reflect')
  if ((unsigned)m < (unsigned)N) return m;
  const int P = 2 * N;
  int mm = m % P;
  if (mm < 0) mm += P;
  return mm >= N ? P - 1 - mm : mm;
}

// Order-preserving unsigned keys of the counts: uint16 as is; float32 by the
// usual sign flip (negative values: all bits inverted, others: sign bit set).
template <typename T>
struct Key;
template <>
struct Key<uint16_t> {
  static constexpr int bits = 16;
  __device__ static uint32_t of(uint16_t v) { return v; }
  __device__ static uint16_t back(uint32_t k) { return (uint16_t)k; }
};
template <>
struct Key<float> {
  static constexpr int bits = 32;
  __device__ static uint32_t of(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
  __device__ static float back(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
  }
};

// Median of the w x w window of every pixel of one slice plane, written to
// `med` (same layout as `data`).  The workgroup stages its (16+w-1) x (64+w-1)
// input tile (reflected at the borders) in LDS; each thread then selects the
// rank-(w*w/2) value of its window bit by bit: the largest key t with
// #{key < t} <= rank, over the 16 (uint16) or 32 (float) bits of the keys.
template <int W, typename T>
__global__ __launch_bounds__(256) void k_pre_median(const T *__restrict__ data,
                                                    T *__restrict__ med, int nrows, int nz,
                                                    int n) {
  constexpr int TR = kOutRows + W - 1, TC = kOutCols + W - 1;
  __shared__ T tile[TR][TC];
  const int z = blockIdx.z;
  const int p0 = blockIdx.y * kOutRows, x0 = blockIdx.x * kOutCols;
  const int r0 = p0 - W / 2, c0 = x0 - W / 2;
  for (int idx = threadIdx.x; idx < TR * TC; idx += 256) {
    const int r = idx / TC, c = idx - r * TC;
    const int p = refl_idx(r0 + r, nrows), x = refl_idx(c0 + c, n);
    tile[r][c] = data[((size_t)p * nz + z) * n + x];
  }
  __syncthreads();
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x = x0 + tx;
  if (x >= n) return;
  constexpr int K = (W * W) / 2;
#pragma unroll 1
  for (int j = 0; j < kOutRows / 4; ++j) {
    const int r = ty + 4 * j, p = p0 + r;
    if (p >= nrows) break;
    uint32_t v[W * W];
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
      for (int b = 0; b < W; ++b) v[a * W + b] = Key<T>::of(tile[r + a][tx + b]);
    uint32_t m = 0;
#pragma unroll
    for (int bit = Key<T>::bits - 1; bit >= 0; --bit) {
      const uint32_t t = m | (1u << bit);
      int c = 0;
#pragma unroll
      for (int i = 0; i < W * W; ++i) c += v[i] < t;
      if (c <= K) m = t;
    }
    med[((size_t)p * nz + z) * n + x] = Key<T>::back(m);
  }
}

// data = (data > med and data - med > threshold) ? med : data, elementwise
// (integer arithmetic for uint16 counts, float32 for float counts).
__device__ __forceinline__ bool is_zinger(uint16_t v, uint16_t f, int threshold) {
  return v > f && (int)v - (int)f > threshold;
}
__device__ __forceinline__ bool is_zinger(float v, float f, int threshold) {
  return v > f && v - f > (float)threshold;
}

template <typename T>
__global__ __launch_bounds__(256) void k_pre_dezinger(T *__restrict__ data,
                                                      const T *__restrict__ med, size_t count,
                                                      int threshold) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const T v = data[i], f = med[i];
  if (is_zinger(v, f, threshold)) data[i] = f;
}

template <typename F>
void with_width(int w, F &&f) {
  switch (w) {
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 5: f(std::integral_constant<int, 5>()); break;
    case 6: f(std::integral_constant<int, 6>()); break;
    case 7: f(std::integral_constant<int, 7>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    default: f(std::integral_constant<int, 9>()); break;
  }
}

template <typename T>
static int remove_outliers(const char *who, tomo_pre_plan *P, T *data, int64_t nrows, int width,
                           int threshold, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "%s: plan is NULL", who);
  TOMO_REQUIRE(data != nullptr, "%s: data is NULL", who);
  TOMO_REQUIRE(nrows >= 1 && nrows <= 65535, "%s: nrows=%lld out of range", who,
               (long long)nrows);
  TOMO_REQUIRE(width >= 0 && width <= kMaxDezinger, "%s: width=%d not supported (0..%d)", who,
               width, kMaxDezinger);
  if (width <= 1) return TOMO_OK;  // a 1 x 1 median is the pixel itself
  const size_t count = (size_t)nrows * P->nz * P->n;
  if (P->outl.bytes < sizeof(T) * count) {
    // scratch for the medians, sized on first use for the largest array seen
    // (projections, darks or flats of the chunk)
    const int64_t rows = std::max<int64_t>(nrows, std::max(P->nproj, std::max(P->ndark, P->nflat)));
    P->outl.release();
    int rc = P->outl.alloc(sizeof(T) * (size_t)rows * P->nz * P->n);
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  T *med = P->outl.as<T>();
  const dim3 grid((unsigned)ceil_div(P->n, kOutCols), (unsigned)ceil_div(nrows, kOutRows),
                  (unsigned)P->nz);
  {
    ProfScope ps_("k_pre_median", st);
    with_width(width, [&](auto wc) {
      hipLaunchKernelGGL((k_pre_median<decltype(wc)::value, T>), grid, dim3(256), 0, st, data,
                         med, (int)nrows, (int)P->nz, (int)P->n);
    });
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_pre_dezinger", st);
    hipLaunchKernelGGL(k_pre_dezinger<T>, dim3((unsigned)ceil_div((int64_t)count, 256)),
                       dim3(256), 0, st, data, med, count, threshold);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

}  // namespace

extern "C" int pre_remove_outliers(tomo_pre_plan *P, uint16_t *data, int64_t nrows, int width,
                                   int threshold, void *stream) {
  return remove_outliers("pre_remove_outliers", P, data, nrows, width, threshold, stream);
}

extern "C" int pre_remove_outliers_f32(tomo_pre_plan *P, float *data, int64_t nrows, int width,
                                       int threshold, void *stream) {
  return remove_outliers("pre_remove_outliers_f32", P, data, nrows, width, threshold, stream);
}
