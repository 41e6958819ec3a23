// Zinger removal (tomocupy ProcFunctions.remove_outliers,
// src/tomocupy/processing/proc_functions.py:96-107) for gfx950:
//   f = median_filter(data, [w, 1, w])      (angle, slice, detector window)
//   data = where(data > f and data - f > threshold, f, data)
// on uint16 counts, bit-exact.  The median follows scipy/cupyx
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

// Median of the w x w window of every pixel of one slice plane, written to
// `med` (same layout as `data`).  The workgroup stages its (16+w-1) x (64+w-1)
// input tile (reflected at the borders) in LDS; each thread then selects the
// rank-(w*w/2) value of its window bit by bit: the largest t with
// #{v < t} <= rank, over the 16 bits of the counts.
template <int W>
__global__ __launch_bounds__(256) void k_pre_median(const uint16_t *__restrict__ data,
                                                    uint16_t *__restrict__ med, int nrows,
                                                    int nz, int n) {
  constexpr int TR = kOutRows + W - 1, TC = kOutCols + W - 1;
  __shared__ uint16_t tile[TR][TC];
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
    int v[W * W];
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
      for (int b = 0; b < W; ++b) v[a * W + b] = tile[r + a][tx + b];
    int m = 0;
#pragma unroll
    for (int bit = 15; bit >= 0; --bit) {
      const int t = m | (1 << bit);
      int c = 0;
#pragma unroll
      for (int i = 0; i < W * W; ++i) c += v[i] < t;
      if (c <= K) m = t;
    }
    med[((size_t)p * nz + z) * n + x] = (uint16_t)m;
  }
}

// data = (data > med and data - med > threshold) ? med : data, elementwise.
__global__ __launch_bounds__(256) void k_pre_dezinger(uint16_t *__restrict__ data,
                                                      const uint16_t *__restrict__ med,
                                                      size_t count, int threshold) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const int v = data[i], f = med[i];
  if (v > f && v - f > threshold) data[i] = (uint16_t)f;
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

}  // namespace

extern "C" int pre_remove_outliers(tomo_pre_plan *P, uint16_t *data, int64_t nrows, int width,
                                   int threshold, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "pre_remove_outliers: plan is NULL");
  TOMO_REQUIRE(data != nullptr, "pre_remove_outliers: data is NULL");
  TOMO_REQUIRE(nrows >= 1 && nrows <= 65535, "pre_remove_outliers: nrows=%lld out of range",
               (long long)nrows);
  TOMO_REQUIRE(width >= 0 && width <= kMaxDezinger,
               "pre_remove_outliers: width=%d not supported (0..%d)", width, kMaxDezinger);
  if (width <= 1) return TOMO_OK;  // a 1 x 1 median is the pixel itself
  const size_t count = (size_t)nrows * P->nz * P->n;
  if (P->outl.bytes < sizeof(uint16_t) * count) {
    // scratch for the medians, sized on first use for the largest array seen
    // (projections, darks or flats of the chunk)
    const int64_t rows = std::max<int64_t>(nrows, std::max(P->nproj, std::max(P->ndark, P->nflat)));
    P->outl.release();
    int rc = P->outl.alloc(sizeof(uint16_t) * (size_t)rows * P->nz * P->n);
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  uint16_t *med = P->outl.as<uint16_t>();
  const dim3 grid((unsigned)ceil_div(P->n, kOutCols), (unsigned)ceil_div(nrows, kOutRows),
                  (unsigned)P->nz);
  {
    ProfScope ps_("k_pre_median", st);
    with_width(width, [&](auto wc) {
      hipLaunchKernelGGL((k_pre_median<decltype(wc)::value>), grid, dim3(256), 0, st, data, med,
                         (int)nrows, (int)P->nz, (int)P->n);
    });
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_pre_dezinger", st);
    hipLaunchKernelGGL(k_pre_dezinger, dim3((unsigned)ceil_div((int64_t)count, 256)), dim3(256), 0,
                       st, data, med, count, threshold);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}
