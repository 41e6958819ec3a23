// Binning of raw frames for gfx950 (tomocupy utils.downsample, src/tomocupy/
// utils.py:157-167, applied by the reader to projections, darks and flats,
// dataio/reader.py:350-389): every output pixel is the sum of a 2^b x 2^b
// block of counts (b detector columns halvings, then b row halvings).  The
// sums are formed in uint32 and stored as float32; for b <= 3 they stay below
// 2^24 (64 * 65535), so they are exact and equal the reference's float sums
// in any order of addition.
#include <cstdint>
#include <type_traits>

#include "common.h"
#include "preproc.h"

using namespace tomo;

namespace {

// One thread per output pixel x of output row (r, z); blockIdx.y = z,
// blockIdx.z = r.  Input (rows, nz << B, n << B), output (rows, nz, n).
template <int B>
__global__ __launch_bounds__(256) void k_pre_bin(const uint16_t *__restrict__ src,
                                                 float *__restrict__ dst, int nz, int n) {
  constexpr int S = 1 << B;
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= n) return;
  const int z = blockIdx.y, r = blockIdx.z;
  const size_t wi = (size_t)n << B;
  const uint16_t *s = src + ((size_t)r * ((size_t)nz << B) + ((size_t)z << B)) * wi + ((size_t)x << B);
  uint32_t acc = 0;
#pragma unroll
  for (int a = 0; a < S; ++a)
#pragma unroll
    for (int b = 0; b < S; ++b) acc += s[(size_t)a * wi + b];
  dst[((size_t)r * nz + z) * n + x] = (float)acc;
}

}  // namespace

extern "C" int pre_bin(tomo_pre_plan *P, const uint16_t *src, float *dst, int64_t nrows,
                       int binning, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "pre_bin: plan is NULL");
  TOMO_REQUIRE(src && dst, "pre_bin: NULL array");
  TOMO_REQUIRE(binning >= 0 && binning <= 3, "pre_bin: binning=%d not in 0..3", binning);
  TOMO_REQUIRE(nrows >= 1 && nrows <= 65535, "pre_bin: nrows=%lld out of range",
               (long long)nrows);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)ceil_div(P->n, 256), (unsigned)P->nz, (unsigned)nrows);
  const int nz = (int)P->nz, n = (int)P->n;
  {
    ProfScope ps_("k_pre_bin", st);
    switch (binning) {
      case 0: hipLaunchKernelGGL(k_pre_bin<0>, grid, dim3(256), 0, st, src, dst, nz, n); break;
      case 1: hipLaunchKernelGGL(k_pre_bin<1>, grid, dim3(256), 0, st, src, dst, nz, n); break;
      case 2: hipLaunchKernelGGL(k_pre_bin<2>, grid, dim3(256), 0, st, src, dst, nz, n); break;
      default: hipLaunchKernelGGL(k_pre_bin<3>, grid, dim3(256), 0, st, src, dst, nz, n); break;
    }
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}
