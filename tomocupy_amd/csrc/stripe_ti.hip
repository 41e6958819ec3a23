// Titarenko ring removal (tomocupy remove_stripe_ti,
// src/tomocupy/processing/remove_stripe.py:296-309) for gfx950, on the float32
// transmission chunk res (nproj, nz, n) in place:
//   v = mean over angles (per slice and column) - its first column
//   c = v (*) gamma, circular convolution of length n with
//       gamma_k = beta ((1-beta)/(1+beta))^|k'| (k' = k for k <= n/2, n-k
//       otherwise; fftfreq(n)*n), gamma_0 - 1
//   res += c on the central mask*n columns
// The reference forms c with rfft/irfft (:303); the direct sum here is the
// same circular convolution, in FP64.  Like the reference, even n only.
#include <cmath>
#include <cstdint>

#include "common.h"
#include "preproc.h"

using namespace tomo;

namespace {

// v[z][x] = mean_p res[p][z][x] in FP64 (one thread per column of a slice)
__global__ __launch_bounds__(256) void k_ti_mean(const float *__restrict__ res,
                                                 double *__restrict__ v, int nproj, int nz,
                                                 int n) {
  const int x = blockIdx.x * 256 + threadIdx.x, z = blockIdx.y;
  if (x >= n) return;
  double s = 0.0;
  for (int p = 0; p < nproj; ++p) s += (double)res[((size_t)p * nz + z) * n + x];
  v[(size_t)z * n + x] = s / nproj;
}

// c[z][x] = mask(x) * sum_j (v[z][j] - v[z][0]) gamma[(x - j) mod n], FP64;
// the slice's v row and the kernel gamma are staged in LDS (2n doubles).
__global__ __launch_bounds__(256) void k_ti_conv(const double *__restrict__ v,
                                                 double *__restrict__ c, int n, double beta,
                                                 int lo, int hi) {
  extern __shared__ double sm[];
  double *vr = sm, *gam = sm + n;
  const int z = blockIdx.y;
  const double r = (1.0 - beta) / (1.0 + beta);
  const double v0 = v[(size_t)z * n];
  for (int j = threadIdx.x; j < n; j += 256) {
    vr[j] = v[(size_t)z * n + j] - v0;
    const int k = j <= n / 2 ? j : n - j;
    gam[j] = beta * std::pow(r, (double)k) - (j == 0 ? 1.0 : 0.0);
  }
  __syncthreads();
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= n) return;
  double s = 0.0;
  if (x >= lo && x < hi) {
    // j <= x: gamma[x - j]; j > x: gamma[x - j + n]
    for (int j = 0; j <= x; ++j) s += vr[j] * gam[x - j];
    for (int j = x + 1; j < n; ++j) s += vr[j] * gam[x - j + n];
  }
  c[(size_t)z * n + x] = s;
}

// res[p][z][x] += c[z][x] (in FP64, rounded once)
__global__ __launch_bounds__(256) void k_ti_add(float *__restrict__ res,
                                                const double *__restrict__ c, int nz, int n,
                                                size_t count) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const size_t zx = i % ((size_t)nz * n);
  res[i] = (float)((double)res[i] + c[zx]);
}

}  // namespace

extern "C" int pre_remove_stripe_ti(tomo_pre_plan *P, float *res, double beta, double mask,
                                    void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr && res != nullptr, "pre_remove_stripe_ti: NULL argument");
  TOMO_REQUIRE(std::isfinite(beta) && beta > 0.0 && beta < 1.0,
               "pre_remove_stripe_ti: beta=%g must be in (0, 1)", beta);
  TOMO_REQUIRE(std::isfinite(mask) && mask >= 0.0 && mask <= 1.0,
               "pre_remove_stripe_ti: mask=%g must be in [0, 1]", mask);
  const int n = (int)P->n, nz = (int)P->nz, nproj = (int)P->nproj;
  TOMO_REQUIRE(n % 2 == 0 && n >= 2, "pre_remove_stripe_ti: n=%d must be even", n);
  TOMO_REQUIRE((size_t)2 * n * sizeof(double) <= max_lds_per_block(),
               "pre_remove_stripe_ti: n=%d too large for one workgroup", n);
  if (P->ti.bytes < 2 * sizeof(double) * nz * n) {
    P->ti.release();
    int rc = P->ti.alloc(2 * sizeof(double) * nz * n);
    if (rc) return rc;
  }
  // the central mask*n columns: [n/2 - m/2, n/2 + m/2), m = floor(mask*n)
  const int m = (int)std::floor(mask * n);
  const int lo = n / 2 - m / 2, hi = n / 2 + m / 2;
  hipStream_t st = (hipStream_t)stream;
  double *v = P->ti.as<double>(), *c = v + (size_t)nz * n;
  const dim3 g2((unsigned)ceil_div(n, 256), (unsigned)nz);
  {
    ProfScope ps_("k_ti_mean", st);
    hipLaunchKernelGGL(k_ti_mean, g2, dim3(256), 0, st, res, v, nproj, nz, n);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_ti_conv", st);
    const size_t lds = 2 * sizeof(double) * n;
    TOMO_HIP(hipFuncSetAttribute((const void *)k_ti_conv,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_ti_conv, g2, dim3(256), lds, st, v, c, n, beta, lo, hi);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_ti_add", st);
    const size_t count = (size_t)nproj * nz * n;
    hipLaunchKernelGGL(k_ti_add, dim3((unsigned)ceil_div((int64_t)count, 256)), dim3(256), 0, st,
                       res, c, nz, n, count);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}
