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

// Partial column sums over angles, FP64: part[s][z][x] = sum of res[p][z][x]
// over the s-th of gridDim.z blocks of angles (one thread per column and
// block, coalesced along the detector; the blocks keep enough loads in
// flight to stream the chunk).
__global__ __launch_bounds__(256) void k_ti_mean(const float *__restrict__ res,
                                                 double *__restrict__ part, int nproj, int nz,
                                                 int n) {
  const int x = blockIdx.x * 256 + threadIdx.x, z = blockIdx.y, sb = blockIdx.z;
  if (x >= n) return;
  const int len = (nproj + gridDim.z - 1) / gridDim.z;
  const int p0 = sb * len, p1 = min(nproj, p0 + len);
  double s = 0.0;
  for (int p = p0; p < p1; ++p) s += (double)res[((size_t)p * nz + z) * n + x];
  part[((size_t)sb * nz + z) * n + x] = s;
}

// v[z][x] = column mean (the nsb partial sums over nproj); gamma (one row,
// written by the z = 0 blocks): gamma_k = beta r^k' - [k = 0]
__global__ __launch_bounds__(256) void k_ti_vfin(const double *__restrict__ part, int nsb,
                                                 int nproj, int nz, int n, double beta,
                                                 double *__restrict__ v,
                                                 double *__restrict__ gam) {
  const int x = blockIdx.x * 256 + threadIdx.x, z = blockIdx.y;
  if (x >= n) return;
  double s = 0.0;
  for (int b = 0; b < nsb; ++b) s += part[((size_t)b * nz + z) * n + x];
  v[(size_t)z * n + x] = s / nproj;
  if (z == 0) {
    const double r = (1.0 - beta) / (1.0 + beta);
    const int k = x <= n / 2 ? x : n - x;
    gam[x] = beta * std::pow(r, (double)k) - (x == 0 ? 1.0 : 0.0);
  }
}

// c[z][x] = mask(x) * sum_j (v[z][j] - v[z][0]) gamma[(x - j) mod n], FP64.
// The slice's v row and gamma are staged in LDS (2n doubles); 64 outputs per
// workgroup, four threads (adjacent lanes) per output over quarters of j,
// combined by lane shuffles.
__global__ __launch_bounds__(256) void k_ti_conv(const double *__restrict__ v,
                                                 const double *__restrict__ gam_,
                                                 double *__restrict__ c, int n, int lo, int hi) {
  extern __shared__ double sm[];
  double *vr = sm, *gam = sm + n;
  const int z = blockIdx.y;
  const double v0 = v[(size_t)z * n];
  for (int j = threadIdx.x; j < n; j += 256) {
    vr[j] = v[(size_t)z * n + j] - v0;
    gam[j] = gam_[j];
  }
  __syncthreads();
  const int x = blockIdx.x * 64 + (threadIdx.x >> 2), q = threadIdx.x & 3;
  const int j0 = (int)(((long long)n * q) / 4), j1 = (int)(((long long)n * (q + 1)) / 4);
  double s = 0.0;
  if (x < n && x >= lo && x < hi) {
    for (int j = j0; j < j1; ++j) {
      const int d = x - j;
      s += vr[j] * gam[d >= 0 ? d : d + n];
    }
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if (x < n && q == 0) c[(size_t)z * n + x] = s;
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
  const int nsb = nproj < 32 ? nproj : 32;  // blocks of angles of the column sums
  const size_t need = sizeof(double) * ((size_t)(nsb + 2) * nz * n + n);
  if (P->ti.bytes < need) {
    P->ti.release();
    int rc = P->ti.alloc(need);
    if (rc) return rc;
  }
  // the central mask*n columns: [n/2 - m/2, n/2 + m/2), m = floor(mask*n)
  const int m = (int)std::floor(mask * n);
  const int lo = n / 2 - m / 2, hi = n / 2 + m / 2;
  hipStream_t st = (hipStream_t)stream;
  double *part = P->ti.as<double>(), *v = part + (size_t)nsb * nz * n;
  double *c = v + (size_t)nz * n, *gam = c + (size_t)nz * n;
  const dim3 g2((unsigned)ceil_div(n, 256), (unsigned)nz);
  {
    ProfScope ps_("k_ti_mean", st);
    hipLaunchKernelGGL(k_ti_mean, dim3(g2.x, g2.y, (unsigned)nsb), dim3(256), 0, st, res, part,
                       nproj, nz, n);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_ti_conv", st);
    const size_t lds = 2 * sizeof(double) * n;
    TOMO_HIP(hipFuncSetAttribute((const void *)k_ti_conv,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_ti_vfin, g2, dim3(256), 0, st, part, nsb, nproj, nz, n, beta, v, gam);
    hipLaunchKernelGGL(k_ti_conv, dim3((unsigned)ceil_div(n, 64), (unsigned)nz), dim3(256), lds,
                       st, v, gam, c, n, lo, hi);
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
