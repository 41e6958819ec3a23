// Chunk preprocessing for gfx950: dark/flat correction, minus-log and the swap
// from projection order (nproj, nz, n) to sinogram order (nz, nproj, n)
// (tomocupy processing/proc_functions.py:55-83, 140-198; rec.py:172-174).
#include <cmath>
#include <mutex>
#include <string>

#include "common.h"
#include "precision.h"
#include "preproc.h"

using namespace tomo;

namespace {

// Sum of frames [f0, f1) of pixel i (frame stride zn) as conv(value), four
// independent FP64 partial sums so a thread's loads are in flight together.
// The terms are integers (dark) or float32 quotients (flat) well inside
// FP64's 53 bits, so every partial sum and their total are exact and the
// result does not depend on the order.
template <typename TI, typename F>
__device__ __forceinline__ double frame_sum(const TI *__restrict__ p, int f0, int f1, int zn,
                                            int i, F conv) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int f = f0;
  for (; f + 4 <= f1; f += 4) {
    const double a = conv(p[(size_t)f * zn + i]), b = conv(p[(size_t)(f + 1) * zn + i]);
    const double c = conv(p[(size_t)(f + 2) * zn + i]), d = conv(p[(size_t)(f + 3) * zn + i]);
    s0 += a;
    s1 += b;
    s2 += c;
    s3 += d;
  }
  for (; f < f1; ++f) s0 += conv(p[(size_t)f * zn + i]);
  return (s0 + s1) + (s2 + s3);
}

// dark0 = mean(dark); flat means (one, or two halves for flat_linear)
// (proc_functions.py:55-72).  Means accumulate in FP64.  TI is the type of
// the counts: uint16_t (raw frames) or float (binned frames).  One thread
// per pixel, 64-thread workgroups so that the launch spreads over all CUs.
constexpr int kMeansThreads = 64;

template <typename TI>
__global__ __launch_bounds__(kMeansThreads) void k_pre_means(const TI *__restrict__ dark,
                                                  const TI *__restrict__ flat,
                                                  float *__restrict__ dark0,
                                                  float *__restrict__ flatA,
                                                  float *__restrict__ flatB, int zn, int ndark,
                                                  int nflat, float bright, int flat_linear) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= zn) return;
  auto raw = [](TI v) { return (double)v; };
  auto rel = [bright](TI v) { return (double)((float)v / bright); };
  dark0[i] = (float)(frame_sum(dark, 0, ndark, zn, i, raw) / ndark);
  if (!flat_linear) {
    flatA[i] = (float)(frame_sum(flat, 0, nflat, zn, i, rel) / nflat);
  } else {
    const int h = nflat / 2;
    flatA[i] = (float)(frame_sum(flat, 0, h, zn, i, rel) / h);
    flatB[i] = (float)(frame_sum(flat, h + 1, nflat, zn, i, rel) / (nflat - h - 1));
  }
}

// One workgroup per (p, z) detector row: read data[p][z][:], write out[z][p][:].
// TO is the storage type of the output, TI that of the counts.
template <int MODE, typename TO, typename TI>  // 0: darkflat (+ optional -log) + swap; 1: darkflat only, projection order
__global__ void k_pre_rows(const TI *__restrict__ data, const float *__restrict__ dark0,
                           const float *__restrict__ flatA, const float *__restrict__ flatB,
                           TO *__restrict__ out, int nproj, int nz, int n, int flat_linear,
                           int take_log) {
  const int p = blockIdx.x, z = blockIdx.y;
  const TI *src = data + ((size_t)p * nz + z) * n;
  TO *dst = MODE == 0 ? out + ((size_t)z * nproj + p) * n : out + ((size_t)p * nz + z) * n;
  const float v = nproj > 1 ? (float)((double)p / (double)(nproj - 1)) : 0.f;
  const float *d0 = dark0 + (size_t)z * n;
  const float *fa = flatA + (size_t)z * n;
  const float *fb = flatB + (size_t)z * n;
  for (int x = threadIdx.x; x < n; x += blockDim.x) {
    float r = darkflat((float)src[x], d0[x], fa[x], flat_linear ? fb[x] : 0.f, v, flat_linear);
    if (MODE == 0 && take_log) r = minus_log(r);
    str(dst + x, r);
  }
}

// out[z][p][:] = (-log) res[p][z0 + z][:] for z < lz, 0 for lz <= z < gridDim.y;
// res has ld rows per projection (the chunk itself: ld = nz, z0 = 0).
template <typename TO>
__global__ void k_pre_log_swap(const float *__restrict__ res, TO *__restrict__ out, int nproj,
                               int ld, int z0, int lz, int n, int take_log) {
  const int p = blockIdx.x, z = blockIdx.y;
  TO *dst = out + ((size_t)z * nproj + p) * n;
  if (z >= lz) {
    for (int x = threadIdx.x; x < n; x += blockDim.x) str(dst + x, 0.f);
    return;
  }
  const float *src = res + ((size_t)p * ld + z0 + z) * n;
  for (int x = threadIdx.x; x < n; x += blockDim.x) {
    const float r = src[x];
    str(dst + x, take_log ? minus_log(r) : r);
  }
}

}  // namespace


extern "C" int pre_create(tomo_pre_plan **plan, int64_t nproj, int64_t nz, int64_t n,
                          int64_t ndark, int64_t nflat) {
  clear_error();
  TOMO_REQUIRE(plan != nullptr, "pre_create: plan pointer is NULL");
  *plan = nullptr;
  TOMO_REQUIRE(nproj >= 1 && nz >= 1 && n >= 1, "pre_create: sizes must be >= 1");
  TOMO_REQUIRE(ndark >= 1 && nflat >= 1, "pre_create: need at least one dark and one flat frame");
  TOMO_REQUIRE(nproj <= 65535 && nz <= 65535, "pre_create: nproj and nz must be <= 65535");
  auto *P = new (std::nothrow) tomo_pre_plan();
  if (!P) {
    set_error("pre_create: host allocation failed");
    return TOMO_ENOMEM;
  }
  P->nproj = nproj;
  P->nz = nz;
  P->n = n;
  P->ndark = ndark;
  P->nflat = nflat;
  int rc = P->dark0.alloc(sizeof(float) * nz * n);
  if (!rc) rc = P->flatA.alloc(sizeof(float) * nz * n);
  if (!rc) rc = P->flatB.alloc(sizeof(float) * nz * n);
  if (rc) {
    pre_destroy(P);
    return rc;
  }
  *plan = P;
  return TOMO_OK;
}

template <typename TI>
static int launch_means(tomo_pre_plan *P, const TI *dark, const TI *flat, double bright_ratio,
                        int flat_linear, hipStream_t st) {
  TOMO_REQUIRE(bright_ratio > 0 && std::isfinite(bright_ratio), "bright_ratio must be > 0");
  TOMO_REQUIRE(!flat_linear || P->nflat >= 3, "flat_linear needs at least 3 flat frames");
  const int zn = (int)(P->nz * P->n);
  {
    ProfScope ps_("k_pre_means", st);
    hipLaunchKernelGGL(k_pre_means<TI>, dim3((unsigned)ceil_div(zn, kMeansThreads)),
                       dim3(kMeansThreads), 0, st, dark, flat, P->dark0.as<float>(),
                       P->flatA.as<float>(), P->flatB.as<float>(), zn, (int)P->ndark,
                       (int)P->nflat, (float)bright_ratio, flat_linear);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

int tomo::pre_launch_means(tomo_pre_plan *P, const uint16_t *dark, const uint16_t *flat,
                           double bright_ratio, int flat_linear, hipStream_t st) {
  return launch_means(P, dark, flat, bright_ratio, flat_linear, st);
}

template <typename TO, typename TI>
static void launch_rows0(tomo_pre_plan *P, const TI *data, void *out, int flat_linear,
                         int take_log, hipStream_t st) {
  hipLaunchKernelGGL((k_pre_rows<0, TO, TI>), dim3((unsigned)P->nproj, (unsigned)P->nz), dim3(256),
                     0, st, data, P->dark0.as<float>(), P->flatA.as<float>(), P->flatB.as<float>(),
                     (TO *)out, (int)P->nproj, (int)P->nz, (int)P->n, flat_linear, take_log);
}

template <typename TO>
static void launch_log_swap(tomo_pre_plan *P, const float *res, int64_t ld, int64_t z0,
                            int64_t lz, void *out, int take_log, hipStream_t st) {
  hipLaunchKernelGGL(k_pre_log_swap<TO>, dim3((unsigned)P->nproj, (unsigned)P->nz), dim3(256), 0,
                     st, res, (TO *)out, (int)P->nproj, (int)ld, (int)z0, (int)lz, (int)P->n,
                     take_log);
}

template <typename TI>
static int process(const char *who, tomo_pre_plan *P, const TI *data, const TI *dark,
                   const TI *flat, void *out, int out_dtype, double bright_ratio, int flat_linear,
                   int minus_log, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "%s: plan is NULL", who);
  TOMO_REQUIRE(data && dark && flat && out, "%s: NULL array", who);
  TOMO_REQUIRE(out_dtype == TOMO_F32 || out_dtype == TOMO_F16 || out_dtype == TOMO_BF16,
               "%s: unknown out_dtype %d", who, out_dtype);
  hipStream_t st = (hipStream_t)stream;
  int rc = launch_means(P, dark, flat, bright_ratio, flat_linear, st);
  if (rc) return rc;
  {
    ProfScope ps_("k_pre_rows", st);
    if (out_dtype == TOMO_F16)
      launch_rows0<_Float16>(P, data, out, flat_linear, minus_log, st);
    else if (out_dtype == TOMO_BF16)
      launch_rows0<__bf16>(P, data, out, flat_linear, minus_log, st);
    else
      launch_rows0<float>(P, data, out, flat_linear, minus_log, st);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

template <typename TI>
static int darkflat(const char *who, tomo_pre_plan *P, const TI *data, const TI *dark,
                    const TI *flat, float *res, double bright_ratio, int flat_linear,
                    void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "%s: plan is NULL", who);
  TOMO_REQUIRE(data && dark && flat && res, "%s: NULL array", who);
  hipStream_t st = (hipStream_t)stream;
  int rc = launch_means(P, dark, flat, bright_ratio, flat_linear, st);
  if (rc) return rc;
  {
    ProfScope ps_("k_pre_rows", st);
    hipLaunchKernelGGL((k_pre_rows<1, float, TI>), dim3((unsigned)P->nproj, (unsigned)P->nz),
                       dim3(256), 0, st, data, P->dark0.as<float>(), P->flatA.as<float>(),
                       P->flatB.as<float>(), res, (int)P->nproj, (int)P->nz, (int)P->n,
                       flat_linear, 0);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

extern "C" int pre_process(tomo_pre_plan *P, const uint16_t *data, const uint16_t *dark,
                           const uint16_t *flat, void *out, int out_dtype, double bright_ratio,
                           int flat_linear, int minus_log, void *stream) {
  return process("pre_process", P, data, dark, flat, out, out_dtype, bright_ratio, flat_linear,
                 minus_log, stream);
}

extern "C" int pre_process_f32(tomo_pre_plan *P, const float *data, const float *dark,
                               const float *flat, void *out, int out_dtype, double bright_ratio,
                               int flat_linear, int minus_log, void *stream) {
  return process("pre_process_f32", P, data, dark, flat, out, out_dtype, bright_ratio,
                 flat_linear, minus_log, stream);
}

extern "C" int pre_darkflat(tomo_pre_plan *P, const uint16_t *data, const uint16_t *dark,
                            const uint16_t *flat, float *res, double bright_ratio, int flat_linear,
                            void *stream) {
  return darkflat("pre_darkflat", P, data, dark, flat, res, bright_ratio, flat_linear, stream);
}

extern "C" int pre_darkflat_f32(tomo_pre_plan *P, const float *data, const float *dark,
                                const float *flat, float *res, double bright_ratio,
                                int flat_linear, void *stream) {
  return darkflat("pre_darkflat_f32", P, data, dark, flat, res, bright_ratio, flat_linear,
                  stream);
}

static int log_swap(const char *who, tomo_pre_plan *P, const float *res, int64_t ld, int64_t z0,
                    int64_t lz, void *out, int out_dtype, int minus_log, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "%s: plan is NULL", who);
  TOMO_REQUIRE(res && out, "%s: NULL array", who);
  TOMO_REQUIRE(out_dtype == TOMO_F32 || out_dtype == TOMO_F16 || out_dtype == TOMO_BF16,
               "%s: unknown out_dtype %d", who, out_dtype);
  TOMO_REQUIRE(res != out, "%s: in-place swap is not supported", who);
  TOMO_REQUIRE(ld >= 1 && ld <= 0x7fffffff && z0 >= 0 && lz >= 0 && lz <= P->nz && z0 + lz <= ld,
               "%s: rows %lld..%lld outside the %lld rows of the volume", who, (long long)z0,
               (long long)(z0 + lz), (long long)ld);
  hipStream_t st = (hipStream_t)stream;
  {
    ProfScope ps_("k_pre_log_swap", st);
    if (out_dtype == TOMO_F16)
      launch_log_swap<_Float16>(P, res, ld, z0, lz, out, minus_log, st);
    else if (out_dtype == TOMO_BF16)
      launch_log_swap<__bf16>(P, res, ld, z0, lz, out, minus_log, st);
    else
      launch_log_swap<float>(P, res, ld, z0, lz, out, minus_log, st);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

extern "C" int pre_minus_log_swap(tomo_pre_plan *P, const float *res, void *out, int out_dtype,
                                  int minus_log, void *stream) {
  return log_swap("pre_minus_log_swap", P, res, P ? P->nz : 0, 0, P ? P->nz : 0, out, out_dtype,
                  minus_log, stream);
}

extern "C" int pre_minus_log_swap_rows(tomo_pre_plan *P, const float *vol, int64_t ld,
                                       int64_t z0, int64_t lz, void *out, int out_dtype,
                                       int minus_log, void *stream) {
  return log_swap("pre_minus_log_swap_rows", P, vol, ld, z0, lz, out, out_dtype, minus_log,
                  stream);
}

// ---------------------------------------------------------------------------
// Allocation-free preprocessing: pre_process without a plan, its dark/flat
// means in a static device array instead of plan-owned scratch.

constexpr int64_t kNoAllocMaxZN = (int64_t)1 << 20;  // nz * n of one call

// dark0, flatA, flatB of one call, kNoAllocMaxZN floats each (12 MiB per
// GPU, held until the process ends).  Host code never passes this name to a
// kernel: it is a host-side stand-in; the per-device address comes from
// hipGetSymbolAddress below.
static __device__ __align__(256) float scratch[3 * kNoAllocMaxZN];

namespace {

constexpr int kNoAllocMaxDevices = 64;

std::mutex g_scratch_mu;
float *g_scratch_dev[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_address(float **out) {
  int dev = -1;
  TOMO_HIP(hipGetDevice(&dev));
  TOMO_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "pre_process_noalloc: device %d out of range",
               dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    TOMO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TOMO_REQUIRE(p != nullptr, "pre_process_noalloc: no address for the scratch array");
    g_scratch_dev[dev] = static_cast<float *>(p);
  }
  *out = g_scratch_dev[dev];
  return TOMO_OK;
}

}  // namespace

extern "C" int pre_process_noalloc(int64_t nproj, int64_t nz, int64_t n, int64_t ndark,
                                   int64_t nflat, const uint16_t *data, const uint16_t *dark,
                                   const uint16_t *flat, void *out, int out_dtype,
                                   double bright_ratio, int flat_linear, int minus_log,
                                   void *stream) {
  clear_error();
  TOMO_REQUIRE(nproj >= 1 && nz >= 1 && n >= 1, "pre_process_noalloc: sizes must be >= 1");
  TOMO_REQUIRE(ndark >= 1 && nflat >= 1,
               "pre_process_noalloc: need at least one dark and one flat frame");
  TOMO_REQUIRE(nproj <= 65535 && nz <= 65535, "pre_process_noalloc: nproj and nz must be <= 65535");
  TOMO_REQUIRE(nz * n <= kNoAllocMaxZN,
               "pre_process_noalloc: nz*n=%lld exceeds the static scratch (nz*n <= %lld)",
               (long long)(nz * n), (long long)kNoAllocMaxZN);
  TOMO_REQUIRE(data && dark && flat && out, "pre_process_noalloc: NULL array");
  TOMO_REQUIRE(out_dtype == TOMO_F32 || out_dtype == TOMO_F16 || out_dtype == TOMO_BF16,
               "pre_process_noalloc: unknown out_dtype %d", out_dtype);
  TOMO_REQUIRE(bright_ratio > 0 && std::isfinite(bright_ratio),
               "pre_process_noalloc: bright_ratio must be > 0");
  TOMO_REQUIRE(!flat_linear || nflat >= 3,
               "pre_process_noalloc: flat_linear needs at least 3 flat frames");
  float *sc = nullptr;
  int rc = scratch_address(&sc);
  if (rc) return rc;
  // k_pre_means writes every mean it produces (flatB only with flat_linear,
  // the only case that reads it), so nothing needs clearing between calls
  float *d0 = sc, *fa = sc + kNoAllocMaxZN, *fb = sc + 2 * kNoAllocMaxZN;
  hipStream_t st = (hipStream_t)stream;
  const int zn = (int)(nz * n);
  hipLaunchKernelGGL(k_pre_means<uint16_t>, dim3((unsigned)ceil_div(zn, kMeansThreads)),
                     dim3(kMeansThreads), 0, st, dark, flat, d0, fa, fb, zn, (int)ndark,
                     (int)nflat, (float)bright_ratio, flat_linear);
  TOMO_LAUNCH_CHECK();
  const dim3 grid((unsigned)nproj, (unsigned)nz);
  if (out_dtype == TOMO_F16)
    hipLaunchKernelGGL((k_pre_rows<0, _Float16, uint16_t>), grid, dim3(256), 0, st, data, d0, fa,
                       fb, (_Float16 *)out, (int)nproj, (int)nz, (int)n, flat_linear, minus_log);
  else if (out_dtype == TOMO_BF16)
    hipLaunchKernelGGL((k_pre_rows<0, __bf16, uint16_t>), grid, dim3(256), 0, st, data, d0, fa,
                       fb, (__bf16 *)out, (int)nproj, (int)nz, (int)n, flat_linear, minus_log);
  else
    hipLaunchKernelGGL((k_pre_rows<0, float, uint16_t>), grid, dim3(256), 0, st, data, d0, fa, fb,
                       (float *)out, (int)nproj, (int)nz, (int)n, flat_linear, minus_log);
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

extern "C" void pre_destroy(tomo_pre_plan *P) {
  if (!P) return;
  P->dark0.release();
  P->flatA.release();
  P->flatB.release();
  for (auto &l : P->lv) {
    l.tw.release();
    l.gk.release();
    l.ll.release();
    l.part.release();
    l.X.release();
  }
  P->rowlo.release();
  P->rowhi.release();
  P->synlo.release();
  P->synhi.release();
  P->syn[0].release();
  P->syn[1].release();
  P->outl.release();
  P->ti.release();
  P->vo.release();
  delete P;
}
