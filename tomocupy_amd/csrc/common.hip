// Status handling, version string and FFT plan construction for libtomohip.
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "common.h"
#include "fft_lds.h"

namespace tomo {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

size_t max_lds_per_block() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 65536;
  return (size_t)v;
}

int FftPlan::radices(int N, FftDesc &d) {
  TOMO_REQUIRE(fft_size_ok(N) && N <= 16384,
               "FFT length %d is not <= 16384 with prime factors <= %d", N, kFftMaxPrime);
  d.N = N;
  d.nr = 0;
  d.tw = nullptr;
  int m = N;
  // radix 16 first (fewest passes), then 8, 4, 2, then 3s, 5s, 7s, then the
  // larger primes
  while (m % 16 == 0) { d.r[d.nr++] = 16; m /= 16; }
  while (m % 8 == 0) { d.r[d.nr++] = 8; m /= 8; }
  while (m % 4 == 0) { d.r[d.nr++] = 4; m /= 4; }
  while (m % 2 == 0) { d.r[d.nr++] = 2; m /= 2; }
  while (m % 3 == 0) { d.r[d.nr++] = 3; m /= 3; }
  while (m % 5 == 0) { d.r[d.nr++] = 5; m /= 5; }
  while (m % 7 == 0) { d.r[d.nr++] = 7; m /= 7; }
  for (int p = 11; p <= kFftMaxPrime && m > 1; p += 2)  // generic prime passes
    while (m % p == 0) { d.r[d.nr++] = p; m /= p; }
  return TOMO_OK;
}

int FftPlan::init(int N) {
  int rc = radices(N, d);
  if (rc) return rc;
  std::vector<float2> h(N);
  for (int k = 0; k < N; ++k) {
    double a = -2.0 * M_PI * (double)k / (double)N;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  rc = tw.alloc(sizeof(float2) * N);
  if (rc) return rc;
  TOMO_HIP(hipMemcpy(tw.p, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  d.tw = tw.as<float2>();
  return TOMO_OK;
}

static std::mutex g_prof_mu;
static std::atomic<bool> g_prof_on{false};
static std::string g_prof_only;  // empty: every kernel
static std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> g_prof_pending;
static std::vector<hipEvent_t> g_prof_pool;
static std::map<std::string, std::pair<double, int64_t>> g_prof_acc;

bool prof_wants(const char *name) {
  if (!g_prof_on.load(std::memory_order_relaxed)) return false;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  return g_prof_only.empty() || g_prof_only == name;
}

bool prof_take(hipEvent_t *a, hipEvent_t *b) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  hipEvent_t e[2];
  for (auto &x : e) {
    if (!g_prof_pool.empty()) {
      x = g_prof_pool.back();
      g_prof_pool.pop_back();
    } else if (hipEventCreate(&x) != hipSuccess) {
      if (&x != e) g_prof_pool.push_back(e[0]);
      return false;
    }
  }
  *a = e[0];
  *b = e[1];
  return true;
}

void prof_push(const char *name, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_pending.emplace_back(name, a, b);
}

static int prof_drain() {
  for (auto &e : g_prof_pending) {
    hipEvent_t a = std::get<1>(e), b = std::get<2>(e);
    float ms = 0.f;
    hipError_t err = hipEventSynchronize(b);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
    if (err == hipSuccess) {
      auto &acc = g_prof_acc[std::get<0>(e)];
      acc.first += ms;
      acc.second += 1;
    }
    g_prof_pool.push_back(a);
    g_prof_pool.push_back(b);
  }
  g_prof_pending.clear();
  return TOMO_OK;
}

}  // namespace tomo

extern "C" int tomo_prof_enable(int on) {
  tomo::g_prof_on.store(on != 0);
  return TOMO_OK;
}

extern "C" int tomo_prof_only(const char *kernel) {
  std::lock_guard<std::mutex> lk(tomo::g_prof_mu);
  tomo::g_prof_only = kernel ? kernel : "";
  return TOMO_OK;
}

extern "C" int tomo_prof_reset(void) {
  std::lock_guard<std::mutex> lk(tomo::g_prof_mu);
  tomo::prof_drain();
  tomo::g_prof_acc.clear();
  return TOMO_OK;
}

extern "C" int tomo_prof_read(const char *kernel, double *total_ms, int64_t *launches) {
  tomo::clear_error();
  TOMO_REQUIRE(kernel != nullptr, "tomo_prof_read: kernel name is NULL");
  std::lock_guard<std::mutex> lk(tomo::g_prof_mu);
  tomo::prof_drain();
  auto it = tomo::g_prof_acc.find(kernel);
  if (total_ms) *total_ms = it == tomo::g_prof_acc.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == tomo::g_prof_acc.end() ? 0 : it->second.second;
  return TOMO_OK;
}

extern "C" const char *tomo_last_error(void) { return tomo::g_err; }
extern "C" const char *tomo_version(void) { return "tomohip 0.1.0 (gfx950)"; }

namespace {

constexpr int kDotBlocks = 1024;

// Block b sums a[i] b[i] for i = b*256 + t + k*(1024*256), in FP64, then a
// fixed-order tree over the block's threads.
__global__ __launch_bounds__(256) void k_dot_partial(const float *__restrict__ a,
                                                     const float *__restrict__ b, int64_t n,
                                                     double *__restrict__ part) {
  __shared__ double sd[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kDotBlocks * 256)
    s = fma((double)a[i], (double)b[i], s);
  sd[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sd[threadIdx.x] += sd[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sd[0];
}

__global__ __launch_bounds__(256) void k_dot_final(double *__restrict__ out) {
  __shared__ double sd[256];
  const double *part = out + 1;
  double s = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += 256) s += part[i];
  sd[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sd[threadIdx.x] += sd[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sd[0];
}

// CG step: x += alpha p, r -= alpha q, then the FP64 partial sums of r.r over
// the same elements, in the same order and tree as k_dot_partial.
__global__ __launch_bounds__(256) void k_cg_step(float *__restrict__ x, float *__restrict__ r,
                                                 const float *__restrict__ p,
                                                 const float *__restrict__ q, int64_t n,
                                                 const double *__restrict__ rr,
                                                 const double *__restrict__ pq,
                                                 double *__restrict__ part) {
  __shared__ double sd[256];
  const double den = pq[0];
  const float alpha = den > 0.0 ? (float)(rr[0] / den) : 0.f;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)kDotBlocks * 256) {
    x[i] = fmaf(alpha, p[i], x[i]);
    const float ri = fmaf(-alpha, q[i], r[i]);
    r[i] = ri;
    s = fma((double)ri, (double)ri, s);
  }
  sd[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sd[threadIdx.x] += sd[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sd[0];
}

// CG direction: p = r + beta p.
__global__ __launch_bounds__(256) void k_cg_direction(float *__restrict__ p,
                                                      const float *__restrict__ r, int64_t n,
                                                      const double *__restrict__ rr,
                                                      const double *__restrict__ rr_new) {
  const double den = rr[0];
  const float beta = den > 0.0 ? (float)(rr_new[0] / den) : 0.f;
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x, stride = (int64_t)gridDim.x * 256;
  // 16-byte accesses for the aligned bulk, then the tail element by element
  const bool al = (((uintptr_t)p | (uintptr_t)r) & 15) == 0;
  const int64_t n4 = al ? n / 4 : 0;
  float4 *p4 = reinterpret_cast<float4 *>(p);
  const float4 *r4 = reinterpret_cast<const float4 *>(r);
  for (int64_t i = gid; i < n4; i += stride) {
    float4 a = p4[i];
    const float4 b = r4[i];
    a.x = fmaf(beta, a.x, b.x);
    a.y = fmaf(beta, a.y, b.y);
    a.z = fmaf(beta, a.z, b.z);
    a.w = fmaf(beta, a.w, b.w);
    p4[i] = a;
  }
  for (int64_t i = 4 * n4 + gid; i < n; i += stride) p[i] = fmaf(beta, p[i], r[i]);
}

}  // namespace

extern "C" int tomo_cg_step(float *x, float *r, const float *p, const float *q, int64_t n,
                            const double *rr, const double *pq, double *out, void *stream) {
  tomo::clear_error();
  TOMO_REQUIRE(x && r && p && q && rr && pq && out, "tomo_cg_step: NULL array");
  TOMO_REQUIRE(n >= 0, "tomo_cg_step: n < 0");
  TOMO_REQUIRE(out != rr && out != pq, "tomo_cg_step: out aliases rr or pq");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cg_step, dim3(kDotBlocks), dim3(256), 0, st, x, r, p, q, n, rr, pq,
                     out + 1);
  TOMO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, st, out);
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

extern "C" int tomo_cg_direction(float *p, const float *r, int64_t n, const double *rr,
                                 const double *rr_new, void *stream) {
  tomo::clear_error();
  TOMO_REQUIRE(p && r && rr && rr_new, "tomo_cg_direction: NULL array");
  TOMO_REQUIRE(n >= 0, "tomo_cg_direction: n < 0");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cg_direction, dim3(kDotBlocks), dim3(256), 0, st, p, r, n, rr, rr_new);
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

extern "C" int tomo_dot_f64(const float *a, const float *b, int64_t n, double *out,
                            void *stream) {
  tomo::clear_error();
  TOMO_REQUIRE(a != nullptr && b != nullptr && out != nullptr, "tomo_dot_f64: NULL array");
  TOMO_REQUIRE(n >= 0, "tomo_dot_f64: n < 0");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(256), 0, st, a, b, n, out + 1);
  TOMO_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, st, out);
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}
