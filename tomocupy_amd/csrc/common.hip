// Status handling, version string and FFT plan construction for libtomohip.
#include <cmath>
#include <cstring>

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

int FftPlan::init(int N) {
  TOMO_REQUIRE(fft_size_ok(N) && N <= 16384, "FFT length %d is not 2^a 3^b <= 16384", N);
  d.N = N;
  d.nr = 0;
  int m = N;
  // radix 8 first (fewest passes), then 4, 2, then 3s
  while (m % 8 == 0 && m > 8) { d.r[d.nr++] = 8; m /= 8; }
  while (m % 4 == 0) { d.r[d.nr++] = 4; m /= 4; }
  while (m % 2 == 0) { d.r[d.nr++] = 2; m /= 2; }
  while (m % 3 == 0) { d.r[d.nr++] = 3; m /= 3; }
  std::vector<float2> h(N);
  for (int k = 0; k < N; ++k) {
    double a = -2.0 * M_PI * (double)k / (double)N;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  int rc = tw.alloc(sizeof(float2) * N);
  if (rc) return rc;
  TOMO_HIP(hipMemcpy(tw.p, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  d.tw = tw.as<float2>();
  return TOMO_OK;
}

}  // namespace tomo

extern "C" const char *tomo_last_error(void) { return tomo::g_err; }
extern "C" const char *tomo_version(void) { return "tomohip 0.1.0 (gfx950)"; }
