// Shared host/device helpers for libtomohip: status handling, complex
// arithmetic, launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/tomohip.h"

namespace tomo {

// ---- status --------------------------------------------------------------
void set_error(const char *fmt, ...);
void clear_error();

#define TOMO_REQUIRE(cond, ...)              \
  do {                                       \
    if (!(cond)) {                           \
      ::tomo::set_error(__VA_ARGS__);        \
      return TOMO_EINVAL;                    \
    }                                        \
  } while (0)

#define TOMO_HIP(expr)                                                            \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::tomo::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                      \
      return e_ == hipErrorOutOfMemory ? TOMO_ENOMEM : TOMO_EHIP;                \
    }                                                                             \
  } while (0)

#define TOMO_LAUNCH_CHECK() TOMO_HIP(hipGetLastError())

// Device buffer owned by a plan.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int alloc(size_t nbytes) {
    release();
    if (nbytes == 0) return TOMO_OK;
    TOMO_HIP(hipMalloc(&p, nbytes));
    bytes = nbytes;
    return TOMO_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// ---- complex helpers -----------------------------------------------------
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__host__ __device__ inline float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__host__ __device__ inline float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Largest dynamic LDS a workgroup may use on the current device.
size_t max_lds_per_block();

// ---- kernel timing (tomo_prof_*) -----------------------------------------
// When enabled, every launch wrapped in a ProfScope whose kernel name is
// selected (all names, or the one given to tomo_prof_only) is bracketed by
// two HIP events recorded on the launch stream; durations are collected
// lazily by tomo_prof_read().  Events come from a pool and are reused.
bool prof_wants(const char *name);
bool prof_take(hipEvent_t *a, hipEvent_t *b);
void prof_push(const char *name, hipEvent_t a, hipEvent_t b);

struct ProfScope {
  const char *name;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(const char *nm, hipStream_t s) : name(nm), st(s) {
    if (prof_wants(name) && prof_take(&a, &b))
      (void)hipEventRecord(a, st);
    else
      a = b = nullptr;
  }
  ~ProfScope() {
    if (a && b) {
      (void)hipEventRecord(b, st);
      prof_push(name, a, b);
    }
  }
};

}  // namespace tomo
