// Double field of view (tomocupy ProcFunctions.pad360,
// src/tomocupy/processing/proc_functions.py:109-124, applied by proc_proj for
// file_type double_fov, :193-194) for gfx950, on sinograms of width ni:
//   flip the detector axis when the rotation axis is in the left half,
//   taper the last w = max(1, int(2 (ni - center))) columns by
//   v = t^5 (126 - 420 t + 540 t^2 - 315 t^3 + 70 t^4), t = 1 - i/w,
//   and pad with zeros to width 2 ni.
// Written from sino (nz, nproj, ni) into out (nz, nproj, 2 ni), same dtype;
// the taper is evaluated and applied in FP64 and rounded once.
#include <cmath>
#include <cstdint>

#include "common.h"
#include "precision.h"
#include "preproc.h"

using namespace tomo;

namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_pre_pad360(const T *__restrict__ sino,
                                                    T *__restrict__ out, int rows, int ni,
                                                    int flip, int w) {
  const int row = blockIdx.x;  // (z, p) flattened
  if (row >= rows) return;
  const T *src = sino + (size_t)row * ni;
  T *dst = out + (size_t)row * 2 * ni;
  for (int x = threadIdx.x; x < 2 * ni; x += 256) {
    double v = 0.0;
    if (x < ni) {
      v = (double)ldr(src + (flip ? ni - 1 - x : x));
      const int i = x - (ni - w);
      if (i >= 0) {
        const double t = 1.0 - (double)i / (double)w;
        const double t2 = t * t;
        v *= t2 * t2 * t * (126.0 - 420.0 * t + 540.0 * t2 - 315.0 * t2 * t + 70.0 * t2 * t2);
      }
    }
    str(dst + x, (float)v);
  }
}

}  // namespace

extern "C" int pre_pad360(tomo_pre_plan *P, const void *sino, void *out, int dtype,
                          double center, int flip, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "pre_pad360: plan is NULL");
  TOMO_REQUIRE(sino != nullptr && out != nullptr && sino != out, "pre_pad360: bad arrays");
  TOMO_REQUIRE(dtype == TOMO_F32 || dtype == TOMO_F16 || dtype == TOMO_BF16,
               "pre_pad360: unknown dtype %d", dtype);
  const int ni = (int)P->n;
  TOMO_REQUIRE(std::isfinite(center) && center >= ni / 2,
               "pre_pad360: center=%g must be >= ni/2 (ni=%d) after the flip rule", center, ni);
  // w = max(1, int(2 (ni - center))) with Python's truncation toward zero
  const int w = std::max(1, (int)(2.0 * ((double)ni - center)));
  TOMO_REQUIRE(w <= ni, "pre_pad360: taper width %d exceeds ni=%d", w, ni);
  const int rows = (int)(P->nz * P->nproj);
  hipStream_t st = (hipStream_t)stream;
  {
    ProfScope ps_("k_pre_pad360", st);
    if (dtype == TOMO_F16)
      hipLaunchKernelGGL(k_pre_pad360<_Float16>, dim3(rows), dim3(256), 0, st,
                         (const _Float16 *)sino, (_Float16 *)out, rows, ni, flip, w);
    else if (dtype == TOMO_BF16)
      hipLaunchKernelGGL(k_pre_pad360<__bf16>, dim3(rows), dim3(256), 0, st, (const __bf16 *)sino,
                         (__bf16 *)out, rows, ni, flip, w);
    else
      hipLaunchKernelGGL(k_pre_pad360<float>, dim3(rows), dim3(256), 0, st, (const float *)sino,
                         (float *)out, rows, ni, flip, w);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}
