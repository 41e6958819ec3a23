// Storage precisions of the fourierrec path.  Arithmetic is float32 in
// registers and LDS; sinograms, spectra, grids and images are stored in the
// plan's precision (float, IEEE half or bfloat16) and converted on load and
// store (round to nearest even).
#pragma once

#include <hip/hip_runtime.h>

namespace tomo {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

// Conversion of a loaded (raw) complex value; loading raw values first and
// converting afterwards lets a thread keep several loads in flight.
template <typename T>
__device__ __forceinline__ float2 tof2(cpx<T> v) {
  return make_float2((float)v.x, (float)v.y);
}

template <typename T>
__device__ __forceinline__ float ldr(const T *p) {
  return (float)*p;
}
template <typename T>
__device__ __forceinline__ void str(T *p, float v) {
  *p = (T)v;
}
template <typename T>
__device__ __forceinline__ float2 ldc(const cpx<T> *p) {
  const cpx<T> v = *p;
  return make_float2((float)v.x, (float)v.y);
}
template <typename T>
__device__ __forceinline__ void stc(cpx<T> *p, float2 v) {
  *p = cpx<T>{(T)v.x, (T)v.y};
}
template <>
__device__ __forceinline__ float2 ldc<float>(const cpx<float> *p) {
  return *reinterpret_cast<const float2 *>(p);
}
template <>
__device__ __forceinline__ void stc<float>(cpx<float> *p, float2 v) {
  *reinterpret_cast<float2 *>(p) = v;
}

}  // namespace tomo
