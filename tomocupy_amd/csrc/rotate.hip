// Projection rotation for gfx950 (tomocupy ProcFunctions.rotate_proj,
// src/tomocupy/processing/proc_functions.py:126-138: ndimage.rotate(data,
// angle, reshape=False, order=order, axes=(2, 1), mode='nearest'), applied by
// proc_proj, :182-184): every projection (rows y, columns x) is resampled at
//   (iy, ix) = R (y, x) + offset,  R = [[c, s], [-s, c]],
//   offset = centre - R centre,  centre = ((ny - 1)/2, (nx - 1)/2),
// c, s = cos, sin of the angle in degrees, with order 0 (nearest sample) or
// 1 (bilinear), neighbours clamped to the image ('nearest' extension).
// Coordinates in FP64, interpolation in float32.
#include <cmath>
#include <cstdint>

#include "common.h"

using namespace tomo;

namespace {

template <int ORDER>
__global__ __launch_bounds__(256) void k_pre_rotate(const float *__restrict__ src,
                                                    float *__restrict__ dst, int ny, int nx,
                                                    double c, double s, double oy, double ox) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, p = blockIdx.z;
  if (x >= nx) return;
  const float *img = src + (size_t)p * ny * nx;
  const double iy = c * y + s * x + oy, ix = -s * y + c * x + ox;
  float v;
  if (ORDER == 0) {
    const int jy = min(max((int)floor(iy + 0.5), 0), ny - 1);
    const int jx = min(max((int)floor(ix + 0.5), 0), nx - 1);
    v = img[(size_t)jy * nx + jx];
  } else {
    const double fy0 = floor(iy), fx0 = floor(ix);
    const float wy = (float)(iy - fy0), wx = (float)(ix - fx0);
    // clamp in FP64 before the conversion: far coordinates stay in int range
    const int y0 = (int)fmin(fmax(fy0, 0.0), (double)(ny - 1));
    const int y1 = (int)fmin(fmax(fy0 + 1.0, 0.0), (double)(ny - 1));
    const int x0 = (int)fmin(fmax(fx0, 0.0), (double)(nx - 1));
    const int x1 = (int)fmin(fmax(fx0 + 1.0, 0.0), (double)(nx - 1));
    const float a = img[(size_t)y0 * nx + x0], b = img[(size_t)y0 * nx + x1];
    const float cc = img[(size_t)y1 * nx + x0], d = img[(size_t)y1 * nx + x1];
    v = (1.f - wy) * ((1.f - wx) * a + wx * b) + wy * ((1.f - wx) * cc + wx * d);
  }
  dst[((size_t)p * ny + y) * nx + x] = v;
}

// cos and sin of an angle in degrees, exact at multiples of 90 (as
// scipy.special.cosdg / sindg)
void cos_sin_deg(double a, double *c, double *s) {
  double r = std::fmod(a, 360.0);
  if (r < 0) r += 360.0;
  if (r == 0.0) { *c = 1; *s = 0; return; }
  if (r == 90.0) { *c = 0; *s = 1; return; }
  if (r == 180.0) { *c = -1; *s = 0; return; }
  if (r == 270.0) { *c = 0; *s = -1; return; }
  const double t = r * (M_PI / 180.0);
  *c = std::cos(t);
  *s = std::sin(t);
}

}  // namespace

extern "C" int pre_rotate_proj(const float *src, float *dst, int64_t nproj, int64_t ny,
                               int64_t nx, double angle, int order, void *stream) {
  clear_error();
  TOMO_REQUIRE(src && dst, "pre_rotate_proj: NULL array");
  TOMO_REQUIRE(src != dst, "pre_rotate_proj: in-place rotation is not supported");
  TOMO_REQUIRE(nproj >= 0 && nproj <= 65535 && ny >= 1 && ny <= 65535 && nx >= 1 &&
                   nx <= 0x7fffffff,
               "pre_rotate_proj: sizes out of range");
  TOMO_REQUIRE(std::isfinite(angle), "pre_rotate_proj: angle must be finite");
  if (order < 0 || order > 1) {
    set_error("pre_rotate_proj: interpolation order %d not supported (0 or 1; the spline "
              "orders 2-5 are not implemented)", order);
    return order < 0 ? TOMO_EINVAL : TOMO_EUNSUPPORTED;
  }
  if (nproj == 0) return TOMO_OK;
  double c, s;
  cos_sin_deg(angle, &c, &s);
  const double cy = (double)(ny - 1) / 2, cx = (double)(nx - 1) / 2;
  const double oy = cy - (c * cy + s * cx), ox = cx - (-s * cy + c * cx);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)ceil_div(nx, 256), (unsigned)ny, (unsigned)nproj);
  {
    ProfScope ps_("k_pre_rotate", st);
    if (order == 0)
      hipLaunchKernelGGL(k_pre_rotate<0>, grid, dim3(256), 0, st, src, dst, (int)ny, (int)nx, c,
                         s, oy, ox);
    else
      hipLaunchKernelGGL(k_pre_rotate<1>, grid, dim3(256), 0, st, src, dst, (int)ny, (int)nx, c,
                         s, oy, ox);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}
