// Projection rotation for gfx950 (tomocupy ProcFunctions.rotate_proj,
// src/tomocupy/processing/proc_functions.py:126-138: ndimage.rotate(data,
// angle, reshape=False, order=order, axes=(2, 1), mode='nearest'), applied by
// proc_proj, :182-184): every projection (rows y, columns x) is resampled at
//   (iy, ix) = R (y, x) + offset,  R = [[c, s], [-s, c]],
//   offset = centre - R centre,  centre = ((ny - 1)/2, (nx - 1)/2),
// c, s = cos, sin of the angle in degrees, with order 0 (nearest sample),
// 1 (bilinear), neighbours clamped to the image ('nearest' extension), or 2
// and 3 (quadratic, cubic B-splines on the prefiltered, edge-padded plane, as
// SciPy's spline path).  Coordinates in FP64, interpolation in float32.
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

// B-spline prefilter of order 2 or 3 (single pole z: sqrt(8) - 3, sqrt(3) - 2)
// on the plane edge-padded by kSplinePad samples (as SciPy's
// _prepad_for_spline_filter does for mode 'nearest'), the recursive filter of
// Unser, Aldroubi and Eden (IEEE TSP 41, 1993) with mirror (whole-sample
// symmetric) initial values as in Thevenaz, Blu and Unser (IEEE TMI 19,
// 2000): gain (1 - z)(1 - 1/z), causal c[k] += z c[k-1] from a truncated
// mirror sum, anticausal c[k] = z (c[k+1] - c[k]) from
// c[N-1] = z/(z^2 - 1) (z c[N-2] + c[N-1]).  The initial values only reach the
// data through |z|^12 (<= 1.4e-7).
constexpr int kSplinePad = 12;

__device__ __forceinline__ float spline_pole(int order) {
  return order == 2 ? -0.17157287525380990f : -0.26794919243112270f;
}

// in place along a strided sequence c[0], c[st], ... of length N (gain applied)
__device__ void spline_1d(float *c, int N, size_t st, float z) {
  const float lambda = (1.f - z) * (1.f - 1.f / z);
  if (N < 2) return;
  for (int k = 0; k < N; ++k) c[k * st] *= lambda;
  // causal initial value: sum_k z^k c[k] until z^k < 1e-9
  float zk = 1.f, sum = 0.f;
  for (int k = 0; k < N && fabsf(zk) > 1e-9f; ++k) {
    sum += zk * c[k * st];
    zk *= z;
  }
  c[0] = sum;
  for (int k = 1; k < N; ++k) c[k * st] += z * c[(k - 1) * st];
  c[(N - 1) * st] = (z / (z * z - 1.f)) * (z * c[(N - 2) * st] + c[(N - 1) * st]);
  for (int k = N - 2; k >= 0; --k) c[k * st] = z * (c[(k + 1) * st] - c[k * st]);
}

// columns of the padded plane: work[p][yc][xc] = spline along yc of
// data[p][clamp(yc - pad)][clamp(xc - pad)], one thread per padded column
__global__ __launch_bounds__(256) void k_pre_spline_cols(const float *__restrict__ src,
                                                         float *__restrict__ work, int ny,
                                                         int nx, int order) {
  const int Py = ny + 2 * kSplinePad, Px = nx + 2 * kSplinePad;
  const int xc = blockIdx.x * 256 + threadIdx.x, p = blockIdx.y;
  if (xc >= Px) return;
  const int x = min(max(xc - kSplinePad, 0), nx - 1);
  const float *img = src + (size_t)p * ny * nx;
  float *w = work + (size_t)p * Py * Px + xc;
  for (int yc = 0; yc < Py; ++yc)
    w[(size_t)yc * Px] = img[(size_t)min(max(yc - kSplinePad, 0), ny - 1) * nx + x];
  spline_1d(w, Py, (size_t)Px, spline_pole(order));
}

// rows of the padded plane, in place, one thread per padded row
__global__ __launch_bounds__(256) void k_pre_spline_rows(float *__restrict__ work, int ny,
                                                         int nx, int order) {
  const int Py = ny + 2 * kSplinePad, Px = nx + 2 * kSplinePad;
  const int yc = blockIdx.x * 256 + threadIdx.x, p = blockIdx.y;
  if (yc >= Py) return;
  spline_1d(work + ((size_t)p * Py + yc) * Px, Px, 1, spline_pole(order));
}

// B-spline evaluation of order 2 or 3 at the rotated coordinates, on the
// coefficients of the padded plane (indices clamped to it)
template <int ORDER>
__global__ __launch_bounds__(256) void k_pre_rotate_spline(const float *__restrict__ work,
                                                           float *__restrict__ dst, int ny,
                                                           int nx, double c, double s, double oy,
                                                           double ox) {
  const int Py = ny + 2 * kSplinePad, Px = nx + 2 * kSplinePad;
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, p = blockIdx.z;
  if (x >= nx) return;
  const float *cf = work + (size_t)p * Py * Px;
  const double iy = c * y + s * x + oy + kSplinePad, ix = -s * y + c * x + ox + kSplinePad;
  constexpr int K = ORDER + 1;
  float wy[K], wx[K];
  double by, bx;  // first neighbour
  auto weights = [&](double u, float *w, double &b) {
    if (ORDER == 3) {
      const double f = floor(u);
      const float t = (float)(u - f);
      b = f - 1.0;
      const float t2 = t * t, t3 = t2 * t, m = 1.f - t;
      w[0] = m * m * m / 6.f;
      w[1] = (4.f - 6.f * t2 + 3.f * t3) / 6.f;
      w[2] = (1.f + 3.f * t + 3.f * t2 - 3.f * t3) / 6.f;
      w[3] = t3 / 6.f;
    } else {
      const double f = floor(u + 0.5);
      const float t = (float)(u - f);
      b = f - 1.0;
      w[0] = 0.5f * (0.5f - t) * (0.5f - t);
      w[1] = 0.75f - t * t;
      w[2] = 0.5f * (0.5f + t) * (0.5f + t);
    }
  };
  weights(iy, wy, by);
  weights(ix, wx, bx);
  float v = 0.f;
#pragma unroll
  for (int a = 0; a < K; ++a) {
    const int yy = (int)fmin(fmax(by + a, 0.0), (double)(Py - 1));
    float r = 0.f;
#pragma unroll
    for (int b = 0; b < K; ++b) {
      const int xx = (int)fmin(fmax(bx + b, 0.0), (double)(Px - 1));
      r += wx[b] * cf[(size_t)yy * Px + xx];
    }
    v += wy[a] * r;
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
                               int64_t nx, double angle, int order, float *work, void *stream) {
  clear_error();
  TOMO_REQUIRE(src && dst, "pre_rotate_proj: NULL array");
  TOMO_REQUIRE(src != dst, "pre_rotate_proj: in-place rotation is not supported");
  TOMO_REQUIRE(nproj >= 0 && nproj <= 65535 && ny >= 1 && ny <= 65535 && nx >= 1 &&
                   nx <= 0x7fffffff,
               "pre_rotate_proj: sizes out of range");
  TOMO_REQUIRE(std::isfinite(angle), "pre_rotate_proj: angle must be finite");
  if (order < 0 || order > 3) {
    set_error("pre_rotate_proj: interpolation order %d not supported (0 to 3; the spline "
              "orders 4 and 5 are not implemented)", order);
    return order < 0 ? TOMO_EINVAL : TOMO_EUNSUPPORTED;
  }
  TOMO_REQUIRE(order < 2 || work != nullptr,
               "pre_rotate_proj: orders 2 and 3 need the work buffer");
  TOMO_REQUIRE(order < 2 || (work != src && work != dst),
               "pre_rotate_proj: work must not alias src or dst");
  TOMO_REQUIRE(order < 2 || nx <= 0x7fffffff - 2 * kSplinePad,
               "pre_rotate_proj: sizes out of range");
  if (nproj == 0) return TOMO_OK;
  double c, s;
  cos_sin_deg(angle, &c, &s);
  const double cy = (double)(ny - 1) / 2, cx = (double)(nx - 1) / 2;
  const double oy = cy - (c * cy + s * cx), ox = cx - (-s * cy + c * cx);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)ceil_div(nx, 256), (unsigned)ny, (unsigned)nproj);
  if (order >= 2) {
    const int Py = (int)ny + 2 * kSplinePad, Px = (int)nx + 2 * kSplinePad;
    {
      ProfScope ps_("k_pre_spline", st);
      hipLaunchKernelGGL(k_pre_spline_cols, dim3((unsigned)ceil_div(Px, 256), (unsigned)nproj),
                         dim3(256), 0, st, src, work, (int)ny, (int)nx, order);
      hipLaunchKernelGGL(k_pre_spline_rows, dim3((unsigned)ceil_div(Py, 256), (unsigned)nproj),
                         dim3(256), 0, st, work, (int)ny, (int)nx, order);
    }
    TOMO_LAUNCH_CHECK();
    {
      ProfScope ps_("k_pre_rotate", st);
      if (order == 2)
        hipLaunchKernelGGL(k_pre_rotate_spline<2>, grid, dim3(256), 0, st, work, dst, (int)ny,
                           (int)nx, c, s, oy, ox);
      else
        hipLaunchKernelGGL(k_pre_rotate_spline<3>, grid, dim3(256), 0, st, work, dst, (int)ny,
                           (int)nx, c, s, oy, ox);
    }
    TOMO_LAUNCH_CHECK();
    return TOMO_OK;
  }
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
