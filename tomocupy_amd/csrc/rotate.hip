// Projection rotation for gfx950 (tomocupy ProcFunctions.rotate_proj,
// src/tomocupy/processing/proc_functions.py:126-138: ndimage.rotate(data,
// angle, reshape=False, order=order, axes=(2, 1), mode='nearest'), applied by
// proc_proj, :182-184): every projection (rows y, columns x) is resampled at
//   (iy, ix) = R (y, x) + offset,  R = [[c, s], [-s, c]],
//   offset = centre - R centre,  centre = ((ny - 1)/2, (nx - 1)/2),
// c, s = cos, sin of the angle in degrees, with order 0 (nearest sample),
// 1 (bilinear), neighbours clamped to the image ('nearest' extension), or 2
// to 5 (B-splines on the prefiltered, edge-padded plane, as SciPy's spline
// path).  Coordinates in FP64, interpolation in float32.
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

// B-spline prefilter of order 2 to 5 on the plane edge-padded by kSplinePad
// samples (as SciPy's _prepad_for_spline_filter does for mode 'nearest'):
// the inverse of the sampled B-spline, whose poles z in (-1, 0) are one for
// orders 2, 3 and two for 4, 5 (Unser, Aldroubi and Eden, IEEE TSP 41, 1993).
// Its impulse response is, per pole, g[k] = (1 - z)/(1 + z) z^|k| (the
// causal/anticausal pair with gain (1 - z)(1 - 1/z)), and for two poles the
// convolution of the two.  It is applied as a symmetric FIR truncated where
// |z_max|^M < 1e-8, separably along rows then columns of an LDS tile, with
// whole-sample mirror extension at the padded plane's edges: the solution the
// recursive filter with mirror initial values (Thevenaz, Blu and Unser, IEEE
// TMI 19, 2000) computes.  The padding keeps the boundary |z|^12 from the data.
constexpr int kSplinePad = 12;
constexpr int kSplineTy = 32, kSplineTx = 64;  // output tile of the prefilter
constexpr double kSplinePoles[4][2] = {{-0.17157287525380990, 0.0},   // order 2: sqrt(8) - 3
                                       {-0.26794919243112270, 0.0},   // order 3: sqrt(3) - 2
                                       {-0.36134122590022033, -0.013725429297339118},
                                       {-0.43057534709997360, -0.043096288203264665}};
// FIR half-lengths ceil(log(1e-8) / log|z_max|) for orders 2..5
constexpr int kSplineHalf[4] = {11, 14, 19, 22};

struct SplineTaps {
  float h[24];  // h[0..M]
};

// h[k] of the prefilter of `order` for k = 0..M (FP64 on the host)
SplineTaps spline_taps(int order) {
  const double *z = kSplinePoles[order - 2];
  const int M = kSplineHalf[order - 2];
  auto g = [](double zz, int k) { return (1 - zz) / (1 + zz) * std::pow(zz, std::abs(k)); };
  SplineTaps t{};
  for (int k = 0; k <= M; ++k) {
    double v;
    if (z[1] == 0.0) {
      v = g(z[0], k);
    } else {
      v = 0.0;  // |z| < 0.44: 0.44^120 vanishes in FP64
      for (int j = -120; j <= 120; ++j) v += g(z[0], j) * g(z[1], k - j);
    }
    t.h[k] = (float)v;
  }
  return t;
}

// whole-sample mirror into [0, N), then clamped (indices that only feed
// outputs outside the plane may reflect more than once)
__device__ __forceinline__ int spline_mirror(int q, int N) {
  q = q < 0 ? -q : q;
  q = q > N - 1 ? 2 * (N - 1) - q : q;
  return min(max(q, 0), N - 1);
}

// spline coefficients of one kSplineTy x kSplineTx tile of the padded plane
// of projection blockIdx.z: the tile and an M halo (mirrored, then edge
// padding read from the projection) into LDS, the row FIR into a second LDS
// array, the column FIR out to work[p][Py][Px]
template <int M>
__global__ __launch_bounds__(256) void k_pre_spline_coef(const float *__restrict__ src,
                                                         float *__restrict__ work, int ny,
                                                         int nx, SplineTaps taps) {
  constexpr int AH = kSplineTy + 2 * M, AW = kSplineTx + 2 * M;
  __shared__ float A[AH * AW];
  __shared__ float B[AH * kSplineTx];
  const int Py = ny + 2 * kSplinePad, Px = nx + 2 * kSplinePad;
  const int X0 = blockIdx.x * kSplineTx, Y0 = blockIdx.y * kSplineTy, p = blockIdx.z;
  const float *img = src + (size_t)p * ny * nx;
  for (int i = threadIdx.x; i < AH * AW; i += 256) {
    const int r = i / AW, cc = i - r * AW;
    const int yy = min(max(spline_mirror(Y0 - M + r, Py) - kSplinePad, 0), ny - 1);
    const int xx = min(max(spline_mirror(X0 - M + cc, Px) - kSplinePad, 0), nx - 1);
    A[i] = img[(size_t)yy * nx + xx];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < AH * kSplineTx; i += 256) {
    const int r = i / kSplineTx, j = i % kSplineTx;
    const float *a = A + r * AW + j + M;
    float v = taps.h[0] * a[0];
#pragma unroll
    for (int k = 1; k <= M; ++k) v += taps.h[k] * (a[k] + a[-k]);
    B[i] = v;
  }
  __syncthreads();
  float *w = work + (size_t)p * Py * Px;
  for (int i = threadIdx.x; i < kSplineTy * kSplineTx; i += 256) {
    const int r = i / kSplineTx, j = i % kSplineTx;
    const int y = Y0 + r, x = X0 + j;
    if (y >= Py || x >= Px) continue;
    const float *bb = B + (r + M) * kSplineTx + j;
    float v = taps.h[0] * bb[0];
#pragma unroll
    for (int k = 1; k <= M; ++k) v += taps.h[k] * (bb[k * kSplineTx] + bb[-k * kSplineTx]);
    w[(size_t)y * Px + x] = v;
  }
}

// the ORDER + 1 weights beta_ORDER(u - b - k), k = 0..ORDER, of the samples
// b, b + 1, ... around u (closed forms for 2 and 3; for 4 and 5 the
// truncated-power sum beta_n(x) = 1/n! sum_k (-1)^k C(n+1, k)
// max(x + (n+1)/2 - k, 0)^n, in FP64)
template <int ORDER>
__device__ __forceinline__ void spline_weights(double u, float *w, double &b) {
  if (ORDER == 3) {
    const double f = floor(u);
    const float t = (float)(u - f);
    b = f - 1.0;
    const float t2 = t * t, t3 = t2 * t, m = 1.f - t;
    w[0] = m * m * m / 6.f;
    w[1] = (4.f - 6.f * t2 + 3.f * t3) / 6.f;
    w[2] = (1.f + 3.f * t + 3.f * t2 - 3.f * t3) / 6.f;
    w[3] = t3 / 6.f;
  } else if (ORDER == 2) {
    const double f = floor(u + 0.5);
    const float t = (float)(u - f);
    b = f - 1.0;
    w[0] = 0.5f * (0.5f - t) * (0.5f - t);
    w[1] = 0.75f - t * t;
    w[2] = 0.5f * (0.5f + t) * (0.5f + t);
  } else {
    b = (ORDER & 1) ? floor(u) - (ORDER - 1) / 2 : floor(u + 0.5) - ORDER / 2;
    constexpr double binom[7] = {1, ORDER + 1.0, (ORDER + 1.0) * ORDER / 2,
                                 (ORDER + 1.0) * ORDER * (ORDER - 1) / 6,
                                 (ORDER + 1.0) * ORDER * (ORDER - 1) * (ORDER - 2) / 24,
                                 (ORDER + 1.0) * ORDER * (ORDER - 1) * (ORDER - 2) * (ORDER - 3) / 120,
                                 1};
    constexpr double inv_fact = ORDER == 4 ? 1.0 / 24 : 1.0 / 120;
#pragma unroll
    for (int k = 0; k <= ORDER; ++k) {
      const double x = u - (b + k) + 0.5 * (ORDER + 1);
      double sum = 0.0;
#pragma unroll
      for (int j = 0; j <= ORDER + 1; ++j) {
        const double t = fmax(x - j, 0.0);
        double tp = t;
#pragma unroll
        for (int e = 1; e < ORDER; ++e) tp *= t;
        sum += ((j & 1) ? -binom[j] : binom[j]) * tp;
      }
      w[k] = (float)(sum * inv_fact);
    }
  }
}

// B-spline evaluation of order 2 to 5 at the rotated coordinates, on the
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
  spline_weights<ORDER>(iy, wy, by);
  spline_weights<ORDER>(ix, wx, bx);
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
  TOMO_REQUIRE(order >= 0 && order <= 5,
               "pre_rotate_proj: interpolation order must be in 0..5");
  TOMO_REQUIRE(order < 2 || work != nullptr,
               "pre_rotate_proj: the spline orders 2 to 5 need the work buffer");
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
      ProfScope ps_("k_pre_spline_coef", st);
      const dim3 g2((unsigned)ceil_div(Px, kSplineTx), (unsigned)ceil_div(Py, kSplineTy),
                    (unsigned)nproj);
      const SplineTaps taps = spline_taps(order);
      auto kern = order == 2   ? k_pre_spline_coef<kSplineHalf[0]>
                  : order == 3 ? k_pre_spline_coef<kSplineHalf[1]>
                  : order == 4 ? k_pre_spline_coef<kSplineHalf[2]>
                               : k_pre_spline_coef<kSplineHalf[3]>;
      hipLaunchKernelGGL(kern, g2, dim3(256), 0, st, src, work, (int)ny, (int)nx, taps);
    }
    TOMO_LAUNCH_CHECK();
    {
      ProfScope ps_("k_pre_rotate", st);
      auto kern = order == 2   ? k_pre_rotate_spline<2>
                  : order == 3 ? k_pre_rotate_spline<3>
                  : order == 4 ? k_pre_rotate_spline<4>
                               : k_pre_rotate_spline<5>;
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, work, dst, (int)ny, (int)nx, c, s, oy,
                         ox);
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
