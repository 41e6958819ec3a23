// FBP filter (tomocupy cfunc_filter + FBPFilter + fbp_filter_center) for gfx950.
//
// Two rows of the same slice (same filter) are filtered together as one
// complex row z = x1 + i*x2: FFT_ne(z) * W * IFFT_ne gives y1 + i*y2 because
// the full-length filter W (W[k] = w[k] for k <= ne/2 with the Nyquist bin's
// real part, W[ne-k] = conj(w[k])) is Hermitian, which is exactly the
// R2C -> mulw -> C2R of cfunc_filter::filter (src/cuda/cfunc_filter.cu:49-61).
// The edge-replicate padding of fbp_filter_center
// (src/tomocupy/reconstruction/backproj_functions.py:97-100) is applied while
// loading, so the (nz, nproj, ne) temporary of the reference never exists.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "fft_lds.h"
#include "internal.h"
#include "precision.h"

using namespace tomo;

namespace {

// w[z][k] = wfilter[k] * exp(-2 pi i (-center + sht[z] + n/2) k / ne), FP64
// phase (backproj_functions.py:103-104).
__global__ void k_flt_weights(const float *__restrict__ wfilter, const float *__restrict__ sht,
                              float2 *__restrict__ w, int ne, int n, int nz, double center) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int z = blockIdx.y;
  const int nh = ne / 2 + 1;
  if (k >= nh) return;
  const double shift = -center + (sht ? (double)sht[z] : 0.0) + (double)(n / 2);
  // reduce k*shift/ne modulo 1 before the trig for accuracy
  double f = (double)k * shift / (double)ne;
  f -= std::floor(f);
  double s, c;
  sincospi(-2.0 * f, &s, &c);
  const double a = (double)wfilter[k];
  w[(size_t)z * nh + k] = make_float2((float)(a * c), (float)(a * s));
}

// mode 0: rows of width ne already padded (cfunc_filter::filter);
// mode 1: rows of width n, edge-replicate padded on load and cropped on store
template <int MODE, typename T, int LV = 0>
__global__ __launch_bounds__(1024) void k_flt_rows(T *__restrict__ g, const float2 *__restrict__ w,
                                                   FftDesc fd, int ne, int n, int nproj) {
  extern __shared__ float2 sm[];
  const int pp = blockIdx.x, z = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int p0 = 2 * pp, p1 = p0 + 1;
  const bool has1 = p1 < nproj;
  const int width = MODE == 0 ? ne : n;
  T *r0 = g + ((size_t)z * nproj + p0) * width;
  T *r1 = g + ((size_t)z * nproj + (has1 ? p1 : p0)) * width;
  const int pad = ne / 2 - n / 2;
  stage_lds<8, cpx<T>>(
      ne, tid, nt,
      [&](int j) {
        const int src = MODE == 1 ? min(max(j - pad, 0), n - 1) : j;
        return cpx<T>{r0[src], has1 ? r1[src] : T(0)};
      },
      [&](int j, cpx<T> v) { sm[lds_pad(j)] = tof2(v); });
  __syncthreads();
  lds_fft<-1, kFftEPT, LV>(sm, fd, 0, tid, nt);
  const float2 *wz = w + (size_t)z * (ne / 2 + 1);
  const int nh = ne / 2;
  for (int k = tid; k < ne; k += nt) {
    float2 W;
    if (k == 0 || k == nh) {
      W = make_float2(wz[k].x, 0.f);
    } else if (k < nh) {
      W = wz[k];
    } else {
      W = cconj(wz[ne - k]);
    }
    sm[lds_pad(k)] = cmul(sm[lds_pad(k)], W);
  }
  __syncthreads();
  lds_fft<1, kFftEPT, LV>(sm, fd, 0, tid, nt);
  if (MODE == 0) {
    for (int j = tid; j < ne; j += nt) {
      const float2 y = sm[lds_pad(j)];
      str(r0 + j, y.x);
      if (has1) str(r1 + j, y.y);
    }
  } else {
    for (int j = tid; j < n; j += nt) {
      const float2 y = sm[lds_pad(pad + j)];
      str(r0 + j, y.x);
      if (has1) str(r1 + j, y.y);
    }
  }
}


// ---- fused path: the padded filter as a 2n circular convolution ------------
//
// With h = C2R(w_z) (length ne, real) the reference computes, for detector
// sample j in [0, n) of an edge-replicated row x,
//   y[j] = sum_{i<n} h[(j-i) mod ne] x[i] + x[0] HL[j] + x[n-1] HR[j],
//   HL[j] = sum_{d=j+1}^{j+pad} h[d],  HR[j] = sum_{d=pad+j+1}^{ne-n+j} h[d]
// (pad = ne/2 - n/2).  The first sum only sees lags |j-i| < n, so it is the
// circular convolution of length 2n of the zero-padded row with h2 (h wrapped
// to lags (-n, n), h2[n] = 0): FFT_2n, multiply by H2 = FFT_2n(h2)/(2n), and
// the unnormalised IFFT_2n.
// H2, HL and HR are built in FP64 and rounded once: the low-frequency values
// of H2 are tiny next to max|h| (the ramp vanishes at DC), so a float32
// round trip h -> H2 loses them (measured 4e-5 rel-L2 in the image at
// n = 4096).  With shift = -center + sht[z] + n/2 and phi_d = 2 pi (d - shift)/ne,
// the C2R of w[k] = wfilter[k] exp(-2 pi i k shift/ne) (imaginary parts of the
// DC and Nyquist bins dropped) is
//   h[d] = wf[0] + (-1)^d wf[ne/2] cos(pi shift) + 2 sum_{0<k<ne/2} wf[k] cos(k phi_d),
// summed with a rotating phasor.
// Each sum is split into kSeg contiguous segments, one thread each, whose
// rotating phasor is seeded from an exactly reduced phase; the segment sums
// are added in order.
constexpr int kSeg = 16;

__device__ __forceinline__ double flt_shift(const float *sht, int z, double center, int n) {
  return -center + (sht ? (double)sht[z] : 0.0) + (double)(n / 2);
}

__global__ __launch_bounds__(1024) void k_flt_h64(const float *__restrict__ wfilter,
                                                  const float *__restrict__ sht,
                                                  double *__restrict__ h, int ne, int n,
                                                  double center) {
  __shared__ double part[kSeg][64];
  const int lane = threadIdx.x & 63, s = threadIdx.x >> 6;
  const int d = blockIdx.x * 64 + lane, z = blockIdx.y;
  const int nh = ne / 2;
  const double shift = flt_shift(sht, z, center, n);
  const int len = (nh - 1 + kSeg - 1) / kSeg;
  const int k0 = 1 + s * len, k1 = min(nh, k0 + len);
  double acc = 0.0;
  if (d < ne && k0 < k1) {
    // phase k0 (d - shift)/ne mod 1, with k0 d reduced in integers
    double fs = (double)k0 * shift / (double)ne;
    fs -= std::floor(fs);
    double f = (double)(((long long)k0 * d) % ne) / (double)ne - fs;
    f -= std::floor(f);
    double rs, rc;
    sincospi(2.0 * f, &rs, &rc);
    double g = ((double)d - shift) / (double)ne;  // the step phi_d / (2 pi)
    g -= std::floor(g);
    double sn, cs;
    sincospi(2.0 * g, &sn, &cs);
    for (int k = k0; k < k1; ++k) {
      acc += (double)wfilter[k] * rc;
      const double t = rc * cs - rs * sn;
      rs = rc * sn + rs * cs;
      rc = t;
    }
  }
  part[s][lane] = acc;
  __syncthreads();
  if (s != 0 || d >= ne) return;
  double sum = 0.0;
#pragma unroll
  for (int q = 0; q < kSeg; ++q) sum += part[q][lane];
  double g = shift;  // cos(pi shift) with the argument reduced mod 2
  g -= 2.0 * std::floor(g / 2.0);
  const double cnyq = cospi(g);
  h[(size_t)z * ne + d] =
      (double)wfilter[0] + ((d & 1) ? -1.0 : 1.0) * (double)wfilter[nh] * cnyq + 2.0 * sum;
}

// FP64 prefix sums P[d] = sum_{e<d} h[e] (one workgroup per slice), then the
// edge responses HL[j] = P[pad+j+1] - P[j+1], HR[j] = P[j-n+ne+1] - P[pad+j+1].
__global__ __launch_bounds__(1024) void k_flt_edges64(const double *__restrict__ h,
                                                      double *__restrict__ P,
                                                      float *__restrict__ HL,
                                                      float *__restrict__ HR, int ne, int n) {
  __shared__ double csum[1024];
  const int z = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const double *hz = h + (size_t)z * ne;
  const int chunk = (ne + nt - 1) / nt;
  const int d0 = tid * chunk, d1 = min(ne, d0 + chunk);
  double acc = 0.0;
  for (int d = d0; d < d1; ++d) acc += hz[d];
  csum[tid] = acc;
  __syncthreads();
  if (tid == 0) {
    double run = 0.0;
    for (int t = 0; t < nt; ++t) {
      const double c = csum[t];
      csum[t] = run;
      run += c;
    }
  }
  __syncthreads();
  double *Pz = P + (size_t)z * (ne + 1);
  acc = csum[tid];
  for (int d = d0; d < d1; ++d) {
    Pz[d] = acc;
    acc += hz[d];
  }
  if (d1 == ne && d0 < d1) Pz[ne] = acc;
  __syncthreads();  // the global prefix sums are visible to the workgroup
  const int pad = ne / 2 - n / 2;
  for (int j = tid; j < n; j += nt) {
    HL[(size_t)z * n + j] = (float)(Pz[pad + j + 1] - Pz[j + 1]);
    HR[(size_t)z * n + j] = (float)(Pz[j - n + ne + 1] - Pz[pad + j + 1]);
  }
}

// H2[k] = FFT_2n(h2)[k] / (2n), h2 = h on lags (-n, n) wrapped to length 2n
// (h2[n] = 0); FP64 segment sums as above, rounded once.  Also, at
// H2 + nz 2n, the spectra for a differenced input: with d[j] = z[j] - z[j-1]
// (circular, length 2n), D[k] = Z[k] (1 - W^k), W = exp(-2 pi i/2n), so
// Z[k] H2[k] = D[k] H2[k] / (1 - W^k) for k > 0; slot 0 keeps H2[0] (the DC
// term is formed from the row sum).  A high-pass filter applied to a smooth
// row this way rounds relative to the small differences instead of the
// row's large low frequencies.
__global__ __launch_bounds__(1024) void k_flt_H2_64(const double *__restrict__ h,
                                                    float2 *__restrict__ H2, int ne, int n) {
  __shared__ double pre[kSeg][64], pim[kSeg][64];
  const int lane = threadIdx.x & 63, s = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + lane, z = blockIdx.y;
  const int n2 = 2 * n;
  const double *hz = h + (size_t)z * ne;
  const int len = (n2 + kSeg - 1) / kSeg;
  const int d0 = s * len, d1 = min(n2, d0 + len);
  double ar = 0.0, ai = 0.0;
  if (k < n2 && d0 < d1) {
    double ri, rr;
    sincospi(-2.0 * (double)(((long long)k * d0) % n2) / (double)n2, &ri, &rr);
    double sn, cs;
    sincospi(-2.0 * (double)k / (double)n2, &sn, &cs);
    for (int d = d0; d < d1; ++d) {
      const double v = d < n ? hz[d] : (d > n ? hz[ne - n2 + d] : 0.0);
      ar += v * rr;
      ai += v * ri;
      const double t = rr * cs - ri * sn;
      ri = rr * sn + ri * cs;
      rr = t;
    }
  }
  pre[s][lane] = ar;
  pim[s][lane] = ai;
  __syncthreads();
  if (s != 0 || k >= n2) return;
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int q = 0; q < kSeg; ++q) {
    sr += pre[q][lane];
    si += pim[q][lane];
  }
  sr /= n2;
  si /= n2;
  H2[(size_t)z * n2 + k] = make_float2((float)sr, (float)si);
  double gr = sr, gi = si;
  if (k > 0) {
    // 1 - W^k = 2 sin^2(pi k/2n) + i sin(pi k/n)
    const double sh = sinpi((double)k / (double)n2), s1 = sinpi((double)k / (double)n);
    const double dr = 2.0 * sh * sh, di = s1, den = dr * dr + di * di;
    gr = (sr * dr + si * di) / den;
    gi = (si * dr - sr * di) / den;
  }
  H2[((size_t)gridDim.y + z) * n2 + k] = make_float2((float)gr, (float)gi);
}

}  // namespace

int tomo::flt_prepare_2n(tomo_flt_plan *P, const float *wfilter, double center, const float *sht,
                         hipStream_t st) {
  TOMO_REQUIRE(P->fft2n.d.N == 2 * P->n && P->H2.p != nullptr,
               "fused filter path not available for n=%lld, ne=%lld", (long long)P->n,
               (long long)P->ne);
  const int ne = (int)P->ne, n = (int)P->n, nz = (int)P->nz;
  {
    ProfScope ps_("k_flt_prepare", st);
    hipLaunchKernelGGL(k_flt_h64, dim3((unsigned)ceil_div(ne, 64), nz), dim3(1024), 0, st,
                       wfilter, sht, P->Hd.as<double>(), ne, n, center);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_flt_prepare", st);
    hipLaunchKernelGGL(k_flt_edges64, dim3(nz), dim3(1024), 0, st, P->Hd.as<double>(),
                       P->P.as<double>(), P->HL.as<float>(), P->HR.as<float>(), ne, n);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_flt_prepare", st);
    hipLaunchKernelGGL(k_flt_H2_64, dim3((unsigned)ceil_div(2 * n, 64), nz), dim3(1024), 0, st,
                       P->Hd.as<double>(), P->H2.as<float2>(), ne, n);
  }
  TOMO_LAUNCH_CHECK();
  P->prepared = 1;
  return TOMO_OK;
}

extern "C" int flt_prepare(tomo_flt_plan *P, const float *wfilter, double center,
                           const float *sht, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "flt_prepare: plan is NULL");
  TOMO_REQUIRE(wfilter != nullptr, "flt_prepare: wfilter is NULL");
  TOMO_REQUIRE(std::isfinite(center), "flt_prepare: center is not finite");
  return flt_prepare_2n(P, wfilter, center, sht, (hipStream_t)stream);
}

namespace {

// ---- filter design (fbp_filter.py:60-118), host FP64 ----------------------

// Solve A X = B (n x n, n x m) by LU with partial pivoting; B overwritten by X.
bool lu_solve(std::vector<double> &A, std::vector<double> &B, int n, int m) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > best) best = std::fabs(A[i * n + k]), p = i;
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      for (int j = 0; j < m; ++j) std::swap(B[k * m + j], B[p * m + j]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      A[i * n + k] = f;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      for (int j = 0; j < m; ++j) B[i * m + j] -= f * B[k * m + j];
    }
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = 0; j < m; ++j) {
      double s = B[i * m + j];
      for (int q = i + 1; q < n; ++q) s -= A[i * n + q] * B[q * m + j];
      B[i * m + j] = s / A[i * n + i];
    }
  return true;
}

// _wint(n, t) (fbp_filter.py:90-118): local order-n quadrature weights for the
// integral of x*f(x), with the last 40 entries replaced by a linear ramp.
bool wint(int n, const std::vector<double> &t, std::vector<double> &wn) {
  const int N = (int)t.size();
  std::vector<double> s(n);
  // numpy.linspace(1e-40, 1, n): i*step + start, endpoint set exactly
  const double step = (1.0 - 1e-40) / (double)(n - 1);
  for (int i = 0; i < n; ++i) s[i] = (double)i * step + 1e-40;
  s[n - 1] = 1.0;
  // V[i][j] = exp(i*log(s_j)) = s_j^i ; u[i][j] = diff_j(exp((i+1) log s)/(i+1)), i < n+1
  std::vector<double> V(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i * n + j] = std::exp((double)i * std::log(s[j]));
  std::vector<double> U((n + 1) * (n - 1));
  for (int i = 0; i < n + 1; ++i) {
    std::vector<double> e(n);
    for (int j = 0; j < n; ++j) e[j] = std::exp((double)(i + 1) * std::log(s[j])) * (1.0 / (i + 1));
    for (int j = 0; j < n - 1; ++j) U[i * (n - 1) + j] = e[j + 1] - e[j];
  }
  // iv = inv(V) (LU with partial pivoting against the identity, as LAPACK
  // gesv), then W1 = iv @ U[1:n+1], W2 = iv @ U[0:n] (fbp_filter.py:97-101)
  std::vector<double> iv(n * n, 0.0);
  for (int i = 0; i < n; ++i) iv[i * n + i] = 1.0;
  std::vector<double> A = V;
  if (!lu_solve(A, iv, n, n)) return false;
  std::vector<double> W1(n * (n - 1), 0.0), W2(n * (n - 1), 0.0);
  for (int r = 0; r < n; ++r)
    for (int k = 0; k < n - 1; ++k) {
      double s1 = 0.0, s2 = 0.0;
      for (int q = 0; q < n; ++q) {
        s1 += iv[r * n + q] * U[(q + 1) * (n - 1) + k];
        s2 += iv[r * n + q] * U[q * (n - 1) + k];
      }
      W1[r * (n - 1) + k] = s1;
      W2[r * (n - 1) + k] = s2;
    }
  std::vector<double> p;
  for (int i = 1; i < n; ++i) p.push_back(1.0 / i);
  for (int i = 0; i < N - 2 * (n - 1) - 1; ++i) p.push_back(1.0 / (n - 1));
  for (int i = n - 1; i > 0; --i) p.push_back(1.0 / i);
  std::vector<double> w(N, 0.0);
  for (int j = 0; j < N - n + 1; ++j) {
    const double h = t[j + n - 1] - t[j];
    for (int k = 0; k < n - 1; ++k) {
      const double pk = p[j + k];
      for (int r = 0; r < n; ++r)
        w[j + r] += pk * (h * h * W1[r * (n - 1) + k] + h * t[j] * W2[r * (n - 1) + k]);
    }
  }
  wn = w;
  for (int i = N - 40; i < N; ++i) wn[i] = w[N - 40] / (N - 40) * i;
  return true;
}

}  // namespace


extern "C" int flt_calc_filter(int64_t ne, const char *name, double *out) {
  clear_error();
  TOMO_REQUIRE(name != nullptr && out != nullptr, "flt_calc_filter: NULL argument");
  TOMO_REQUIRE(ne >= 80 && ne % 2 == 0, "flt_calc_filter: ne=%lld must be even and >= 80",
               (long long)ne);
  const std::string nm(name);
  const int N = (int)(ne / 2 + 1);
  std::vector<double> t(N);
  for (int j = 0; j < N; ++j) t[j] = (double)j / (double)ne;
  const double d = 0.5;
  std::vector<double> wfa(N);
  const char *known[] = {"none", "ramp", "shepp", "cosine", "cosine2", "hamming", "hann", "parzen"};
  bool ok = false;
  for (const char *k : known) ok |= (nm == k);
  TOMO_REQUIRE(ok, "flt_calc_filter: unknown filter '%s'", name);
  if (nm == "none") {
    for (int j = 0; j < N; ++j) wfa[j] = (double)ne * 0.5;
    wfa[0] *= 2;
  } else {
    std::vector<double> wi;
    if (!wint(12, t, wi)) {
      set_error("flt_calc_filter: singular quadrature system");
      return TOMO_EINVAL;
    }
    for (int j = 0; j < N; ++j) {
      const double x = t[j];
      double win = 1.0;
      if (nm == "shepp") {
        const double a = x / (2 * d);
        win = (a == 0.0 ? 1.0 : std::sin(M_PI * a) / (M_PI * a)) * (x / d <= 2);
      } else if (nm == "cosine") {
        win = std::cos(M_PI * x / (2 * d)) * (x / d <= 1);
      } else if (nm == "cosine2") {
        const double c = std::cos(M_PI * x / (2 * d));
        win = c * c * (x / d <= 1);
      } else if (nm == "hamming") {
        win = (.54 + .46 * std::cos(M_PI * x / d)) * (x / d <= 1);
      } else if (nm == "hann") {
        win = (1 + std::cos(M_PI * x / d)) / 2.0 * (x / d <= 1);
      } else if (nm == "parzen") {
        win = std::pow(1 - x / d, 3) * (x / d <= 1);
      }
      wfa[j] = (double)ne * 0.5 * wi[j] * win;
    }
  }
  for (int j = 0; j < N; ++j) wfa[j] = 2 * wfa[j] * (wfa[j] >= 0 ? 1.0 : 0.0);
  wfa[0] *= 2;
  for (int j = 0; j < N; ++j) out[j] = wfa[j] / (double)ne;
  return TOMO_OK;
}

extern "C" int flt_create(tomo_flt_plan **plan, int64_t n, int64_t ne, int64_t nproj, int64_t nz,
                          int dtype) {
  clear_error();
  TOMO_REQUIRE(plan != nullptr, "flt_create: plan pointer is NULL");
  *plan = nullptr;
  TOMO_REQUIRE(n >= 20 && n % 2 == 0, "flt_create: n=%lld must be even and >= 20", (long long)n);
  TOMO_REQUIRE(ne >= n && ne % 2 == 0, "flt_create: ne=%lld must be even and >= n", (long long)ne);
  TOMO_REQUIRE(fft_size_ok(ne) && ne <= 16384,
               "flt_create: ne=%lld must be <= 16384 with prime factors <= %d", (long long)ne,
               kFftMaxPrime);
  TOMO_REQUIRE(nproj >= 1 && nz >= 1, "flt_create: nproj and nz must be >= 1");
  TOMO_REQUIRE(dtype == TOMO_F32 || dtype == TOMO_F16 || dtype == TOMO_BF16,
               "flt_create: unknown dtype %d", dtype);
  {
    // workgroup of the ne-point row kernels, from the radix sequence alone
    FftDesc d{};
    const int rc0 = FftPlan::radices((int)ne, d);
    if (rc0) return rc0;
    TOMO_REQUIRE(fft_threads(d, ne) > 0,
                 "flt_create: ne=%lld needs more than 1024 threads per FFT workgroup "
                 "(its radix-3/5/7/11/13 passes cover fewer than 16 values per thread)",
                 (long long)ne);
  }
  auto *P = new (std::nothrow) tomo_flt_plan();
  if (!P) {
    set_error("flt_create: host allocation failed");
    return TOMO_ENOMEM;
  }
  P->n = n;
  P->ne = ne;
  P->nproj = nproj;
  P->nz = nz;
  P->dtype = dtype;
  int rc = P->fft.init((int)ne);
  if (!rc) rc = P->w.alloc(sizeof(float2) * nz * (ne / 2 + 1));
  if (rc) {
    flt_destroy(P);
    return rc;
  }
  // fused 2n path (needs 2n <= 8192 and ne >= 2n so that lags |d| < n fit)
  if (2 * n <= 8192 && ne >= 2 * n && ne <= 16384 && !rc) {
    rc = P->fft2n.init((int)(2 * n));
    // H2, then the differenced-input spectra H2 / (1 - W^k) (see k_flt_H2_64)
    if (!rc) rc = P->H2.alloc(sizeof(float2) * 2 * nz * 2 * n);
    if (!rc) rc = P->HL.alloc(sizeof(float) * nz * n);
    if (!rc) rc = P->HR.alloc(sizeof(float) * nz * n);
    if (!rc) rc = P->P.alloc(sizeof(double) * nz * (ne + 1));
    if (!rc) rc = P->Hd.alloc(sizeof(double) * nz * ne);
    if (rc) {
      flt_destroy(P);
      return rc;
    }
  }
  P->threads = fft_threads(P->fft.d, ne);
  P->lds = sizeof(float2) * lds_padded_len((int)ne);
  if (P->threads <= 0 || P->lds > max_lds_per_block()) {
    set_error("flt_create: ne=%lld too large for one workgroup", (long long)ne);
    flt_destroy(P);
    return TOMO_EINVAL;
  }
  const void *rows_kernels[] = {
      (const void *)k_flt_rows<0, float>, (const void *)k_flt_rows<1, float>,
      (const void *)k_flt_rows<0, _Float16>, (const void *)k_flt_rows<1, _Float16>,
      (const void *)k_flt_rows<0, __bf16>, (const void *)k_flt_rows<1, __bf16>,
      (const void *)k_flt_rows<0, float, 1>, (const void *)k_flt_rows<1, float, 1>,
      (const void *)k_flt_rows<0, _Float16, 1>, (const void *)k_flt_rows<1, _Float16, 1>,
      (const void *)k_flt_rows<0, __bf16, 1>, (const void *)k_flt_rows<1, __bf16, 1>,
      (const void *)k_flt_rows<0, float, 2>, (const void *)k_flt_rows<1, float, 2>,
      (const void *)k_flt_rows<0, _Float16, 2>, (const void *)k_flt_rows<1, _Float16, 2>,
      (const void *)k_flt_rows<0, __bf16, 2>, (const void *)k_flt_rows<1, __bf16, 2>};
  hipError_t e = hipSuccess;
  for (const void *k : rows_kernels)
    if (e == hipSuccess) e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->lds);
  if (e != hipSuccess) {
    set_error("flt_create: hipFuncSetAttribute failed: %s", hipGetErrorString(e));
    flt_destroy(P);
    return TOMO_EHIP;
  }
  *plan = P;
  return TOMO_OK;
}

template <int MODE, typename T>
static void launch_rows(tomo_flt_plan *P, void *g, const float2 *w, hipStream_t st) {
  const dim3 grid((unsigned)ceil_div(P->nproj, 2), (unsigned)P->nz);
  const int lv = fft_level(P->ne);  // transforms with the level-1 or level-2 passes
  if (lv == 2)
    hipLaunchKernelGGL((k_flt_rows<MODE, T, 2>), grid, dim3(P->threads), P->lds, st, (T *)g, w,
                       P->fft.d, (int)P->ne, (int)P->n, (int)P->nproj);
  else if (lv == 1)
    hipLaunchKernelGGL((k_flt_rows<MODE, T, 1>), grid, dim3(P->threads), P->lds, st, (T *)g, w,
                       P->fft.d, (int)P->ne, (int)P->n, (int)P->nproj);
  else
    hipLaunchKernelGGL((k_flt_rows<MODE, T>), grid, dim3(P->threads), P->lds, st, (T *)g, w,
                       P->fft.d, (int)P->ne, (int)P->n, (int)P->nproj);
}

template <int MODE>
static int flt_rows(tomo_flt_plan *P, void *g, const float2 *w, hipStream_t st) {
  {
    ProfScope ps_("k_flt_rows", st);
    if (P->dtype == TOMO_F16)
      launch_rows<MODE, _Float16>(P, g, w, st);
    else if (P->dtype == TOMO_BF16)
      launch_rows<MODE, __bf16>(P, g, w, st);
    else
      launch_rows<MODE, float>(P, g, w, st);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

extern "C" int flt_filter(tomo_flt_plan *P, void *g, const float *w, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "flt_filter: plan is NULL");
  TOMO_REQUIRE(g != nullptr && w != nullptr, "flt_filter: NULL array");
  return flt_rows<0>(P, g, reinterpret_cast<const float2 *>(w), (hipStream_t)stream);
}

extern "C" int flt_filter_center(tomo_flt_plan *P, void *data, const float *wfilter, double center,
                                 const float *sht, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "flt_filter_center: plan is NULL");
  TOMO_REQUIRE(data != nullptr && wfilter != nullptr, "flt_filter_center: NULL array");
  TOMO_REQUIRE(std::isfinite(center), "flt_filter_center: center is not finite");
  hipStream_t st = (hipStream_t)stream;
  const int nh = (int)(P->ne / 2 + 1);
  {
    ProfScope ps_("k_flt_weights", st);
    hipLaunchKernelGGL(k_flt_weights, dim3((unsigned)ceil_div(nh, 256), (unsigned)P->nz), dim3(256), 0,
                       st, wfilter, sht, P->w.as<float2>(), (int)P->ne, (int)P->n, (int)P->nz, center);
  }
  TOMO_LAUNCH_CHECK();
  return flt_rows<1>(P, data, P->w.as<float2>(), st);
}

extern "C" void flt_destroy(tomo_flt_plan *P) {
  if (!P) return;
  P->fft.tw.release();
  P->fft2n.tw.release();
  P->w.release();
  P->H2.release();
  P->HL.release();
  P->HR.release();
  P->P.release();
  P->Hd.release();
  delete P;
}
