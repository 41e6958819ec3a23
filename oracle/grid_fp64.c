/* FP64 gridding of the USFFT backprojector and its transpose, in C.
 *
 * TEST INFRASTRUCTURE ONLY: the checker's fast path for the gather step of
 * oracle/fourierrec_fp64.py (same formulas, same FP64 arithmetic, summation
 * in sample order), never part of the product.  It exists because the NumPy
 * scatter (np.bincount over 81 taps per sample) needs ~80 s for one slice
 * pair at n=4096 with 3000 angles.
 *
 * Restates, per sample (p, k) of complex slice c:
 *   takexy (src/include/kernels_fourierrec.cuh:23-36):
 *     x = (k - n/2)/n cos(theta_p), y = -(k - n/2)/n sin(theta_p),
 *     values >= 0.5 clamped to 0.5 - 1e-5;
 *   gather (kernels_fourierrec.cuh:60-119):
 *     b0 = floor(2n x) - m, b1 = floor(2n y) - m,
 *     e0[i] = exp(c1 ((b0+i)/(2n) - x)^2), e1 likewise with y, c1 = -pi^2/mu,
 *     fde[c, n+m+b1+i1, n+m+b0+i0] += c0 e1[i1] e0[i0] * (4/n) G[c,p,k],
 *     c0 = pi/(mu 4 n^2), on the (2n+2m)^2 grid with its m-wide border.
 * grid_interp is the exact transpose (interpolation of the grid at the
 * samples with the same weights), the gather^T of the forward projector.
 *
 * Threads split the grid by rows (each thread applies only the taps that
 * land in its rows, in sample order), so the result does not depend on the
 * thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void sample_geom(double th, int64_t k, int64_t n, int m, double mu, int64_t *b0,
                        int64_t *b1, double *e0, double *e1) {
  const double c1 = -M_PI * M_PI / mu;
  const double t = (double)(k - n / 2) / (double)n;
  double x = t * cos(th), y = -t * sin(th);
  if (x >= 0.5) x = 0.5 - 1e-5;
  if (y >= 0.5) y = 0.5 - 1e-5;
  *b0 = (int64_t)floor(2.0 * (double)n * x) - m;
  *b1 = (int64_t)floor(2.0 * (double)n * y) - m;
  for (int i = 0; i <= 2 * m; ++i) {
    const double w0 = (double)(*b0 + i) / (2.0 * (double)n) - x;
    const double w1 = (double)(*b1 + i) / (2.0 * (double)n) - y;
    e0[i] = exp(c1 * w0 * w0);
    e1[i] = exp(c1 * w1 * w1);
  }
}

/* G: (nzc, nproj, n) complex as interleaved doubles; theta: (nproj,) in FP64
 * (the float32 angles widened); fde: (nzc, S, S) complex, S = 2n+2m, zeroed
 * by the caller and accumulated into. */
void grid_scatter(const double *G, const double *theta, int64_t nzc, int64_t nproj, int64_t n,
                  int m, double mu, double *fde) {
  const int64_t S = 2 * n + 2 * m;
  const double c0 = M_PI / (mu * (4.0 * (double)n * (double)n));
  const double sc = 4.0 / (double)n;
#pragma omp parallel
  {
    int nth = 1, me = 0;
#ifdef _OPENMP
    nth = omp_get_num_threads();
    me = omp_get_thread_num();
#endif
    const int64_t r0 = S * me / nth, r1 = S * (me + 1) / nth;
    double e0[16], e1[16];
    for (int64_t p = 0; p < nproj; ++p) {
      for (int64_t k = 0; k < n; ++k) {
        int64_t b0, b1;
        sample_geom(theta[p], k, n, m, mu, &b0, &b1, e0, e1);
        const int64_t rb = n + m + b1;
        if (rb + 2 * m < r0 || rb >= r1) continue;
        for (int64_t c = 0; c < nzc; ++c) {
          const double *g = G + 2 * ((c * nproj + p) * n + k);
          const double vr = g[0] * sc, vi = g[1] * sc;
          double *F = fde + 2 * c * S * S;
          for (int i1 = 0; i1 <= 2 * m; ++i1) {
            const int64_t r = rb + i1;
            if (r < r0 || r >= r1) continue;
            const double we = c0 * e1[i1];
            double *row = F + 2 * (r * S + n + m + b0);
            for (int i0 = 0; i0 <= 2 * m; ++i0) {
              const double w = we * e0[i0];
              row[2 * i0] += w * vr;
              row[2 * i0 + 1] += w * vi;
            }
          }
        }
      }
    }
  }
}

/* Transpose of grid_scatter without its 4/n factor (the caller applies it):
 * G[c,p,k] = sum_taps c0 e1 e0 fde[c, ...]. */
void grid_interp(const double *fde, const double *theta, int64_t nzc, int64_t nproj, int64_t n,
                 int m, double mu, double *G) {
  const int64_t S = 2 * n + 2 * m;
  const double c0 = M_PI / (mu * (4.0 * (double)n * (double)n));
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < nproj; ++p) {
    double e0[16], e1[16];
    for (int64_t k = 0; k < n; ++k) {
      int64_t b0, b1;
      sample_geom(theta[p], k, n, m, mu, &b0, &b1, e0, e1);
      for (int64_t c = 0; c < nzc; ++c) {
        const double *F = fde + 2 * c * S * S;
        double ar = 0.0, ai = 0.0;
        for (int i1 = 0; i1 <= 2 * m; ++i1) {
          const double we = c0 * e1[i1];
          const double *row = F + 2 * ((n + m + b1 + i1) * S + n + m + b0);
          for (int i0 = 0; i0 <= 2 * m; ++i0) {
            const double w = we * e0[i0];
            ar += w * row[2 * i0];
            ai += w * row[2 * i0 + 1];
          }
        }
        double *g = G + 2 * ((c * nproj + p) * n + k);
        g[0] = ar;
        g[1] = ai;
      }
    }
  }
}
