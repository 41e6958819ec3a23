// Single-step phase retrieval for gfx950 (tomocupy processing/retrieve_phase.py:
// paganin_filter :59-104, farago_filter :106-141, _retrieve_phase :143-161,
// applied by ProcFunctions.proc_proj, proc_functions.py:169-181) on float32
// projections (nproj, dy, dz) in place:
//   prj  = projection edge-replicated to Ny x Nx (pads py rows and pz columns
//          on each side, Ny = dy + 2 py and Nx = dz + 2 pz powers of two)
//   proj = Re ifft2(fft2(prj) * H / max H)[py : py + dy, pz : pz + dz]
// with H = fftshift of the method's factor on the reciprocal grid:
//   paganin   1 / (lambda dist w2 / (4 pi) + alpha),          w2 = u^2 + v^2
//   Gpaganin  1 / (1 - 2 aph / W^2 (kf - 2)),  aph = db dist lambda / (4 pi),
//             kf = cos(2 pi ps u) + cos(2 pi ps v)
//   farago    1 / (cos(pi lambda dist w2) + db sin(pi lambda dist w2))
// u, v the reference's reciprocal coordinates (k - (N-1)/2) / ((N-1) ps),
// half a sample off zero for even N.
//
// Three kernels per batch of projections, through a complex work buffer of
// the dy real rows' spectra (the py padded rows above and below a projection
// are copies of its first and last rows, so their row spectra are those
// rows' spectra): rows (two real rows packed into one complex FFT, split into
// their two spectra), columns (the padded column rebuilt from the edge
// spectra, forward FFT, filter, inverse FFT scaled by 1/(Ny Nx), the dy rows
// the crop keeps written back), and inverse rows (the Hermitian parts of two
// row spectra packed into one inverse FFT, whose real and imaginary parts
// are the two real rows).  The
// FFTs are the LDS Stockham transforms of fft_lds.h; the filter is evaluated
// in float32 (the reference's precision) from 1-D tables built in FP64.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.h"
#include "fft_lds.h"

using namespace tomo;

struct tomo_ph_plan {
  int64_t dy = 0, dz = 0, batch = 0;
  int method = 0;
  int py = 0, pz = 0, Ny = 0, Nx = 0;
  int lgC = 0;  // log2 of the columns per workgroup of the column pass
  float A = 0.f, B = 0.f, alpha = 0.f, db = 0.f, hmax = 1.f;
  int rowThreads = 0, colThreads = 0;
  size_t rowLds = 0, colLds = 0;
  FftPlan fy, fx;
  DevBuf ty, tx;  // fftshifted 1-D terms of the filter (rows, columns)
  DevBuf F;       // (batch, dy, Nx) float2: spectra of the real rows
};

namespace {

constexpr double kPi = 3.14159265359;               // retrieve_phase.py:48
constexpr double kPlanck = 6.58211928e-19;          // [keV s], :49
constexpr double kLight = 299792458e+2;             // [cm/s], :47

// the method's factor at one grid point, from the sum of the two 1-D terms
__device__ __forceinline__ float ph_factor(int method, float s, float A, float B, float alpha,
                                           float db) {
  if (method == 0) return 1.f / (A * s + alpha);             // paganin
  if (method == 1) return 1.f / (1.f - B * (s - 2.f));       // Gpaganin
  float sn, cs;                                              // farago
  sincosf(A * s, &sn, &cs);
  return 1.f / (cs + db * sn);
}

// row maxima of the factor over the grid (the normalisation max H)
__global__ __launch_bounds__(256) void k_ph_fmax(const float *__restrict__ ty,
                                                 const float *__restrict__ tx, int Nx,
                                                 int method, float A, float B, float alpha,
                                                 float db, float *__restrict__ rowmax) {
  __shared__ float red[256];
  const int ky = blockIdx.x;
  float m = -INFINITY;
  for (int kx = threadIdx.x; kx < Nx; kx += 256)
    m = fmaxf(m, ph_factor(method, ty[ky] + tx[kx], A, B, alpha, db));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) rowmax[ky] = red[0];
}

// Forward FFTs of the column-padded real rows 2 pr and 2 pr + 1 of
// projection blockIdx.y: z = row0 + i row1, Z = fft(z);
// X0 = (Z[k] + conj Z[-k]) / 2, X1 = (Z[k] - conj Z[-k]) / (2i).
__global__ __launch_bounds__(512) void k_ph_rows_fwd(const float *__restrict__ data,
                                                      float2 *__restrict__ F, FftDesc fd, int dy,
                                                      int dz, int pz) {
  extern __shared__ float2 sm[];
  const int pr = blockIdx.x, b = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int Nx = fd.N;
  const int iy0 = 2 * pr;
  const float *img = data + (size_t)b * dy * dz;
  const float *r0 = img + (size_t)iy0 * dz;
  const float *r1 = r0 + dz;
  for (int ix = tid; ix < Nx; ix += nt) {
    const int sx = min(max(ix - pz, 0), dz - 1);
    sm[lds_pad(ix)] = make_float2(r0[sx], r1[sx]);
  }
  __syncthreads();
  lds_fft<-1>(sm, fd, 0, tid, nt);
  float2 *f0 = F + ((size_t)b * dy + iy0) * Nx;
  float2 *f1 = f0 + Nx;
  for (int k = tid; k < Nx; k += nt) {
    const float2 zk = sm[lds_pad(k)], zm = cconj(sm[lds_pad(k == 0 ? 0 : Nx - k)]);
    f0[k] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y));
    f1[k] = make_float2(0.5f * (zk.y - zm.y), -0.5f * (zk.x - zm.x));
  }
}

// Column pass over 2^lgC adjacent columns of projection blockIdx.y: the
// padded column (padded row ky holds real row clamp(ky - py, 0, dy - 1)),
// forward FFT, times H / max H / (Ny Nx), inverse FFT; rows py .. py+dy-1
// written back in place.
__global__ __launch_bounds__(512) void k_ph_cols(float2 *__restrict__ F, FftDesc fd,
                                                  int lgC, int Nx, int dy, int py,
                                                  const float *__restrict__ ty,
                                                  const float *__restrict__ tx, int method,
                                                  float A, float B, float alpha, float db,
                                                  float hmax, float scale) {
  extern __shared__ float2 sm[];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int Ny = fd.N, C = 1 << lgC;
  // workgroups are dealt to the 8 XCDs in turn: give each XCD a contiguous
  // range of column groups, so that the groups sharing cache lines (2^lgC
  // columns = 8 << lgC bytes of each row) share its L2
  const int G = gridDim.x;
  const int g = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  const int kx0 = g << lgC, b = blockIdx.y;
  float2 *fb = F + (size_t)b * dy * Nx + kx0;
  for (int e = tid; e < (Ny << lgC); e += nt) {
    const int ky = e >> lgC, c = e & (C - 1);
    sm[(lds_pad(ky) << lgC) + c] = fb[(size_t)min(max(ky - py, 0), dy - 1) * Nx + c];
  }
  __syncthreads();
  lds_fft<-1>(sm, fd, lgC, tid, nt);
  for (int e = tid; e < (Ny << lgC); e += nt) {
    const int ky = e >> lgC, c = e & (C - 1);
    const float h = ph_factor(method, ty[ky] + tx[kx0 + c], A, B, alpha, db) / hmax * scale;
    float2 &v = sm[(lds_pad(ky) << lgC) + c];
    v = make_float2(v.x * h, v.y * h);
  }
  __syncthreads();
  lds_fft<1>(sm, fd, lgC, tid, nt);
  for (int e = tid; e < (dy << lgC); e += nt) {
    const int y = e >> lgC, c = e & (C - 1);
    fb[(size_t)y * Nx + c] = sm[(lds_pad(y + py) << lgC) + c];
  }
}

// Inverse row FFTs for output rows 2 pr and 2 pr + 1:
// Re ifft(G) = ifft(Herm G), Herm G[k] = (G[k] + conj G[-k]) / 2, so
// ifft(Herm G0 + i Herm G1) = Re ifft(G0) + i Re ifft(G1); cropped columns
// written back to the projection.
__global__ __launch_bounds__(512) void k_ph_rows_inv(float *__restrict__ data,
                                                      const float2 *__restrict__ F, FftDesc fd,
                                                      int dy, int dz, int pz) {
  extern __shared__ float2 sm[];
  const int pr = blockIdx.x, b = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int Nx = fd.N;
  const int oy0 = 2 * pr, oy1 = oy0 + 1;  // dy is even: both rows exist
  const float2 *g0 = F + ((size_t)b * dy + oy0) * Nx;
  const float2 *g1 = g0 + Nx;
  for (int k = tid; k < Nx; k += nt) {
    const int km = k == 0 ? 0 : Nx - k;
    const float2 a = g0[k], am = g0[km], c = g1[k], cm = g1[km];
    const float2 h0 = make_float2(0.5f * (a.x + am.x), 0.5f * (a.y - am.y));
    const float2 h1 = make_float2(0.5f * (c.x + cm.x), 0.5f * (c.y - cm.y));
    sm[lds_pad(k)] = make_float2(h0.x - h1.y, h0.y + h1.x);
  }
  __syncthreads();
  lds_fft<1>(sm, fd, 0, tid, nt);
  float *img = data + (size_t)b * dy * dz;
  for (int x = tid; x < dz; x += nt) {
    const float2 v = sm[lds_pad(pz + x)];
    img[(size_t)oy0 * dz + x] = v.x;
    if (oy1 < dy) img[(size_t)oy1 * dz + x] = v.y;
  }
}

// the reference's _calc_pad_width (retrieve_phase.py:214-216)
int pad_width(int64_t dim, double pad_pix) {
  const double p2 = std::pow(2.0, std::ceil(std::log2((double)dim + pad_pix)));
  return (int)((p2 - (double)dim) * 0.5);
}

bool pow2(int v) { return v >= 2 && (v & (v - 1)) == 0; }

}  // namespace

extern "C" void ph_destroy(tomo_ph_plan *P) {
  if (!P) return;
  P->fy.tw.release();
  P->fx.tw.release();
  P->ty.release();
  P->tx.release();
  P->F.release();
  delete P;
}

extern "C" int ph_create(tomo_ph_plan **plan, int64_t dy, int64_t dz, int64_t batch, int method,
                         double pixel_size, double dist, double energy, double alpha, double db,
                         double W) {
  clear_error();
  TOMO_REQUIRE(plan != nullptr, "ph_create: plan pointer is NULL");
  *plan = nullptr;
  TOMO_REQUIRE(dy >= 2 && dz >= 2 && dy % 2 == 0 && dz % 2 == 0,
               "ph_create: projection size %lld x %lld must be even (powers of two after "
               "padding)", (long long)dy, (long long)dz);
  TOMO_REQUIRE(batch >= 1 && batch <= 65535, "ph_create: batch out of range");
  TOMO_REQUIRE(method >= 0 && method <= 2, "ph_create: method %d (0 paganin, 1 Gpaganin, 2 farago)",
               method);
  TOMO_REQUIRE(pixel_size > 0 && dist > 0 && energy > 0 && std::isfinite(pixel_size) &&
                   std::isfinite(dist) && std::isfinite(energy),
               "ph_create: pixel size, distance and energy must be > 0");
  TOMO_REQUIRE(std::isfinite(alpha) && std::isfinite(db) && std::isfinite(W) &&
                   (method != 1 || W > 0),
               "ph_create: invalid alpha, delta/beta or W");
  auto *P = new (std::nothrow) tomo_ph_plan();
  if (!P) {
    set_error("ph_create: host allocation failed");
    return TOMO_ENOMEM;
  }
  P->dy = dy;
  P->dz = dz;
  P->batch = batch;
  P->method = method;
  const double lambda = 2 * kPi * kPlanck * kLight / energy;  // _wavelength, :52-53
  const double pad_pix = std::ceil(kPi * lambda * dist / (pixel_size * pixel_size));
  P->py = pad_width(dy, pad_pix);
  P->pz = pad_width(dz, pad_pix);
  P->Ny = (int)(dy + 2 * P->py);
  P->Nx = (int)(dz + 2 * P->pz);
  if (!pow2(P->Ny) || !pow2(P->Nx) || P->Ny > 8192 || P->Nx > 8192) {
    set_error("ph_create: padded size %d x %d not a power of two <= 8192", P->Ny, P->Nx);
    ph_destroy(P);
    return TOMO_EINVAL;
  }
  if (method == 0) {
    P->A = (float)(lambda * dist / (4 * kPi));
  } else if (method == 1) {
    const double aph = db * (dist * lambda) / (4 * kPi);
    P->B = (float)(2 * aph / (W * W));
  } else {
    P->A = (float)(kPi * lambda * dist);
  }
  P->alpha = (float)alpha;
  P->db = (float)db;
  // fftshifted 1-D terms: u_k = (2k - (N-1)) * 0.5 / ((N-1) ps) at index
  // (k - N/2) mod N (_reciprocal_coord :296-304, fftshift :89-95)
  auto terms = [&](int N) {
    std::vector<float> t(N);
    for (int k = 0; k < N; ++k) {
      const int j = ((k - N / 2) % N + N) % N;
      const double u = (2.0 * j - (N - 1)) * 0.5 / ((N - 1) * pixel_size);
      t[k] = (float)(method == 1 ? std::cos(u * 2 * kPi * pixel_size) : u * u);
    }
    return t;
  };
  const std::vector<float> hy = terms(P->Ny), hx = terms(P->Nx);
  int rc = P->fy.init(P->Ny);
  if (!rc) rc = P->fx.init(P->Nx);
  if (!rc) rc = P->ty.alloc(sizeof(float) * P->Ny);
  if (!rc) rc = P->tx.alloc(sizeof(float) * P->Nx);
  if (!rc) rc = P->F.alloc(sizeof(float2) * (size_t)batch * dy * P->Nx);
  if (rc) {
    ph_destroy(P);
    return rc;
  }
  TOMO_HIP(hipMemcpy(P->ty.p, hy.data(), sizeof(float) * P->Ny, hipMemcpyHostToDevice));
  TOMO_HIP(hipMemcpy(P->tx.p, hx.data(), sizeof(float) * P->Nx, hipMemcpyHostToDevice));
  // max H over the grid, with the kernels' float32 factor
  {
    DevBuf rowmax;
    rc = rowmax.alloc(sizeof(float) * P->Ny);
    if (rc) {
      ph_destroy(P);
      return rc;
    }
    hipLaunchKernelGGL(k_ph_fmax, dim3((unsigned)P->Ny), dim3(256), 0, 0, P->ty.as<float>(),
                       P->tx.as<float>(), P->Nx, method, P->A, P->B, P->alpha, P->db,
                       rowmax.as<float>());
    std::vector<float> rm(P->Ny);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpy(rm.data(), rowmax.p, sizeof(float) * P->Ny, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_error("ph_create: %s", hipGetErrorString(e));
      ph_destroy(P);
      return TOMO_EHIP;
    }
    P->hmax = *std::max_element(rm.begin(), rm.end());
    if (!(std::isfinite(P->hmax) && P->hmax != 0.f)) {
      set_error("ph_create: the filter maximum is %g (check the parameters)", (double)P->hmax);
      ph_destroy(P);
      return TOMO_EINVAL;
    }
  }
  // launch shapes: rows one transform per workgroup; columns 2^lgC
  // interleaved transforms, at most 8192 elements (64 KB of LDS)
  P->lgC = 0;
  while (P->lgC < 3 && ((int64_t)P->Ny << (P->lgC + 1)) <= 8192) ++P->lgC;
  P->rowThreads = fft_threads(P->fx.d, P->Nx);
  P->colThreads = fft_threads(P->fy.d, (int64_t)P->Ny << P->lgC);
  P->rowLds = sizeof(float2) * lds_padded_len(P->Nx);
  P->colLds = sizeof(float2) * ((size_t)lds_padded_len(P->Ny) << P->lgC);
  if (P->rowThreads <= 0 || P->rowThreads > 512 || P->colThreads <= 0 ||
      P->colThreads > 512 || P->rowLds > max_lds_per_block() ||
      P->colLds > max_lds_per_block()) {
    set_error("ph_create: padded size %d x %d too large for one workgroup", P->Ny, P->Nx);
    ph_destroy(P);
    return TOMO_EINVAL;
  }
  hipError_t e = hipFuncSetAttribute((const void *)k_ph_rows_fwd,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->rowLds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_ph_rows_inv,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->rowLds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_ph_cols, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)P->colLds);
  if (e != hipSuccess) {
    set_error("ph_create: hipFuncSetAttribute failed: %s", hipGetErrorString(e));
    ph_destroy(P);
    return TOMO_EHIP;
  }
  *plan = P;
  return TOMO_OK;
}

extern "C" int ph_get_pads(const tomo_ph_plan *P, int64_t *py, int64_t *pz, int64_t *Ny,
                           int64_t *Nx) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "ph_get_pads: plan is NULL");
  if (py) *py = P->py;
  if (pz) *pz = P->pz;
  if (Ny) *Ny = P->Ny;
  if (Nx) *Nx = P->Nx;
  return TOMO_OK;
}

extern "C" int ph_apply(tomo_ph_plan *P, float *data, int64_t nproj, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr && data != nullptr, "ph_apply: NULL argument");
  TOMO_REQUIRE(nproj >= 0, "ph_apply: nproj < 0");
  hipStream_t st = (hipStream_t)stream;
  const int dy = (int)P->dy, dz = (int)P->dz;
  const float scale = 1.f / ((float)P->Ny * (float)P->Nx);
  for (int64_t b0 = 0; b0 < nproj; b0 += P->batch) {
    const unsigned nb = (unsigned)std::min<int64_t>(P->batch, nproj - b0);
    float *img = data + (size_t)b0 * dy * dz;
    {
      ProfScope ps_("k_ph_rows_fwd", st);
      hipLaunchKernelGGL(k_ph_rows_fwd, dim3((unsigned)(dy / 2), nb), dim3(P->rowThreads),
                         P->rowLds, st, img, P->F.as<float2>(), P->fx.d, dy, dz, P->pz);
    }
    TOMO_LAUNCH_CHECK();
    {
      ProfScope ps_("k_ph_cols", st);
      hipLaunchKernelGGL(k_ph_cols, dim3((unsigned)(P->Nx >> P->lgC), nb), dim3(P->colThreads),
                         P->colLds, st, P->F.as<float2>(), P->fy.d, P->lgC, P->Nx, dy, P->py,
                         P->ty.as<float>(), P->tx.as<float>(), P->method, P->A, P->B, P->alpha,
                         P->db, P->hmax, scale);
    }
    TOMO_LAUNCH_CHECK();
    {
      ProfScope ps_("k_ph_rows_inv", st);
      hipLaunchKernelGGL(k_ph_rows_inv, dim3((unsigned)(dy / 2), nb), dim3(P->rowThreads),
                         P->rowLds, st, img, P->F.as<float2>(), P->fx.d, dy, dz, P->pz);
    }
    TOMO_LAUNCH_CHECK();
  }
  return TOMO_OK;
}
