// Fourier-wavelet stripe removal (tomocupy remove_stripe_fw,
// src/tomocupy/processing/remove_stripe.py:253-293) for gfx950.
//
// Per slice of a (nproj, nz, n) chunk: zero-pad the angles to
// nproj + nproj/8 (centred), `level` separable DWT levels (detector axis, then
// angle axis; half-sample symmetric extension, stride-2 correlation with the
// reversed decomposition taps, as afb1d :91-129), notch filter of the
// detector-high/angle-low band along angles, then the synthesis filter banks
// (sfb1d :132-156) back up and the crop to the original rows.
//
// The notch irfft(rfft(x) * (1 - g), my) with g_k = exp(-yhat_k^2/(2 sigma^2))
// (:270-281) equals x - irfft(rfft(x) * g); g falls below 1e-9 after a few
// frequencies, so the correction is a rank-nk sum over the kept frequencies
// (host-side FP64 twiddle tables), exact to far below float32 resolution.
#include <cmath>
#include <cstring>
#include <string>

#include "common.h"
#include "preproc.h"
#include "wavelets.h"

using namespace tomo;

namespace {

constexpr int kMaxNotch = 32;

struct Taps {
  float a[32], b[32];
  int L;
};

__device__ __forceinline__ int refl(int m, int N) {
  // half-sample symmetric reflection into [0, N) (_reflect(x, -0.5, N-0.5), :51-68)
  const int P = 2 * N;
  int mm = m % P;
  if (mm < 0) mm += P;
  return mm >= N ? P - 1 - mm : mm;
}

// Analysis along the detector axis (rows).  MODE 0 reads the chunk layout
// res[p][z][x] with the zero angle padding; MODE 1 reads a band (z, H, W).
template <int MODE>
__global__ void k_fw_rows(const float *__restrict__ src, float *__restrict__ lo,
                          float *__restrict__ hi, Taps t, int H, int W, int Wo, int nz, int nproj,
                          int st) {
  extern __shared__ float xs[];
  const int h = blockIdx.x, z = blockIdx.y;
  const int L = t.L;
  const int pad = (2 * (Wo - 1) - W + L) / 2;
  const int np = 2 * Wo + L - 2;
  const float *row;
  bool zero = false;
  if (MODE == 0) {
    const int p = h - st;
    zero = (p < 0 || p >= nproj);
    row = src + ((size_t)(zero ? 0 : p) * nz + z) * W;
  } else {
    row = src + ((size_t)z * H + h) * W;
  }
  for (int i = threadIdx.x; i < np; i += blockDim.x) xs[i] = zero ? 0.f : row[refl(i - pad, W)];
  __syncthreads();
  float *plo = lo + ((size_t)z * H + h) * Wo;
  float *phi = hi + ((size_t)z * H + h) * Wo;
  for (int o = threadIdx.x; o < Wo; o += blockDim.x) {
    float a = 0.f, b = 0.f;
    for (int j = 0; j < L; ++j) {
      const float x = xs[2 * o + j];
      a += t.a[j] * x;
      b += t.b[j] * x;
    }
    plo[o] = a;
    phi[o] = b;
  }
}

// Analysis along the angle axis (columns) of both row bands.  A workgroup
// produces kColOut output rows of 64 columns; the 2*kColOut+L-2 input rows
// it needs (reflected at the borders) are staged once in LDS.
constexpr int kColOut = 32;
__global__ __launch_bounds__(256) void k_fw_cols(const float *__restrict__ lo,
                                                 const float *__restrict__ hi,
                                                 float *__restrict__ ll, float *__restrict__ lh,
                                                 float *__restrict__ hl, float *__restrict__ hh,
                                                 Taps t, int H, int W, int Ho) {
  __shared__ float2 xs[2 * kColOut + 30][64];
  const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const int w = blockIdx.x * 64 + lane;
  const int oh0 = blockIdx.y * kColOut, z = blockIdx.z;
  const int L = t.L;
  const int pad = (2 * (Ho - 1) - H + L) / 2;
  const int nin = 2 * kColOut + L - 2;
  const int base = 2 * oh0 - pad;
  const bool live = w < W;
  const float *plo = lo + (size_t)z * H * W + (live ? w : 0);
  const float *phi = hi + (size_t)z * H * W + (live ? w : 0);
  for (int i = sub; i < nin; i += 4) {
    const int r = refl(base + i, H);
    xs[i][lane] = live ? make_float2(plo[(size_t)r * W], phi[(size_t)r * W]) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  if (!live) return;
  for (int q = sub; q < kColOut; q += 4) {
    const int oh = oh0 + q;
    if (oh >= Ho) break;
    float a0 = 0.f, a1 = 0.f, b0 = 0.f, b1 = 0.f;
    for (int j = 0; j < L; ++j) {
      const float2 v = xs[2 * q + j][lane];
      a0 += t.a[j] * v.x;
      a1 += t.b[j] * v.x;
      b0 += t.a[j] * v.y;
      b1 += t.b[j] * v.y;
    }
    const size_t o = ((size_t)z * Ho + oh) * W + w;
    ll[o] = a0;
    lh[o] = a1;
    hl[o] = b0;
    hh[o] = b1;
  }
}

// Notch along angles of the stripe band: x_t -= sum_k gk_k Re(X_k e^{2 pi i k t/my}),
// X_k = sum_t x_t e^{-2 pi i k t/my}.  Pass 1 (k_fw_dft): partial X_k over
// blocks of kNotchRows rows for 64 columns, the 4 waves' partials combined in
// LDS in a fixed order.  Pass 2 (k_fw_notch): sum the row-block partials in
// order and correct the block's rows.
constexpr int kNotchRows = 64;
template <int KM>
__global__ __launch_bounds__(256) void k_fw_dft(const float *__restrict__ hl,
                                                const float2 *__restrict__ tw,
                                                float2 *__restrict__ part, int my, int mx, int nk) {
  __shared__ float2 red[4][KM][64];
  const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const int w = blockIdx.x * 64 + lane, rb = blockIdx.y, z = blockIdx.z;
  const int nrb = gridDim.y;
  const bool live = w < mx;
  const float *col = hl + (size_t)z * my * mx + (live ? w : 0);
  float xr[KM], xi[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) xr[k] = xi[k] = 0.f;
  const int t1 = min(my, (rb + 1) * kNotchRows);
  for (int t = rb * kNotchRows + sub; t < t1; t += 4) {
    const float x = live ? col[(size_t)t * mx] : 0.f;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      if (k < nk) {
        const float2 c = tw[k * my + t];
        xr[k] += x * c.x;
        xi[k] -= x * c.y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < KM; ++k) red[sub][k][lane] = make_float2(xr[k], xi[k]);
  __syncthreads();
  if (sub != 0 || !live) return;
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    if (k < nk) {
      const float2 p0 = red[0][k][lane], p1 = red[1][k][lane];
      const float2 p2 = red[2][k][lane], p3 = red[3][k][lane];
      part[(((size_t)z * nrb + rb) * KM + k) * mx + w] =
          make_float2((p0.x + p1.x) + (p2.x + p3.x), (p0.y + p1.y) + (p2.y + p3.y));
    }
  }
}

template <int KM>
__global__ __launch_bounds__(256) void k_fw_notch(float *__restrict__ hl,
                                                  const float2 *__restrict__ tw,
                                                  const float *__restrict__ gk,
                                                  const float2 *__restrict__ part, int my, int mx,
                                                  int nk) {
  const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const int w = blockIdx.x * 64 + lane, rb = blockIdx.y, z = blockIdx.z;
  const int nrb = gridDim.y;
  if (w >= mx) return;
  float xr[KM], xi[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    float sr = 0.f, si = 0.f;
    if (k < nk) {
      for (int q = 0; q < nrb; ++q) {
        const float2 p = part[(((size_t)z * nrb + q) * KM + k) * mx + w];
        sr += p.x;
        si += p.y;
      }
    }
    const float g = k < nk ? gk[k] : 0.f;
    xr[k] = sr * g;
    xi[k] = si * g;
  }
  float *col = hl + (size_t)z * my * mx + w;
  const int t1 = min(my, (rb + 1) * kNotchRows);
  for (int t = rb * kNotchRows + sub; t < t1; t += 4) {
    float corr = 0.f;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      if (k < nk) {
        const float2 c = tw[k * my + t];
        corr += xr[k] * c.x - xi[k] * c.y;
      }
    }
    col[(size_t)t * mx] -= corr;
  }
}

// Synthesis along angles: lo = sfb(ll, lh), hi = sfb(hl, hh); output rows
// t' in [0, 2H-L+2) (the L-2 crop of sfb1d).  ll has its own pitch (llH, llW)
// because it is the cropped reconstruction of the deeper level.  A workgroup
// produces kSynOut output rows of 64 columns from kSynOut/2 + L/2 band rows
// staged in LDS.
constexpr int kSynOut = 64;
__global__ __launch_bounds__(256) void k_fw_syn_cols(const float *__restrict__ ll, int llH, int llW,
                                                     const float *__restrict__ lh,
                                                     const float *__restrict__ hl,
                                                     const float *__restrict__ hh,
                                                     float *__restrict__ lo, float *__restrict__ hi,
                                                     Taps g, int H, int W, int Hs) {
  __shared__ float4 xs[kSynOut / 2 + 17][64];
  const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const int w = blockIdx.x * 64 + lane;
  const int tp0 = blockIdx.y * kSynOut, z = blockIdx.z;
  const int L = g.L;
  // band rows i with 2i + j = t for t in [tp0+L-2, tp0+kSynOut+L-3], j in [0, L)
  const int i0 = (tp0 - 1) >> 1;  // floor((tp0 + L - 2 - (L - 1)) / 2)
  const int nin = kSynOut / 2 + L / 2 + 1;
  const bool live = w < W;
  for (int q = sub; q < nin; q += 4) {
    const int i = i0 + q;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live && i >= 0 && i < H) {
      const size_t ob = ((size_t)z * H + i) * W + w;
      v = make_float4(ll[((size_t)z * llH + i) * llW + w], lh[ob], hl[ob], hh[ob]);
    }
    xs[q][lane] = v;
  }
  __syncthreads();
  if (!live) return;
  for (int q = sub; q < kSynOut; q += 4) {
    const int tp = tp0 + q;
    if (tp >= Hs) break;
    const int t = tp + L - 2;
    float a = 0.f, b = 0.f;
    for (int j = (t & 1); j < L; j += 2) {
      const int i = (t - j) >> 1;
      const float4 v = xs[i - i0][lane];
      a += g.a[j] * v.x + g.b[j] * v.y;
      b += g.a[j] * v.z + g.b[j] * v.w;
    }
    const size_t o = ((size_t)z * Hs + tp) * W + w;
    lo[o] = a;
    hi[o] = b;
  }
}

// Synthesis along the detector axis, one workgroup per output row with both
// input rows staged in LDS.  LAST writes rows [st, st+nproj) and columns
// [0, n) straight into the chunk layout res[p][z][x].
template <int LAST>
__global__ __launch_bounds__(256) void k_fw_syn_rows(const float *__restrict__ lo,
                                                     const float *__restrict__ hi,
                                                     float *__restrict__ out, Taps g, int Hs, int W,
                                                     int Ws, int nz, int nproj, int st, int n) {
  extern __shared__ float2 rs[];
  const int h = blockIdx.x, z = blockIdx.y;
  if (LAST && (h < st || h >= st + nproj)) return;
  const float *pl = lo + ((size_t)z * Hs + h) * W;
  const float *ph = hi + ((size_t)z * Hs + h) * W;
  for (int i = threadIdx.x; i < W; i += blockDim.x) rs[i] = make_float2(pl[i], ph[i]);
  __syncthreads();
  const int L = g.L;
  const int nout = LAST ? min(Ws, n) : Ws;
  for (int tp = threadIdx.x; tp < nout; tp += blockDim.x) {
    const int t = tp + L - 2;
    float a = 0.f;
    for (int j = (t & 1); j < L; j += 2) {
      const int i = (t - j) >> 1;
      if (i < 0 || i >= W) continue;
      const float2 v = rs[i];
      a += g.a[j] * v.x + g.b[j] * v.y;
    }
    if (LAST)
      out[((size_t)(h - st) * nz + z) * n + tp] = a;
    else
      out[((size_t)z * Hs + h) * Ws + tp] = a;
  }
}

const WaveletTaps *find_wavelet(const char *name) {
  for (const auto &w : kWavelets)
    if (std::strcmp(w.name, name) == 0) return &w;
  return nullptr;
}

int setup_fw(tomo_pre_plan *P, const WaveletTaps *wv, double sigma, int level) {
  if (P->fw_name == wv->name && P->fw_levels == level && P->fw_sigma == sigma) return TOMO_OK;
  for (auto &l : P->lv) {
    l.tw.release();
    l.gk.release();
    l.ll.release();
    l.lh.release();
    l.hl.release();
    l.hh.release();
  }
  P->lv.assign(level, FwLevel());
  const int L = wv->len;
  const int64_t nz = P->nz;
  int H = (int)(P->nproj + P->nproj / 8), W = (int)P->n;
  size_t rowmax = 0, synmax = 0, outmax = 0, partmax = 0;
  for (int k = 0; k < level; ++k) {
    FwLevel &l = P->lv[k];
    l.H = H;
    l.W = W;
    l.Ho = (H + L - 1) / 2;
    l.Wo = (W + L - 1) / 2;
    rowmax = std::max(rowmax, (size_t)H * l.Wo);
    const size_t band = sizeof(float) * nz * l.Ho * l.Wo;
    int rc;
    if ((rc = l.ll.alloc(band)) || (rc = l.lh.alloc(band)) || (rc = l.hl.alloc(band)) ||
        (rc = l.hh.alloc(band)))
      return rc;
    // notch: kept frequencies k <= my/2 with g_k >= 1e-9 (yhat of :276-279)
    const int my = l.Ho;
    std::vector<int> ks;
    std::vector<float> g;
    for (int kk = 0; kk <= my / 2; ++kk) {
      // ifftshift((arange(-my, my, 2) + 1)/2)[kk]
      const int idx = (kk + my / 2) % my;
      const double yh = (double)(-my + 2 * idx + 1) / 2.0;
      const double gv = std::exp(-yh * yh / (2.0 * sigma * sigma));
      if (gv < 1e-9) continue;
      const double ck = (kk == 0 || (my % 2 == 0 && kk == my / 2)) ? 1.0 : 2.0;
      ks.push_back(kk);
      g.push_back((float)(ck * gv / my));
    }
    TOMO_REQUIRE((int)ks.size() <= kMaxNotch,
                 "pre_remove_stripe_fw: sigma=%g keeps %zu frequencies at level %d (max %d)",
                 sigma, ks.size(), k + 1, kMaxNotch);
    l.nk = (int)ks.size();
    if (l.nk > 0) {
      std::vector<float2> tw((size_t)l.nk * my);
      for (int q = 0; q < l.nk; ++q)
        for (int t = 0; t < my; ++t) {
          const double ang = 2.0 * M_PI * (double)(((int64_t)ks[q] * t) % my) / (double)my;
          tw[(size_t)q * my + t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
      if ((rc = l.tw.alloc(sizeof(float2) * tw.size())) || (rc = l.gk.alloc(sizeof(float) * l.nk)))
        return rc;
      TOMO_HIP(hipMemcpy(l.tw.p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
      TOMO_HIP(hipMemcpy(l.gk.p, g.data(), sizeof(float) * l.nk, hipMemcpyHostToDevice));
    }
    partmax = std::max(partmax, (size_t)ceil_div(l.Ho, kNotchRows) * kMaxNotch * l.Wo);
    const size_t Hs = 2 * (size_t)l.Ho - L + 2, Ws = 2 * (size_t)l.Wo - L + 2;
    synmax = std::max(synmax, Hs * l.Wo);
    outmax = std::max(outmax, Hs * Ws);
    H = l.Ho;
    W = l.Wo;
  }
  int rc;
  if ((rc = P->rowlo.alloc(sizeof(float) * nz * rowmax)) ||
      (rc = P->rowhi.alloc(sizeof(float) * nz * rowmax)) ||
      (rc = P->synlo.alloc(sizeof(float) * nz * synmax)) ||
      (rc = P->synhi.alloc(sizeof(float) * nz * synmax)) ||
      (rc = P->syn[0].alloc(sizeof(float) * nz * outmax)) ||
      (rc = P->syn[1].alloc(sizeof(float) * nz * outmax)) ||
      (rc = P->notch.alloc(sizeof(float2) * nz * partmax)))
    return rc;
  P->fw_name = wv->name;
  P->fw_levels = level;
  P->fw_sigma = sigma;
  P->fw_L = L;
  return TOMO_OK;
}

}  // namespace

extern "C" int pre_remove_stripe_fw(tomo_pre_plan *P, float *res, double sigma, const char *wname,
                                    int level, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr && res != nullptr && wname != nullptr,
               "pre_remove_stripe_fw: NULL argument");
  TOMO_REQUIRE(sigma > 0 && std::isfinite(sigma), "pre_remove_stripe_fw: sigma must be > 0");
  TOMO_REQUIRE(level >= 1 && level <= 16, "pre_remove_stripe_fw: level=%d out of range", level);
  const WaveletTaps *wv = find_wavelet(wname);
  TOMO_REQUIRE(wv != nullptr, "pre_remove_stripe_fw: unknown wavelet '%s'", wname);
  int rc = setup_fw(P, wv, sigma, level);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int L = wv->len;
  Taps ana{}, syn{};
  ana.L = syn.L = L;
  for (int j = 0; j < L; ++j) {
    ana.a[j] = (float)wv->dec_lo[L - 1 - j];  // reversed (DWTForward, :176-183)
    ana.b[j] = (float)wv->dec_hi[L - 1 - j];
    syn.a[j] = (float)wv->rec_lo[j];          // as given (DWTInverse, :221-225)
    syn.b[j] = (float)wv->rec_hi[j];
  }
  const int nz = (int)P->nz, nproj = (int)P->nproj, n = (int)P->n;
  const int npad = nproj + nproj / 8;
  const int st0 = (npad - nproj) / 2;
  for (int k = 0; k < level; ++k) {
    FwLevel &l = P->lv[k];
    const int lds = (2 * l.Wo + L - 2) * (int)sizeof(float);
    if (k == 0) {
      ProfScope ps_("k_fw_rows", st);
      hipLaunchKernelGGL(k_fw_rows<0>, dim3(l.H, nz), dim3(256), lds, st, res, P->rowlo.as<float>(),
                         P->rowhi.as<float>(), ana, l.H, l.W, l.Wo, nz, nproj, st0);
    } else {
      ProfScope ps_("k_fw_rows", st);
      hipLaunchKernelGGL(k_fw_rows<1>, dim3(l.H, nz), dim3(256), lds, st,
                         P->lv[k - 1].ll.as<float>(), P->rowlo.as<float>(), P->rowhi.as<float>(),
                         ana, l.H, l.W, l.Wo, nz, nproj, st0);
    }
    TOMO_LAUNCH_CHECK();
    {
      ProfScope ps_("k_fw_cols", st);
      hipLaunchKernelGGL(k_fw_cols, dim3((unsigned)ceil_div(l.Wo, 64), (unsigned)ceil_div(l.Ho, kColOut), nz),
                         dim3(256), 0, st,
                         P->rowlo.as<float>(), P->rowhi.as<float>(), l.ll.as<float>(),
                         l.lh.as<float>(), l.hl.as<float>(), l.hh.as<float>(), ana, l.H, l.Wo,
                         l.Ho);
    }
    TOMO_LAUNCH_CHECK();
    if (l.nk > 0) {
      const dim3 g((unsigned)ceil_div(l.Wo, 64), (unsigned)ceil_div(l.Ho, kNotchRows), nz);
      float2 *part = P->notch.as<float2>();
#define TOMO_NOTCH(KM)                                                                        \
  do {                                                                                        \
    {                                                                                         \
      ProfScope ps_("k_fw_dft", st);                                                          \
      hipLaunchKernelGGL(k_fw_dft<KM>, g, dim3(256), 0, st, l.hl.as<float>(),                 \
                         l.tw.as<float2>(), part, l.Ho, l.Wo, l.nk);                          \
    }                                                                                         \
    TOMO_LAUNCH_CHECK();                                                                      \
    ProfScope ps_("k_fw_notch", st);                                                          \
    hipLaunchKernelGGL(k_fw_notch<KM>, g, dim3(256), 0, st, l.hl.as<float>(),                 \
                       l.tw.as<float2>(), l.gk.as<float>(), part, l.Ho, l.Wo, l.nk);          \
  } while (0)
      if (l.nk <= 8)
        TOMO_NOTCH(8);
      else if (l.nk <= 16)
        TOMO_NOTCH(16);
      else
        TOMO_NOTCH(32);
#undef TOMO_NOTCH
    }
    TOMO_LAUNCH_CHECK();
  }
  // synthesis, deepest level first; `cur` is the current low band (pitch curH x curW)
  const float *cur = P->lv[level - 1].ll.as<float>();
  int curH = P->lv[level - 1].Ho, curW = P->lv[level - 1].Wo;
  for (int k = level - 1; k >= 0; --k) {
    FwLevel &l = P->lv[k];
    const int Hb = l.Ho, Wb = l.Wo;
    TOMO_REQUIRE(curH >= Hb && curW >= Wb, "pre_remove_stripe_fw: internal band shape mismatch");
    const int Hs = 2 * Hb - L + 2, Ws = 2 * Wb - L + 2;
    {
      ProfScope ps_("k_fw_syn_cols", st);
      hipLaunchKernelGGL(k_fw_syn_cols, dim3((unsigned)ceil_div(Wb, 64), (unsigned)ceil_div(Hs, kSynOut), nz),
                         dim3(256), 0, st,
                         cur, curH, curW, l.lh.as<float>(), l.hl.as<float>(), l.hh.as<float>(),
                         P->synlo.as<float>(), P->synhi.as<float>(), syn, Hb, Wb, Hs);
    }
    TOMO_LAUNCH_CHECK();
    float *out = P->syn[k & 1].as<float>();
    {
      ProfScope ps_("k_fw_syn_rows", st);
      const size_t rl = sizeof(float2) * Wb;
      if (k == 0)
        hipLaunchKernelGGL(k_fw_syn_rows<1>, dim3(Hs, nz), dim3(256), rl, st, P->synlo.as<float>(),
                           P->synhi.as<float>(), res, syn, Hs, Wb, Ws, nz, nproj, st0, n);
      else
        hipLaunchKernelGGL(k_fw_syn_rows<0>, dim3(Hs, nz), dim3(256), rl, st, P->synlo.as<float>(),
                           P->synhi.as<float>(), out, syn, Hs, Wb, Ws, nz, nproj, st0, n);
    }
    TOMO_LAUNCH_CHECK();
    cur = out;
    curH = Hs;
    curW = Ws;
  }
  return TOMO_OK;
}
