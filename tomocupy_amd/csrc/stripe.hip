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
#include <type_traits>

#include "common.h"
#include "preproc.h"
#include "wavelets.h"

using namespace tomo;

namespace {

constexpr int kMaxNotch = 64;  // kept notch frequencies per level: sigma up to about 9.8

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

__device__ __forceinline__ int refl_fast(int m, int N) {
  return (unsigned)m < (unsigned)N ? m : refl(m, N);
}

// Dark/flat means for the fused first level (k_pre_means output).
struct DarkFlat {
  const float *d0, *fa, *fb;
  int flat_linear;
};

// Compile-time filter lengths of the supported wavelets (haar 2, db5 and sym5
// 10, sym16 32): the tap loops unroll, the taps are kernel-argument constants
// held in scalar registers, and each thread reuses one set of staged inputs
// for several outputs.
template <int V>
using IC = std::integral_constant<int, V>;

template <typename F>
void with_len(int L, F &&f) {
  if (L == 2)
    f(IC<2>());
  else if (L == 10)
    f(IC<10>());
  else
    f(IC<32>());
}

template <typename F>
void with_km(int nk, F &&f) {
  if (nk <= 8)
    f(IC<8>());
  else if (nk <= 16)
    f(IC<16>());
  else if (nk <= 32)
    f(IC<32>());
  else
    f(IC<64>());
}

template <typename F>
void with_r(int R, F &&f) {
  if (R >= 8)
    f(IC<8>());
  else if (R >= 4)
    f(IC<4>());
  else
    f(IC<2>());
}

// LDS bytes of a row kernel's staged row: `pairs` float2 entries plus the
// pair that the last thread's float4 reads may touch beyond them.
inline size_t row_lds_bytes(int pairs) { return sizeof(float4) * (size_t)((pairs + 3) / 2 + 1); }

// Analysis along the detector axis (rows), one workgroup per (row, slice).
// MODE 0 reads the chunk layout res[p][z][x] with the zero angle padding;
// MODE 1 reads a band (z, H, W); MODE 2 reads the raw uint16 chunk
// data[p][z][x] and applies the dark/flat correction on the fly (the fused
// form of darkflat_correction followed by the first level).  The reflected
// row x is staged in LDS and read as pairs (x[2k], x[2k+1]):
//   lo[o] = sum_m a[2m] x[2(o+m)] + a[2m+1] x[2(o+m)+1].
// A thread produces outputs o, o+1 (o even) from the pairs o .. o+L/2, read as
// float4 (two pairs, consecutive lanes on consecutive 16-byte words).
template <int MODE, int L>
__global__ __launch_bounds__(256) void k_fw_rows(const void *__restrict__ src_,
                                                 float *__restrict__ lo, float *__restrict__ hi,
                                                 Taps t, int H, int W, int Wo, int nz, int nproj,
                                                 int st, DarkFlat df) {
  extern __shared__ float4 xs4[];
  float *xs = reinterpret_cast<float *>(xs4);
  constexpr int L2 = L / 2, NQ = (L2 + 2) / 2;
  const int h = blockIdx.x, z = blockIdx.y;
  const int pad = (2 * (Wo - 1) - W + L) / 2;
  const int np = 2 * Wo + L - 2;
  const float *row = nullptr;
  const uint16_t *rawrow = nullptr;
  bool zero = false;
  const int p = h - st;
  if (MODE == 0 || MODE == 2) {
    zero = (p < 0 || p >= nproj);
    const size_t off = ((size_t)(zero ? 0 : p) * nz + z) * W;
    if (MODE == 0)
      row = static_cast<const float *>(src_) + off;
    else
      rawrow = static_cast<const uint16_t *>(src_) + off;
  } else {
    row = static_cast<const float *>(src_) + ((size_t)z * H + h) * W;
  }
  // linear flat interpolation weight of projection p (proc_functions.py:62-66)
  const float vlin = (MODE == 2 && nproj > 1) ? (float)((double)p / (double)(nproj - 1)) : 0.f;
  // Unreflected interior of a row with W % 4 == 0 and aligned storage: each
  // thread loads two quadruples of columns (8- or 16-byte loads of the counts
  // or values, 16-byte loads of the dark/flat means) before converting them;
  // only the L-2 reflected border samples are then loaded one by one.
  const bool vec = MODE != 1 && !zero && (W & 3) == 0 &&
                   ((MODE == 2 ? (uintptr_t)rawrow & 7 : (uintptr_t)row & 15) == 0);
  if (vec) {
    for (int q0 = threadIdx.x; 4 * q0 < W; q0 += 2 * 256) {
      float v[2][4];
#pragma unroll
      for (int h2 = 0; h2 < 2; ++h2) {
        const int x0 = 4 * (q0 + h2 * 256);
        if (x0 < W) {
          if (MODE == 2) {
            const size_t zx = (size_t)z * W + x0;
            const ushort4 r4 = *reinterpret_cast<const ushort4 *>(rawrow + x0);
            const float4 d4 = *reinterpret_cast<const float4 *>(df.d0 + zx);
            const float4 a4 = *reinterpret_cast<const float4 *>(df.fa + zx);
            const float4 b4 = df.flat_linear ? *reinterpret_cast<const float4 *>(df.fb + zx)
                                             : make_float4(0.f, 0.f, 0.f, 0.f);
            v[h2][0] = darkflat((float)r4.x, d4.x, a4.x, b4.x, vlin, df.flat_linear);
            v[h2][1] = darkflat((float)r4.y, d4.y, a4.y, b4.y, vlin, df.flat_linear);
            v[h2][2] = darkflat((float)r4.z, d4.z, a4.z, b4.z, vlin, df.flat_linear);
            v[h2][3] = darkflat((float)r4.w, d4.w, a4.w, b4.w, vlin, df.flat_linear);
          } else {
            const float4 f4 = *reinterpret_cast<const float4 *>(row + x0);
            v[h2][0] = f4.x;
            v[h2][1] = f4.y;
            v[h2][2] = f4.z;
            v[h2][3] = f4.w;
          }
        }
      }
#pragma unroll
      for (int h2 = 0; h2 < 2; ++h2) {
        const int x0 = 4 * (q0 + h2 * 256);
        if (x0 < W)
#pragma unroll
          for (int c = 0; c < 4; ++c) xs[pad + x0 + c] = v[h2][c];
      }
    }
    for (int b = threadIdx.x; b < np - W; b += 256) {
      const int i = b < pad ? b : b + W;  // staged index of a reflected sample
      const int x = refl_fast(i - pad, W);
      if (MODE == 2) {
        const size_t zx = (size_t)z * W + x;
        xs[i] = darkflat((float)rawrow[x], df.d0[zx], df.fa[zx],
                         df.flat_linear ? df.fb[zx] : 0.f, vlin, df.flat_linear);
      } else {
        xs[i] = row[x];
      }
    }
  }
  constexpr int U = 8;
  for (int i0 = vec ? np : threadIdx.x; i0 < np; i0 += U * 256) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * 256;
      if (MODE == 2) {
        v[u] = 0.f;
        if (!zero && i < np) {
          const int x = refl_fast(i - pad, W);
          const size_t zx = (size_t)z * W + x;
          v[u] = darkflat((float)rawrow[x], df.d0[zx], df.fa[zx],
                          df.flat_linear ? df.fb[zx] : 0.f, vlin, df.flat_linear);
        }
      } else {
        v[u] = (!zero && i < np) ? row[refl_fast(i - pad, W)] : 0.f;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * 256 < np) xs[i0 + u * 256] = v[u];
  }
  __syncthreads();
  float *plo = lo + ((size_t)z * H + h) * Wo;
  float *phi = hi + ((size_t)z * H + h) * Wo;
  for (int q = threadIdx.x; 2 * q < Wo; q += 256) {
    float2 x[2 * NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      const float4 v = xs4[q + u];
      x[2 * u] = make_float2(v.x, v.y);
      x[2 * u + 1] = make_float2(v.z, v.w);
    }
    float a0 = 0.f, b0 = 0.f, a1 = 0.f, b1 = 0.f;
#pragma unroll
    for (int m = 0; m < L2; ++m) {
      a0 = fmaf(t.a[2 * m], x[m].x, a0);
      a0 = fmaf(t.a[2 * m + 1], x[m].y, a0);
      b0 = fmaf(t.b[2 * m], x[m].x, b0);
      b0 = fmaf(t.b[2 * m + 1], x[m].y, b0);
      a1 = fmaf(t.a[2 * m], x[m + 1].x, a1);
      a1 = fmaf(t.a[2 * m + 1], x[m + 1].y, a1);
      b1 = fmaf(t.b[2 * m], x[m + 1].x, b1);
      b1 = fmaf(t.b[2 * m + 1], x[m + 1].y, b1);
    }
    const int o = 2 * q;
    plo[o] = a0;
    phi[o] = b0;
    if (o + 1 < Wo) {
      plo[o + 1] = a1;
      phi[o + 1] = b1;
    }
  }
}

// The stripe removal in correction form.  Only the detector-high/angle-low
// band hl of each level is changed (hl' = hl + D_k, D_k the rank-nk notch
// correction), and the synthesis filter bank is linear and reconstructs the
// analysis input exactly (half-sample symmetric extension, L-2 crop), so with
// x_k the input of level k and r_k its reconstruction,
//   r_k = IDWT(r_{k+1}, lh_k, hl_k + D_k, hh_k) = x_k + IDWT(d_{k+1}, 0, D_k, 0),
// i.e. d_k = r_k - x_k = IDWT(d_{k+1}, 0, D_k, 0) with d_level = 0, and the
// output is x_0 + d_0.  The analysis therefore keeps only ll (the next
// level's input) and the notch DFT of hl; the synthesis filters only the
// correction, and the last pass adds it to x_0.

// Analysis along the angle axis (columns).  A workgroup produces 4R output
// rows of 64 columns; the 8R+L-2 input rows of both row bands it needs
// (reflected at the borders) are staged once in LDS, with every load of a
// thread issued before the first LDS store.  Each wave produces R consecutive
// output rows, reading each of its 2R+L-2 staged rows once: ll (low-pass of
// the low row band), stored, and hl (low-pass of the high row band), of which
// only the partial DFT X_k (k < nk) along angles over the workgroup's rows is
// kept, written per row block for the synthesis; the twiddles of a row are
// the same for the whole wave (scalar loads).
template <int KM, int L, int R>
__global__ __launch_bounds__(256) void k_fw_cols(
    const float *__restrict__ lo, const float *__restrict__ hi, float *__restrict__ ll, Taps t,
    int H, int W, int Ho, const float2 *__restrict__ tw, float2 *__restrict__ part, int nk) {
  constexpr int OT = 4 * R, NIN = 2 * OT + L - 2;
  constexpr int NS = NIN > 2 * KM ? NIN : 2 * KM;  // staging, then the reduction
  __shared__ float2 xs[NS][64];
  const int lane = threadIdx.x & 63;
  const int sub = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = blockIdx.x * 64 + lane;
  const int rb = blockIdx.y, oh0 = rb * OT, z = blockIdx.z;
  const int nrb = gridDim.y;
  const int pad = (2 * (Ho - 1) - H + L) / 2;
  const int base = 2 * oh0 - pad;
  const bool live = w < W;
  const float *plo = lo + (size_t)z * H * W + (live ? w : 0);
  const float *phi = hi + (size_t)z * H * W + (live ? w : 0);
  constexpr int kIt = (NIN + 3) / 4;
  float2 v[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int i = sub + 4 * it;
    v[it] = make_float2(0.f, 0.f);
    if (live && i < NIN) {
      const size_t r = (size_t)refl_fast(base + i, H) * W;
      v[it] = make_float2(plo[r], phi[r]);
    }
  }
#pragma unroll
  for (int it = 0; it < kIt; ++it)
    if (sub + 4 * it < NIN) xs[sub + 4 * it][lane] = v[it];
  __syncthreads();
  const int q0 = sub * R;
  float a0[R], b0[R];
#pragma unroll
  for (int r = 0; r < R; ++r) a0[r] = b0[r] = 0.f;
#pragma unroll
  for (int i = 0; i < 2 * R + L - 2; ++i) {
    const float2 x = xs[2 * q0 + i][lane];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = i - 2 * r;
      if (j >= 0 && j < L) {
        a0[r] = fmaf(t.a[j], x.x, a0[r]);
        b0[r] = fmaf(t.a[j], x.y, b0[r]);
      }
    }
  }
  float xr[KM], xi[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) xr[k] = xi[k] = 0.f;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int oh = oh0 + q0 + r;
    if (oh < Ho) {
      if (live) ll[((size_t)z * Ho + oh) * W + w] = a0[r];
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        if (k < nk) {
          const float2 c = tw[k * Ho + oh];
          xr[k] = fmaf(b0[r], c.x, xr[k]);
          xi[k] = fmaf(-b0[r], c.y, xi[k]);
        }
      }
    }
  }
  if (nk == 0) return;
  // reduce the four waves' partial sums through the (now free) staging LDS:
  // waves 2,3 hand to 0,1, then 1 hands to 0
  float2 *red = &xs[0][0];
  __syncthreads();
  if (sub >= 2) {
#pragma unroll
    for (int k = 0; k < KM; ++k) red[((sub - 2) * KM + k) * 64 + lane] = make_float2(xr[k], xi[k]);
  }
  __syncthreads();
  if (sub < 2) {
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const float2 p = red[(sub * KM + k) * 64 + lane];
      xr[k] += p.x;
      xi[k] += p.y;
    }
  }
  __syncthreads();
  if (sub == 1) {
#pragma unroll
    for (int k = 0; k < KM; ++k) red[k * 64 + lane] = make_float2(xr[k], xi[k]);
  }
  __syncthreads();
  if (sub != 0 || !live) return;
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    if (k < nk) {
      const float2 p = red[k * 64 + lane];
      part[(((size_t)z * nrb + rb) * KM + k) * W + w] = make_float2(xr[k] + p.x, xi[k] + p.y);
    }
  }
}

// Notch coefficients of the stripe band: X_k = (sum over the analysis row
// blocks, in order, of the partial DFTs) * gk_k, gk_k = c_k g_k / my.  One
// launch for every level (the synthesis needs them only after the deepest
// analysis): the levels' blocks are concatenated, level j owning blocks
// [blk0, blk0 + ceil(W/256) * nk * nz); one thread per (column, k, slice).
constexpr int kXsumMaxLevels = 16;
struct XsumLevel {
  const float2 *part;
  const float *gk;
  float2 *X;
  int W, nrb, nk, KM, blk0, pad;
};
struct XsumArgs {
  XsumLevel lv[kXsumMaxLevels];
  int nlev, nz;
};

__global__ __launch_bounds__(256) void k_fw_xsum(const XsumArgs a) {
  int j = 0;
  while (j + 1 < a.nlev && (int)blockIdx.x >= a.lv[j + 1].blk0) ++j;
  const XsumLevel &l = a.lv[j];
  const int nwb = (l.W + 255) / 256;
  int b = (int)blockIdx.x - l.blk0;
  const int wb = b % nwb;
  b /= nwb;
  const int k = b % l.nk, z = b / l.nk;
  const int w = wb * 256 + threadIdx.x;
  if (w >= l.W) return;
  const float2 *p = l.part + ((size_t)z * l.nrb * l.KM + k) * l.W + w;
  const size_t stride = (size_t)l.KM * l.W;
  float sr = 0.f, si = 0.f;
#pragma unroll 8
  for (int q = 0; q < l.nrb; ++q) {
    const float2 v = p[q * stride];
    sr += v.x;
    si += v.y;
  }
  const float s = l.gk[k];
  l.X[((size_t)z * l.KM + k) * l.W + w] = make_float2(sr * s, si * s);
}

// Synthesis along angles of the correction: lo = sfb(d, 0), hi = sfb(D, 0)
// for output rows t' in [0, 2H-L+2) (the L-2 crop of sfb1d).  d is the
// correction of the deeper level (pitch dH x dW, NULL at the deepest level:
// zero).  D is generated while staging from the notch coefficients X_k
// (k_fw_xsum): D[i] = -sum_k Re(X_k e^{2 pi i k i/H}).
// With t = t' + L - 2 even (L even), outputs t and t+1 read the same band
// rows i = t/2 - m, m < L/2, through the even and the odd low-pass taps, so
// outputs come in pairs.  A workgroup produces 4R pairs (8R rows) of 64
// columns from 4R + L/2 - 1 staged rows; each wave produces R consecutive
// pairs, reading each of its R + L/2 - 1 staged rows once.
template <int KM, int L, int R>
__global__ __launch_bounds__(256) void k_fw_syn_cols(
    const float *__restrict__ dl, int dH, int dW, const float2 *__restrict__ X,
    const float2 *__restrict__ tw, int nk, float *__restrict__ lo,
    float *__restrict__ hi, Taps g, int H, int W, int Hs) {
  constexpr int L2 = L / 2, PT = 4 * R, NIN = PT + L2 - 1;
  __shared__ float2 xs[NIN][64];
  const int lane = threadIdx.x & 63;
  const int sub = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = blockIdx.x * 64 + lane;
  const int i0 = blockIdx.y * PT, tp0 = 2 * i0, z = blockIdx.z;
  const bool live = w < W;
  float xr[KM], xi[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const float2 x = (live && k < nk) ? X[((size_t)z * KM + k) * W + w] : make_float2(0.f, 0.f);
    xr[k] = x.x;
    xi[k] = x.y;
  }
  constexpr int kIt = (NIN + 3) / 4;
  float dv[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int q = sub + 4 * it, i = i0 + q;
    dv[it] = (dl && live && q < NIN && i < H) ? dl[((size_t)z * dH + i) * dW + w] : 0.f;
  }
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int q = sub + 4 * it, i = i0 + q;
    if (q < NIN) {
      float c = 0.f;
      if (i < H) {
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          if (k < nk) {
            const float2 e = tw[k * H + i];
            c -= xr[k] * e.x - xi[k] * e.y;
          }
        }
      }
      xs[q][lane] = make_float2(dv[it], c);
    }
  }
  __syncthreads();
  if (!live) return;
  const int pq0 = sub * R;
  float a0[R], b0[R], a1[R], b1[R];
#pragma unroll
  for (int r = 0; r < R; ++r) a0[r] = b0[r] = a1[r] = b1[r] = 0.f;
  // pair pq0+r takes rows pq0+r+L2-1-m for m = 0, 1, ... (in that order)
#pragma unroll
  for (int i = R + L2 - 2; i >= 0; --i) {
    const float2 x = xs[pq0 + i][lane];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = r + L2 - 1 - i;
      if (m >= 0 && m < L2) {
        const float ge = g.a[2 * m], go = g.a[2 * m + 1];
        a0[r] = fmaf(ge, x.x, a0[r]);
        b0[r] = fmaf(ge, x.y, b0[r]);
        a1[r] = fmaf(go, x.x, a1[r]);
        b1[r] = fmaf(go, x.y, b1[r]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int tp = tp0 + 2 * (pq0 + r);
    if (tp < Hs) {  // Hs is even, so row tp + 1 exists too
      const size_t o = ((size_t)z * Hs + tp) * W + w;
      lo[o] = a0[r];
      lo[o + W] = a1[r];
      hi[o] = b0[r];
      hi[o + W] = b1[r];
    }
  }
}

// Synthesis along the detector axis, one workgroup per output row with both
// input rows staged in LDS as (lo, hi) pairs; outputs 2p, 2p+1 read the pairs
// p + L/2 - 1 - m (m < L/2) through the even and odd taps.  A thread produces
// outputs 4u .. 4u+3 from the pairs 2u .. 2u+L/2, read as float4.  LAST = 0
// writes the correction d_k (pitch Hs x Ws).  LAST = 1 adds d_0 to the
// float32 chunk res[p][z][x] in place (rows [st, st+nproj), columns [0, n)).
// LAST = 2 adds d_0 to x_0 = darkflat(raw[p][z][x]) and writes the sinograms
// out[z][p][x] as TO, after minus_log when take_log (the fused form of
// proc_proj and the swap of rec.py:174).
template <int LAST, typename TO, int L>
__global__ __launch_bounds__(256) void k_fw_syn_rows(const float *__restrict__ lo,
                                                     const float *__restrict__ hi,
                                                     const void *__restrict__ src_,
                                                     TO *__restrict__ out, Taps g, int Hs, int W,
                                                     int Ws, int nz, int nproj, int st, int n,
                                                     DarkFlat df, int take_log) {
  extern __shared__ float4 rs4[];
  float2 *rs = reinterpret_cast<float2 *>(rs4);
  constexpr int L2 = L / 2, NQ = (L2 + 2) / 2;
  const int h = blockIdx.x, z = blockIdx.y;
  if (LAST && (h < st || h >= st + nproj)) return;
  const float *pl = lo + ((size_t)z * Hs + h) * W;
  const float *ph = hi + ((size_t)z * Hs + h) * W;
  constexpr int U = 4;
  for (int i0 = threadIdx.x; i0 < W; i0 += U * 256) {
    float2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * 256;
      v[u] = i < W ? make_float2(pl[i], ph[i]) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * 256 < W) rs[i0 + u * 256] = v[u];
  }
  __syncthreads();
  const int p = h - st;
  const int nout = LAST ? min(Ws, n) : Ws;
  TO *orow = LAST == 2   ? out + ((size_t)z * nproj + p) * n
             : LAST == 1 ? out + ((size_t)p * nz + z) * n
                         : out + ((size_t)z * Hs + h) * Ws;
  const uint16_t *rawrow =
      LAST == 2 ? static_cast<const uint16_t *>(src_) + ((size_t)p * nz + z) * n : nullptr;
  // linear flat interpolation weight of projection p (proc_functions.py:62-66)
  const float vlin = (LAST == 2 && nproj > 1) ? (float)((double)p / (double)(nproj - 1)) : 0.f;
  const bool vec4 = (n & 3) == 0 && ((uintptr_t)rawrow & 7) == 0;
  for (int u = threadIdx.x; 4 * u < nout; u += 256) {
    float2 x[2 * NQ];
#pragma unroll
    for (int e = 0; e < NQ; ++e) {
      const float4 v = rs4[u + e];
      x[2 * e] = make_float2(v.x, v.y);
      x[2 * e + 1] = make_float2(v.z, v.w);
    }
    float y[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int m = 0; m < L2; ++m) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {  // pair 2u + s
        const float2 xv = x[s + L2 - 1 - m];
        y[2 * s] = fmaf(g.a[2 * m], xv.x, y[2 * s]);
        y[2 * s] = fmaf(g.b[2 * m], xv.y, y[2 * s]);
        y[2 * s + 1] = fmaf(g.a[2 * m + 1], xv.x, y[2 * s + 1]);
        y[2 * s + 1] = fmaf(g.b[2 * m + 1], xv.y, y[2 * s + 1]);
      }
    }
    float xin[4] = {0.f, 0.f, 0.f, 0.f};  // x_0 of the four outputs (LAST 1, 2)
    if (LAST == 2 && vec4 && 4 * u + 3 < nout) {
      // whole aligned quadruples: one 8-byte load of counts, 16-byte loads of
      // the dark/flat means
      const size_t zx = (size_t)z * n + 4 * u;
      const ushort4 r4 = *reinterpret_cast<const ushort4 *>(rawrow + 4 * u);
      const float4 d4 = *reinterpret_cast<const float4 *>(df.d0 + zx);
      const float4 a4 = *reinterpret_cast<const float4 *>(df.fa + zx);
      const float4 b4 = df.flat_linear ? *reinterpret_cast<const float4 *>(df.fb + zx)
                                       : make_float4(0.f, 0.f, 0.f, 0.f);
      xin[0] = darkflat((float)r4.x, d4.x, a4.x, b4.x, vlin, df.flat_linear);
      xin[1] = darkflat((float)r4.y, d4.y, a4.y, b4.y, vlin, df.flat_linear);
      xin[2] = darkflat((float)r4.z, d4.z, a4.z, b4.z, vlin, df.flat_linear);
      xin[3] = darkflat((float)r4.w, d4.w, a4.w, b4.w, vlin, df.flat_linear);
    } else if (LAST) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int x0 = 4 * u + c;
        if (x0 >= nout) continue;
        if (LAST == 1) {
          xin[c] = (float)orow[x0];
        } else {
          const size_t zx = (size_t)z * n + x0;
          xin[c] = darkflat((float)rawrow[x0], df.d0[zx], df.fa[zx],
                            df.flat_linear ? df.fb[zx] : 0.f, vlin, df.flat_linear);
        }
      }
    }
    float yo[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      yo[c] = y[c] + xin[c];
      if (LAST == 2 && take_log) yo[c] = minus_log(yo[c]);
    }
    if (LAST == 2 && vec4 && 4 * u + 3 < nout &&
        ((uintptr_t)(orow + 4 * u) & (4 * sizeof(TO) - 1)) == 0) {
      // one 16-byte (float) or 8-byte (16-bit) store per thread
      if constexpr (sizeof(TO) == 4) {
        *reinterpret_cast<float4 *>(orow + 4 * u) = make_float4(yo[0], yo[1], yo[2], yo[3]);
      } else {
        TO q[4] = {(TO)yo[0], (TO)yo[1], (TO)yo[2], (TO)yo[3]};
        *reinterpret_cast<uint2 *>(orow + 4 * u) = __builtin_bit_cast(uint2, q);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (4 * u + c < nout) orow[4 * u + c] = (TO)yo[c];
    }
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
    l.part.release();
    l.X.release();
  }
  P->lv.assign(level, FwLevel());
  const int L = wv->len;
  const int64_t nz = P->nz;
  int H = (int)(P->nproj + P->nproj / 8), W = (int)P->n;
  size_t rowmax = 0, synmax = 0, outmax = 0;
  for (int k = 0; k < level; ++k) {
    FwLevel &l = P->lv[k];
    l.H = H;
    l.W = W;
    l.Ho = (H + L - 1) / 2;
    l.Wo = (W + L - 1) / 2;
    rowmax = std::max(rowmax, (size_t)H * l.Wo);
    int rc;
    if ((rc = l.ll.alloc(sizeof(float) * nz * l.Ho * l.Wo))) return rc;
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
    // column analysis: 4R output rows per workgroup (R = 8 for tall bands,
    // fewer, for more workgroups, in the short bands of the deep levels);
    // partial notch DFTs per row block, kept until the synthesis
    l.KM = l.nk <= 8 ? 8 : (l.nk <= 16 ? 16 : (l.nk <= 32 ? 32 : 64));
    l.R = l.Ho >= 256 ? 8 : (l.Ho >= 64 ? 4 : 2);
    l.nrb = (int)ceil_div(l.Ho, 4 * l.R);
    if (l.nk > 0 && ((rc = l.part.alloc(sizeof(float2) * nz * l.nrb * l.KM * l.Wo)) ||
                     (rc = l.X.alloc(sizeof(float2) * nz * l.KM * l.Wo))))
      return rc;
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
      (rc = P->syn[1].alloc(sizeof(float) * nz * outmax)))
    return rc;
  P->fw_name = wv->name;
  P->fw_levels = level;
  P->fw_sigma = sigma;
  P->fw_L = L;
  return TOMO_OK;
}

}  // namespace

namespace {

// The level loop of the Fourier-wavelet stripe removal.  src_mode 0: src is
// the float32 transmission chunk res[p][z][x]; 2: src is the raw uint16 chunk
// (dark/flat applied in the first pass, means in the plan).  dst_mode 1: dst
// is the float32 chunk res[p][z][x] (in place allowed); 2: dst is the
// sinogram-order output of dtype out_dtype, after minus_log when take_log.
int fw_run(tomo_pre_plan *P, const WaveletTaps *wv, double sigma, int level, const void *src,
           int src_mode, DarkFlat df, void *dst, int dst_mode, int out_dtype, int take_log,
           hipStream_t st) {
  int rc = setup_fw(P, wv, sigma, level);
  if (rc) return rc;
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
  TOMO_REQUIRE(dst_mode != 2 || src_mode == 2,
               "fw_run: the sinogram output needs the raw chunk as input");
  for (int k = 0; k < level; ++k) {
    FwLevel &l = P->lv[k];
    const size_t lds = row_lds_bytes(l.Wo + L / 2 - 1);
    with_len(L, [&](auto lc) {
      constexpr int LC = decltype(lc)::value;
      ProfScope ps_("k_fw_rows", st);
      if (k > 0)
        hipLaunchKernelGGL((k_fw_rows<1, LC>), dim3(l.H, nz), dim3(256), lds, st,
                           (const void *)P->lv[k - 1].ll.as<float>(), P->rowlo.as<float>(),
                           P->rowhi.as<float>(), ana, l.H, l.W, l.Wo, nz, nproj, st0, df);
      else if (src_mode == 2)
        hipLaunchKernelGGL((k_fw_rows<2, LC>), dim3(l.H, nz), dim3(256), lds, st, src,
                           P->rowlo.as<float>(), P->rowhi.as<float>(), ana, l.H, l.W, l.Wo, nz,
                           nproj, st0, df);
      else
        hipLaunchKernelGGL((k_fw_rows<0, LC>), dim3(l.H, nz), dim3(256), lds, st, src,
                           P->rowlo.as<float>(), P->rowhi.as<float>(), ana, l.H, l.W, l.Wo, nz,
                           nproj, st0, df);
    });
    TOMO_LAUNCH_CHECK();
    const dim3 gc((unsigned)ceil_div(l.Wo, 64), l.nrb, nz);
    with_km(l.nk, [&](auto kmc) {
      with_len(L, [&](auto lc) {
        with_r(l.R, [&](auto rc) {
          constexpr int KM = decltype(kmc)::value, LC = decltype(lc)::value,
                        RC = decltype(rc)::value;
          ProfScope ps_("k_fw_cols", st);
          hipLaunchKernelGGL((k_fw_cols<KM, LC, RC>), gc, dim3(256), 0, st,
                             P->rowlo.as<float>(), P->rowhi.as<float>(), l.ll.as<float>(), ana,
                             l.H, l.Wo, l.Ho, l.tw.as<float2>(), l.part.as<float2>(), l.nk);
        });
      });
    });
    TOMO_LAUNCH_CHECK();
  }
  {
    TOMO_REQUIRE(level <= kXsumMaxLevels, "pre_remove_stripe_fw: more than %d levels",
                 kXsumMaxLevels);
    XsumArgs xa{};
    xa.nz = nz;
    int blk = 0;
    for (int k = 0; k < level; ++k) {
      const FwLevel &l = P->lv[k];
      if (l.nk == 0) continue;
      XsumLevel &x = xa.lv[xa.nlev++];
      x.part = l.part.as<float2>();
      x.gk = l.gk.as<float>();
      x.X = l.X.as<float2>();
      x.W = l.Wo;
      x.nrb = l.nrb;
      x.nk = l.nk;
      x.KM = l.KM;
      x.blk0 = blk;
      blk += (int)ceil_div(l.Wo, 256) * l.nk * nz;
    }
    if (xa.nlev > 0) {
      ProfScope ps_("k_fw_xsum", st);
      hipLaunchKernelGGL(k_fw_xsum, dim3((unsigned)blk), dim3(256), 0, st, xa);
      TOMO_LAUNCH_CHECK();
    }
  }
  // synthesis of the correction, deepest level first; `cur` is the correction
  // of the level below (pitch curH x curW), NULL (zero) at the deepest level
  const float *cur = nullptr;
  int curH = 0, curW = 0;
  for (int k = level - 1; k >= 0; --k) {
    FwLevel &l = P->lv[k];
    const int Hb = l.Ho, Wb = l.Wo;
    TOMO_REQUIRE(!cur || (curH >= Hb && curW >= Wb),
                 "pre_remove_stripe_fw: internal band shape mismatch");
    const int Hs = 2 * Hb - L + 2, Ws = 2 * Wb - L + 2;
    // 8R output rows per workgroup
    const int Rs = Hs >= 512 ? 8 : (Hs >= 128 ? 4 : 2);
    const dim3 gs((unsigned)ceil_div(Wb, 64), (unsigned)ceil_div(Hs, 8 * Rs), nz);
    with_km(l.nk, [&](auto kmc) {
      with_len(L, [&](auto lc) {
        with_r(Rs, [&](auto rc) {
          constexpr int KM = decltype(kmc)::value, LC = decltype(lc)::value,
                        RC = decltype(rc)::value;
          ProfScope ps_("k_fw_syn_cols", st);
          hipLaunchKernelGGL((k_fw_syn_cols<KM, LC, RC>), gs, dim3(256), 0, st, cur, curH, curW,
                             l.X.as<float2>(), l.tw.as<float2>(), l.nk, P->synlo.as<float>(),
                             P->synhi.as<float>(), syn, Hb, Wb, Hs);
        });
      });
    });
    TOMO_LAUNCH_CHECK();
    float *out = P->syn[k & 1].as<float>();
    {
      const size_t rl = row_lds_bytes(Wb);
      const dim3 gr(Hs, nz);
      const float *slo = P->synlo.as<float>(), *shi = P->synhi.as<float>();
      with_len(L, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        ProfScope ps_("k_fw_syn_rows", st);
        if (k > 0)
          hipLaunchKernelGGL((k_fw_syn_rows<0, float, LC>), gr, dim3(256), rl, st, slo, shi,
                             (const void *)nullptr, out, syn, Hs, Wb, Ws, nz, nproj, st0, n, df,
                             0);
        else if (dst_mode == 1)
          hipLaunchKernelGGL((k_fw_syn_rows<1, float, LC>), gr, dim3(256), rl, st, slo, shi,
                             (const void *)nullptr, (float *)dst, syn, Hs, Wb, Ws, nz, nproj,
                             st0, n, df, 0);
        else if (out_dtype == TOMO_F16)
          hipLaunchKernelGGL((k_fw_syn_rows<2, _Float16, LC>), gr, dim3(256), rl, st, slo, shi,
                             src, (_Float16 *)dst, syn, Hs, Wb, Ws, nz, nproj, st0, n, df,
                             take_log);
        else if (out_dtype == TOMO_BF16)
          hipLaunchKernelGGL((k_fw_syn_rows<2, __bf16, LC>), gr, dim3(256), rl, st, slo, shi,
                             src, (__bf16 *)dst, syn, Hs, Wb, Ws, nz, nproj, st0, n, df,
                             take_log);
        else
          hipLaunchKernelGGL((k_fw_syn_rows<2, float, LC>), gr, dim3(256), rl, st, slo, shi,
                             src, (float *)dst, syn, Hs, Wb, Ws, nz, nproj, st0, n, df,
                             take_log);
      });
    }
    TOMO_LAUNCH_CHECK();
    cur = out;
    curH = Hs;
    curW = Ws;
  }
  return TOMO_OK;
}

const WaveletTaps *check_wavelet(const char *fn, const char *wname, double sigma, int level) {
  if (!(sigma > 0 && std::isfinite(sigma))) {
    set_error("%s: sigma must be > 0", fn);
    return nullptr;
  }
  if (level < 1 || level > 16) {
    set_error("%s: level=%d out of range", fn, level);
    return nullptr;
  }
  const WaveletTaps *wv = find_wavelet(wname);
  if (!wv) {
    set_error("%s: unknown wavelet '%s'", fn, wname);
    return nullptr;
  }
  if (wv->len != 2 && wv->len != 10 && wv->len != 32) {
    set_error("%s: wavelet '%s' has %d taps (2, 10 or 32 supported)", fn, wname, wv->len);
    return nullptr;
  }
  return wv;
}

}  // namespace

extern "C" int pre_remove_stripe_fw(tomo_pre_plan *P, float *res, double sigma, const char *wname,
                                    int level, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr && res != nullptr && wname != nullptr,
               "pre_remove_stripe_fw: NULL argument");
  const WaveletTaps *wv = check_wavelet("pre_remove_stripe_fw", wname, sigma, level);
  if (!wv) return TOMO_EINVAL;
  return fw_run(P, wv, sigma, level, res, 0, DarkFlat{nullptr, nullptr, nullptr, 0}, res, 1,
                TOMO_F32, 0, (hipStream_t)stream);
}

extern "C" int pre_process_fw(tomo_pre_plan *P, const uint16_t *data, const uint16_t *dark,
                              const uint16_t *flat, void *out, int out_dtype, double bright_ratio,
                              int flat_linear, int minus_log, double sigma, const char *wname,
                              int level, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr && data && dark && flat && out && wname,
               "pre_process_fw: NULL argument");
  TOMO_REQUIRE(out_dtype == TOMO_F32 || out_dtype == TOMO_F16 || out_dtype == TOMO_BF16,
               "pre_process_fw: unknown out_dtype %d", out_dtype);
  const WaveletTaps *wv = check_wavelet("pre_process_fw", wname, sigma, level);
  if (!wv) return TOMO_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  int rc = pre_launch_means(P, dark, flat, bright_ratio, flat_linear, st);
  if (rc) return rc;
  const DarkFlat df{P->dark0.as<float>(), P->flatA.as<float>(), P->flatB.as<float>(),
                    flat_linear};
  return fw_run(P, wv, sigma, level, data, 2, df, out, 2, out_dtype, minus_log, st);
}
