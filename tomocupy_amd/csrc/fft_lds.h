// LDS-resident Stockham autosort FFT for sizes N <= 16384 whose prime factors
// are at most kFftMaxPrime (per transform), executed cooperatively by a
// workgroup on 2^lgH interleaved transforms held in LDS.  Radices 2, 3, 4, 5,
// 7, 8, 11, 13, 16 and 17 run as butterflies in registers; any other prime p
// runs as a direct length-p DFT per output (fft_pass_gen).
//
// Pass with radix R and current sub-transform length Ns (1, R1, R1*R2, ...):
//   for j in [0, N/R):  k = j mod Ns
//     a_r = x[j + r N/R] * W^(r k N/(Ns R))         (r = 0..R-1)
//     b   = DFT_R(a)
//     y[(j/Ns) Ns R + k + q Ns] = b_q
// which leaves the output in natural order.  Every thread stages all of its
// butterflies in registers before the barrier, so one LDS buffer suffices.
// Twiddles come from a per-size table W_N^k = exp(-2 pi i k/N) computed on the
// host in FP64; the inverse transform uses their conjugates.  Unnormalised in
// both directions.
#pragma once

#include <algorithm>
#include <type_traits>

#include "common.h"

namespace tomo {

constexpr int kFftMaxRadices = 16;
constexpr int kFftEPT = 16;  // complex elements staged per thread per pass (max)
constexpr int kFftMaxPrime = 127;  // largest prime factor of a transform length

struct FftDesc {
  int N;
  int nr;
  int r[kFftMaxRadices];
  const float2 *tw;  // device, N entries
};

// Host-side plan: radix sequence + device twiddle table.
struct FftPlan {
  FftDesc d{};
  DevBuf tw;
  int init(int N);  // N <= 16384, prime factors <= kFftMaxPrime
  // the radix sequence of init() alone (no device memory, no HIP call)
  static int radices(int N, FftDesc &d);
};

template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
  // multiply by (-i) for the forward transform, (+i) for the inverse
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int R, int DIR>
__device__ __forceinline__ void butterfly(float2 *a) {
  if constexpr (R == 2) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  } else if constexpr (R == 3) {
    const float s3 = 0.86602540378443864676f;  // sqrt(3)/2
    float2 t1 = cadd(a[1], a[2]);
    float2 t2 = csub(a[1], a[2]);
    float2 m = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
    float2 r = cscale(mul_i<DIR>(t2), s3);
    a[0] = cadd(a[0], t1);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  } else if constexpr (R == 5) {
    // y_k = sum_j a_j W^(jk), W = exp(-+2 pi i/5): pairs a_j +- a_(5-j) with
    // c_m = cos(2 pi m/5), s_m = sin(2 pi m/5)
    const float c1 = 0.30901699437494745f, c2 = -0.80901699437494734f;
    const float s1 = 0.95105651629515353f, s2 = 0.58778525229247320f;
    const float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    const float2 t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
    const float2 b1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
    const float2 b2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
    const float2 d1 = mul_i<DIR>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    const float2 d2 = mul_i<DIR>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(b1, d1);
    a[4] = csub(b1, d1);
    a[2] = cadd(b2, d2);
    a[3] = csub(b2, d2);
  } else if constexpr (R == 7) {
    // pairs a_j +- a_(7-j), j = 1..3; y_m = a0 + sum_j cos(2 pi jm/7) t_j
    // -+ i sum_j sin(2 pi jm/7) u_j, y_(7-m) its conjugate combination
    const float C[4] = {1.f, 0.62348980185873353f, -0.22252093395631440f,
                        -0.90096886790241913f};
    const float S[4] = {0.f, 0.78183148246802981f, 0.97492791218182361f,
                        0.43388373911755812f};
    float2 t[4], u[4];
#pragma unroll
    for (int j = 1; j <= 3; ++j) {
      t[j] = cadd(a[j], a[7 - j]);
      u[j] = csub(a[j], a[7 - j]);
    }
    float2 y[7];
    y[0] = cadd(a[0], cadd(t[1], cadd(t[2], t[3])));
#pragma unroll
    for (int m = 1; m <= 3; ++m) {
      float2 b = a[0], d = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 1; j <= 3; ++j) {
        const int q = (j * m) % 7;  // cos(2 pi q/7) = C[min(q, 7-q)], sin = +-S[min(q, 7-q)]
        const float c = C[q <= 3 ? q : 7 - q];
        const float sg = q <= 3 ? S[q] : -S[7 - q];
        b = make_float2(b.x + c * t[j].x, b.y + c * t[j].y);
        d = make_float2(d.x + sg * u[j].x, d.y + sg * u[j].y);
      }
      const float2 di = mul_i<DIR>(d);
      y[m] = cadd(b, di);
      y[7 - m] = csub(b, di);
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) a[k] = y[k];
  } else if constexpr (R == 4) {
    float2 t0 = cadd(a[0], a[2]);
    float2 t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]);
    float2 t3 = mul_i<DIR>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  } else if constexpr (R == 8) {
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    butterfly<4, DIR>(e);
    butterfly<4, DIR>(o);
    const float h = 0.70710678118654752440f;
    // w8^q for q = 1,2,3 with w8 = exp(DIR * i pi/4)
    float2 o1 = make_float2(h * (o[1].x - DIR * o[1].y), h * (o[1].y + DIR * o[1].x));
    float2 o2 = mul_i<DIR>(o[2]);
    float2 o3 = make_float2(h * (-o[3].x - DIR * o[3].y), h * (-o[3].y + DIR * o[3].x));
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
  } else if constexpr (R == 16) {
    // 16 = 4 x 4 in place: radix-4 over r2 for each r1 (results q2 back in
    // slots r1 + 4 q2), internal twiddles w16^(r1 q2), radix-4 over r1 for each
    // q2 (results q1 in slots 4 q2 + q1).  Output q = q2 + 4 q1 therefore sits
    // in slot out_slot<16>(q) = 4 (q & 3) + (q >> 2).
#pragma unroll
    for (int r1 = 0; r1 < 4; ++r1) {
      float2 t[4] = {a[r1], a[r1 + 4], a[r1 + 8], a[r1 + 12]};
      butterfly<4, DIR>(t);
      a[r1] = t[0];
      a[r1 + 4] = t[1];
      a[r1 + 8] = t[2];
      a[r1 + 12] = t[3];
    }
    // w16^m = exp(DIR * 2 pi i m / 16), m = r1*q2 in {1,2,3,4,6,9}
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float h = 0.70710678118654752440f;
    const float2 w1 = make_float2(c1, DIR * s1), w2 = make_float2(h, DIR * h);
    const float2 w3 = make_float2(s1, DIR * c1);
    const float2 w6 = make_float2(-h, DIR * h), w9 = make_float2(-c1, -DIR * s1);
    a[1 + 4 * 1] = cmul(a[1 + 4 * 1], w1);
    a[1 + 4 * 2] = cmul(a[1 + 4 * 2], w2);
    a[1 + 4 * 3] = cmul(a[1 + 4 * 3], w3);
    a[2 + 4 * 1] = cmul(a[2 + 4 * 1], w2);
    a[2 + 4 * 2] = mul_i<DIR>(a[2 + 4 * 2]);  // w16^4 = exp(DIR i pi/2) = DIR * i
    a[2 + 4 * 3] = cmul(a[2 + 4 * 3], w6);
    a[3 + 4 * 1] = cmul(a[3 + 4 * 1], w3);
    a[3 + 4 * 2] = cmul(a[3 + 4 * 2], w6);
    a[3 + 4 * 3] = cmul(a[3 + 4 * 3], w9);
#pragma unroll
    for (int q2 = 0; q2 < 4; ++q2) butterfly<4, DIR>(a + 4 * q2);
  }
}

// Register slot holding output q of butterfly<R>: the identity except for the
// in-place radix-16 butterfly.
template <int R>
__device__ __forceinline__ constexpr int out_slot(int q) {
  return R == 16 ? 4 * (q & 3) + (q >> 2) : q;
}

// LDS index of element i of a transform: one pad slot per 16 elements, which
// breaks the power-of-two strides of the early Stockham passes across banks.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
// LDS elements needed for a padded transform of length N.
__host__ __device__ inline int lds_padded_len(int N) { return N + (N >> 4); }

// a[r] *= w^r for r = 1..R-1 with w = W_N^base (conjugated for the inverse).
// Powers of two of w come from the FP64-built table; the others are products
// of at most three of them (relative error a few float32 ulps).
template <int R, int DIR>
__device__ __forceinline__ void twiddle(float2 *a, const float2 *__restrict__ tw, unsigned base) {
  auto ld = [&](unsigned m) {
    float2 w = tw[base * m];
    if (DIR > 0) w.y = -w.y;
    return w;
  };
  if constexpr (R == 2) {
    a[1] = cmul(a[1], ld(1));
  } else if constexpr (R == 3) {
    a[1] = cmul(a[1], ld(1));
    a[2] = cmul(a[2], ld(2));
  } else if constexpr (R == 5) {
    const float2 w1 = ld(1), w2 = ld(2), w4 = ld(4);
    a[1] = cmul(a[1], w1);
    a[2] = cmul(a[2], w2);
    a[3] = cmul(a[3], cmul(w2, w1));
    a[4] = cmul(a[4], w4);
  } else if constexpr (R == 7) {
    const float2 w1 = ld(1), w2 = ld(2), w4 = ld(4);
    a[1] = cmul(a[1], w1);
    a[2] = cmul(a[2], w2);
    a[3] = cmul(a[3], cmul(w2, w1));
    a[4] = cmul(a[4], w4);
    a[5] = cmul(a[5], cmul(w4, w1));
    a[6] = cmul(a[6], cmul(w4, w2));
  } else if constexpr (R == 4) {
    const float2 w1 = ld(1), w2 = ld(2);
    a[1] = cmul(a[1], w1);
    a[2] = cmul(a[2], w2);
    a[3] = cmul(a[3], cmul(w2, w1));
  } else {
    float2 w[R];
    w[1] = ld(1);
    w[2] = ld(2);
    w[4] = ld(4);
    w[3] = cmul(w[2], w[1]);
    w[5] = cmul(w[4], w[1]);
    w[6] = cmul(w[4], w[2]);
    w[7] = cmul(w[4], w[3]);
    if constexpr (R == 16) {
      w[8] = ld(8);
#pragma unroll
      for (int r = 9; r < 16; ++r) w[r] = cmul(w[8], w[r - 8]);
    }
#pragma unroll
    for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[r]);
  }
}

template <int R, int DIR, int EPT>
__device__ __forceinline__ void fft_pass(float2 *buf, int N, int Ns, const float2 *tw, int lgH,
                                         int tid, int nt) {
  constexpr int MAXB = EPT / R;
  const int H = 1 << lgH;
  const int nbPer = N / R;
  const int nb = nbPer << lgH;
  const int twstep = N / (Ns * R);
  const bool pow2 = (Ns & (Ns - 1)) == 0;
  const int lgNs = pow2 ? __ffs(Ns) - 1 : 0;
  const float invNs = 1.0f / (float)Ns;
  // j -> (j / Ns, j % Ns); shifts when Ns is a power of two (all passes before
  // the radix-3 ones), otherwise a float-reciprocal quotient with correction
  auto split = [&](int j, int &jq, int &k) {
    if (pow2) {
      jq = j >> lgNs;
      k = j & (Ns - 1);
    } else {
      jq = (int)((float)j * invNs);
      k = j - jq * Ns;
      if (k < 0) { jq -= 1; k += Ns; }
      if (k >= Ns) { jq += 1; k -= Ns; }
    }
  };
  float2 v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> lgH;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[(lds_pad(j + r * nbPer) << lgH) + t];
      if (Ns > 1) {
        int jq, k;
        split(j, jq, k);
        twiddle<R, DIR>(v[q], tw, (unsigned)(k * twstep));
      }
      butterfly<R, DIR>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> lgH;
      int jq, k;
      split(j, jq, k);
      const int o = jq * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(lds_pad(o + r * Ns) << lgH) + t] = v[q][out_slot<R>(r)];
    }
  }
  __syncthreads();
}

// Radix-R pass for an odd prime 11 <= R <= 17 with the butterfly in
// registers: inputs twiddled by W_N^(r k N/(Ns R)) from the table, then the
// DFT from the cosine/sine pairs P_j = a_j + a_(R-j), M_j = a_j - a_(R-j):
//   b_q, b_(R-q) = a_0 + sum_j cos(2 pi j q/R) P_j  -/+ i sum_j sin(2 pi j q/R) M_j
// (signs of i swapped for the inverse), cos and sin read from the table
// (W_R^m = W_N^(m N/R)).  One butterfly per thread and round when R > EPT.
template <int R, int DIR, int EPT>
__device__ __noinline__ void fft_pass_prime(float2 *buf, int N, int Ns, const float2 *tw,
                                            int lgH, int tid, int nt) {
  constexpr int MAXB = EPT / R > 0 ? EPT / R : 1;
  constexpr int HR = (R - 1) / 2;
  const int H = 1 << lgH;
  const int nbPer = N / R;
  const int nb = nbPer << lgH;
  const int twstep = N / (Ns * R);
  float cs[HR + 1], sn[HR + 1];  // cos, sin(2 pi m / R), m = 1..HR
#pragma unroll
  for (int m = 1; m <= HR; ++m) {
    const float2 w = tw[m * nbPer];
    cs[m] = w.x;
    sn[m] = -w.y;
  }
  float2 v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> lgH;
      const int k = j % Ns;
      const unsigned base = (unsigned)(k * twstep);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = buf[(lds_pad(j + r * nbPer) << lgH) + t];
        if (r > 0 && Ns > 1) {
          float2 w = tw[base * r];
          if (DIR > 0) w.y = -w.y;
          x = cmul(x, w);
        }
        v[q][r] = x;
      }
      float2 P[HR + 1], M[HR + 1];
      float2 b0 = v[q][0];
#pragma unroll
      for (int jj = 1; jj <= HR; ++jj) {
        P[jj] = make_float2(v[q][jj].x + v[q][R - jj].x, v[q][jj].y + v[q][R - jj].y);
        M[jj] = make_float2(v[q][jj].x - v[q][R - jj].x, v[q][jj].y - v[q][R - jj].y);
        b0.x += P[jj].x;
        b0.y += P[jj].y;
      }
#pragma unroll
      for (int qq = 1; qq <= HR; ++qq) {
        float2 re = v[q][0], im = make_float2(0.f, 0.f);
#pragma unroll
        for (int jj = 1; jj <= HR; ++jj) {
          const int m = (jj * qq) % R;  // compile-time after unrolling
          const float c = m <= HR ? cs[m] : cs[R - m];
          const float sv = m <= HR ? sn[m] : -sn[R - m];
          re.x += c * P[jj].x;
          re.y += c * P[jj].y;
          im.x += sv * M[jj].x;
          im.y += sv * M[jj].y;
        }
        const float2 mi = mul_i<DIR>(im);
        v[q][qq] = make_float2(re.x + mi.x, re.y + mi.y);
        v[q][R - qq] = make_float2(re.x - mi.x, re.y - mi.y);
      }
      v[q][0] = b0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> lgH;
      const int jq = j / Ns, k = j - jq * Ns;
      const int o = jq * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(lds_pad(o + r * Ns) << lgH) + t] = v[q][r];
    }
  }
  __syncthreads();
}

// Pass with a prime radix 19 <= R <= kFftMaxPrime: each output y[o],
// o = jq Ns R + q Ns + k, j = jq Ns + k, is the direct sum
//   y[o] = sum_r x[j + r N/R] W_N^(r e mod N),  e = k N/(Ns R) + q N/R,
// (the twiddle W_N^(r k N/(Ns R)) and the DFT_R factor W_R^(r q) in one table
// entry), at most EPT outputs per thread held in registers across the barrier.
template <int DIR, int EPT>
__device__ __noinline__ void fft_pass_gen(float2 *buf, int N, int Ns, int R, const float2 *tw,
                                          int lgH, int tid, int nt) {
  const int H = 1 << lgH;
  const int nbPer = N / R;
  const int total = N << lgH;
  const int twstep = N / (Ns * R);
  float2 acc[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int g = tid + q * nt;
    if (g < total) {
      const int t = g & (H - 1), o = g >> lgH;
      const int k = o % Ns, rest = o / Ns;
      const int qq = rest % R, jq = rest / R;
      const int j = jq * Ns + k;
      const int e = k * twstep + qq * nbPer;  // < N
      float2 s = make_float2(0.f, 0.f);
      int m = 0;
      for (int r = 0; r < R; ++r) {
        const float2 x = buf[(lds_pad(j + r * nbPer) << lgH) + t];
        float2 w = tw[m];
        if (DIR > 0) w.y = -w.y;
        s.x += x.x * w.x - x.y * w.y;
        s.y += x.x * w.y + x.y * w.x;
        m += e;
        if (m >= N) m -= N;
      }
      acc[q] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int g = tid + q * nt;
    if (g < total) buf[(lds_pad(g >> lgH) << lgH) + (g & (H - 1))] = acc[q];
  }
  __syncthreads();
}

// fft_pass behind a call: the run-time-plan transform calls one per pass, so
// the register demand of the largest radix does not add to the caller's.
template <int R, int DIR, int EPT>
__device__ __noinline__ void fft_pass_call(float2 *buf, int N, int Ns, const float2 *tw, int lgH,
                                           int tid, int nt) {
  fft_pass<R, DIR, EPT>(buf, N, Ns, tw, lgH, tid, nt);
}

// In-place transform of 2^lgH interleaved sequences of length d.N held in LDS:
// element i of sequence t lives at buf[lds_pad(i) * 2^lgH + t].  Requires
// 2^lgH * N <= 15 * nt (16 * nt without radix-3 passes); see fft_threads().
// The caller must __syncthreads() after filling buf; the function ends with a
// barrier.
// LV (fft_level of the sizes a kernel serves): 0 = radices 2, 3, 4, 8, 16;
// 1 = also the butterflies of 5, 7, 11, 13, 17; 2 = also the direct-sum
// primes 19..127.  Each level is compiled only into the kernels instantiated
// for such sizes: the extra passes raise the register demand of every
// instantiation.
template <int DIR, int EPT = kFftEPT, int LV = 0>
__device__ void lds_fft(float2 *buf, const FftDesc &d, int lgH, int tid, int nt) {
  int Ns = 1;
  for (int s = 0; s < d.nr; ++s) {
    const int R = d.r[s];
    if (R == 16) {
      if constexpr (EPT >= 16) fft_pass_call<16, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
    } else if (R == 8) {
      fft_pass_call<8, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
    }
    else if (R == 4)
      fft_pass_call<4, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
    else if (R == 2)
      fft_pass_call<2, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
    else if (R == 3)
      fft_pass_call<3, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
    else if constexpr (LV >= 1) {
      if (R == 5)
        fft_pass_call<5, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
      else if (R == 7)
        fft_pass_call<7, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
      else if (R == 11)
        fft_pass_prime<11, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
      else if (R == 13)
        fft_pass_prime<13, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
      else if (R == 17)
        fft_pass_prime<17, DIR, EPT>(buf, d.N, Ns, d.tw, lgH, tid, nt);
      else if constexpr (LV >= 2)
        fft_pass_gen<DIR, EPT>(buf, d.N, Ns, R, d.tw, lgH, tid, nt);
    }
    Ns *= R;
  }
}

// ---- transforms of a compile-time size ---------------------------------------
// The same Stockham passes with N, the radix sequence, the sub-transform
// lengths, the twiddle steps and the interleave 2^LGH all constants, for
// N = 2^k 3^a 5^b with 2^k >= 256.  The radices are those FftPlan::radices
// picks: radix 16 passes, one radix 8/4/2 pass for the remaining power of
// two, then the radix-3 and radix-5 passes, so every sub-transform length Ns
// is 1 or a multiple of 16 (its index splits are divisions by constants).  With one pad slot per 16 elements, the R reads of a butterfly
// (stride N/R, a multiple of 16) and its R writes (stride Ns, a multiple of
// 16, or consecutive when Ns = 1) sit at constant offsets from one base
// address, so a pass costs one address computation per butterfly instead of
// one per element.
constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

constexpr int count_factor(int x, int p) { return x % p == 0 ? 1 + count_factor(x / p, p) : 0; }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <int N>
struct StaticPlan {
  static constexpr int n3 = count_factor(N, 3), n5 = count_factor(N, 5);
  static constexpr int P2 = N / (ipow(3, n3) * ipow(5, n5));
  static_assert(P2 >= 256 && (P2 & (P2 - 1)) == 0, "2^k 3^a 5^b with 2^k >= 256");
  static constexpr int n16 = ilog2c(P2) / 4;
  static constexpr int rest = 1 << (ilog2c(P2) % 4);
  static constexpr int n2p = n16 + (rest > 1 ? 1 : 0);  // passes of the power of two
  static constexpr int passes = n2p + n3 + n5;
  static constexpr int radix(int s) {
    return s < n16 ? 16 : s < n2p ? rest : s < n2p + n3 ? 3 : 5;
  }
};

template <int R, int Ns, int N, int DIR, int EPT, int LGH>
__device__ __forceinline__ void fft_pass_c(float2 *buf, const float2 *__restrict__ tw, int tid,
                                           int nt) {
  constexpr int MAXB = EPT / R;
  constexpr int H = 1 << LGH;
  constexpr int nbPer = N / R;
  constexpr int nb = nbPer << LGH;
  constexpr int twstep = N / (Ns * R);
  static_assert(MAXB >= 1, "EPT must cover the radix");
  static_assert(nbPer % 16 == 0 && (Ns == 1 || Ns % 16 == 0), "constant-offset layout");
  constexpr int rstride = (nbPer + nbPer / 16) << LGH;
  constexpr int wstride = Ns == 1 ? H : (Ns + Ns / 16) << LGH;
  float2 v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> LGH;
      const int base = (lds_pad(j) << LGH) + t;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[base + r * rstride];
      if constexpr (Ns > 1) twiddle<R, DIR>(v[q], tw, (unsigned)((j % Ns) * twstep));
      butterfly<R, DIR>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> LGH;
      const int k = j % Ns;
      const int base = (lds_pad((j - k) * R + k) << LGH) + t;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * wstride] = v[q][out_slot<R>(r)];
    }
  }
  __syncthreads();
}

// lds_fft for a compile-time size N (StaticPlan); `tw` is the twiddle table
// of size N.  Same layout, requirements and barriers as lds_fft.
// Passes S .. SEnd-1 only (Ns: the sub-transform length entering pass S).
template <int N, int DIR, int EPT, int LGH, int S = 0, int Ns = 1,
          int SEnd = StaticPlan<N>::passes>
__device__ __forceinline__ void lds_fft_c(float2 *buf, const float2 *__restrict__ tw, int tid,
                                          int nt) {
  if constexpr (S < SEnd) {
    constexpr int R = StaticPlan<N>::radix(S);
    fft_pass_c<R, Ns, N, DIR, EPT, LGH>(buf, tw, tid, nt);
    lds_fft_c<N, DIR, EPT, LGH, S + 1, Ns * R, SEnd>(buf, tw, tid, nt);
  }
}

// ---- compile-time transforms with global input and output --------------
// fft_c_global runs the transform of lds_fft_c, but its first pass takes its
// inputs from ld(i, t) (element i of interleaved sequence t, straight into
// registers) and its last pass hands each output to st(i, t, value) instead
// of writing it to LDS: two LDS sweeps and four barriers fewer than staging
// the input in LDS and reading the result back.  The caller must not use
// `buf` concurrently (it is written without a barrier first) and gets no
// barrier at the end.  Requires at least two passes.

// Pass 0 (Ns = 1, no twiddles) from ld into LDS.
template <int R, int N, int DIR, int EPT, int LGH, typename Ld>
__device__ __forceinline__ void fft_pass_c_first(float2 *buf, int tid, int nt, Ld ld) {
  constexpr int MAXB = EPT / R;
  constexpr int H = 1 << LGH;
  constexpr int nbPer = N / R;
  constexpr int nb = nbPer << LGH;
  static_assert(MAXB >= 1 && nbPer % 16 == 0, "layout");
  float2 v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> LGH;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = ld(j + r * nbPer, t);
    }
  }
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> LGH;
      butterfly<R, DIR>(v[q]);
      const int base = (lds_pad(j * R) << LGH) + t;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * H] = v[q][out_slot<R>(r)];
    }
  }
}

// The last pass (sub-transform length Ns) from LDS to st.
template <int R, int Ns, int N, int DIR, int EPT, int LGH, typename St>
__device__ __forceinline__ void fft_pass_c_last(const float2 *buf, const float2 *__restrict__ tw,
                                                int tid, int nt, St st) {
  constexpr int MAXB = EPT / R;
  constexpr int H = 1 << LGH;
  constexpr int nbPer = N / R;
  constexpr int nb = nbPer << LGH;
  constexpr int twstep = N / (Ns * R);
  static_assert(MAXB >= 1 && nbPer % 16 == 0 && Ns > 1 && Ns % 16 == 0, "layout");
  constexpr int rstride = (nbPer + nbPer / 16) << LGH;
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int g = tid + q * nt;
    if (g < nb) {
      const int t = g & (H - 1), j = g >> LGH;
      const int base = (lds_pad(j) << LGH) + t;
      float2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = buf[base + r * rstride];
      const int k = j % Ns;
      twiddle<R, DIR>(v, tw, (unsigned)(k * twstep));
      butterfly<R, DIR>(v);
      const int o = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) st(o + r * Ns, t, v[out_slot<R>(r)]);
    }
  }
}

// First pass from ld, the rest in LDS (ends with a barrier, result in buf).
template <int N, int DIR, int EPT, int LGH, typename Ld>
__device__ __forceinline__ void fft_c_from_global(float2 *buf, const float2 *__restrict__ tw,
                                                  int tid, int nt, Ld ld) {
  constexpr int R0 = StaticPlan<N>::radix(0);
  fft_pass_c_first<R0, N, DIR, EPT, LGH>(buf, tid, nt, ld);
  __syncthreads();
  lds_fft_c<N, DIR, EPT, LGH, 1, R0>(buf, tw, tid, nt);
}

template <int N, int DIR, int EPT, int LGH, typename Ld, typename St>
__device__ __forceinline__ void fft_c_global(float2 *buf, const float2 *__restrict__ tw, int tid,
                                             int nt, Ld ld, St st) {
  using SP = StaticPlan<N>;
  static_assert(SP::passes >= 2, "two passes at least");
  constexpr int R0 = SP::radix(0);
  constexpr int RL = SP::radix(SP::passes - 1);
  fft_pass_c_first<R0, N, DIR, EPT, LGH>(buf, tid, nt, ld);
  __syncthreads();
  lds_fft_c<N, DIR, EPT, LGH, 1, R0, SP::passes - 1>(buf, tw, tid, nt);
  fft_pass_c_last<RL, N / RL, N, DIR, EPT, LGH>(buf, tw, tid, nt, st);
}

// Calls f(std::integral_constant<int, N>) for the sizes with a compiled
// specialisation (2^k; 3, 9, 5 and 15 times a power of two: the detector
// widths 2^k, 3 2^k, 1152/2304, 1280/2560 and 1920), f(std::integral_constant
// <int, 0>) otherwise.
template <typename F>
inline void with_static_size(int N, F &&f) {
  switch (N) {
    case 512: f(std::integral_constant<int, 512>()); break;
    case 768: f(std::integral_constant<int, 768>()); break;
    case 1280: f(std::integral_constant<int, 1280>()); break;
    case 1536: f(std::integral_constant<int, 1536>()); break;
    case 2304: f(std::integral_constant<int, 2304>()); break;
    case 2560: f(std::integral_constant<int, 2560>()); break;
    case 3072: f(std::integral_constant<int, 3072>()); break;
    case 3840: f(std::integral_constant<int, 3840>()); break;
    case 4608: f(std::integral_constant<int, 4608>()); break;
    case 5120: f(std::integral_constant<int, 5120>()); break;
    case 6144: f(std::integral_constant<int, 6144>()); break;
    case 1024: f(std::integral_constant<int, 1024>()); break;
    case 2048: f(std::integral_constant<int, 2048>()); break;
    case 4096: f(std::integral_constant<int, 4096>()); break;
    case 8192: f(std::integral_constant<int, 8192>()); break;
    case -1: f(std::integral_constant<int, -1>()); break;  // run-time sizes, fft_level 1
    case -2: f(std::integral_constant<int, -2>()); break;  // run-time sizes, fft_level 2
    default: f(std::integral_constant<int, 0>()); break;
  }
}

// True if N has a compiled specialisation.
inline bool static_fft_size(int N) {
  bool ok = false;
  with_static_size(N, [&](auto ns) { ok = decltype(ns)::value > 0; });
  return ok;
}

// Maps a flattened workgroup id to a logical one such that the workgroups the
// dispatcher sends to one XCD (round robin over 8 XCDs) receive consecutive
// logical ids: neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int id, int total) {
  const int q = total / 8, r = total % 8;
  const int x = id % 8, i = id / 8;
  return x * q + min(x, r) + i;
}

// Copies `count` items into LDS with U global loads of each thread in flight
// at once (all loads of a round are issued before the first LDS store, so
// the round costs one memory latency rather than U): item i is produced by
// ld(i) and written by st(i, value).
template <int U, typename T, typename Ld, typename St>
__device__ __forceinline__ void stage_lds(int count, int tid, int nt, Ld ld, St st) {
  for (int base = tid; base < count; base += U * nt) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * nt;
      if (i < count) v[u] = ld(i);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * nt;
      if (i < count) st(i, v[u]);
    }
  }
}

// Workgroup size for transforming `elems` complex values (all interleaved
// sequences) with plan d: 16 values per thread for radices 2/4/8/16, fewer
// with a radix-3, 5, 7, 11 or 13 pass (15, 15, 14, 11, 13: whole butterflies;
// the other prime passes take any count); at least 256 and a multiple of 64.  Returns -1
// if more than 1024 threads would be needed.
inline int fft_threads(const FftDesc &d, int64_t elems, int ept = kFftEPT) {
  int64_t per = ept;  // values a thread covers in every pass: whole butterflies
  for (int i = 0; i < d.nr; ++i) {
    const int R = d.r[i];
    // whole butterflies; a register-butterfly prime pass above ept runs one
    // per thread (nt >= N/16 > N/R butterflies), the direct-sum pass any count
    if (R <= 8 || R == 16 || (R <= 17 && R <= ept)) per = std::min<int64_t>(per, (ept / R) * R);
  }
  int64_t t = (elems + per - 1) / per;
  t = ((t + 63) / 64) * 64;
  if (t > 1024) return -1;
  return (int)(t < 256 ? 256 : t);
}

// Transform level of a length N (the LV of lds_fft): 0 if N = 2^a 3^b, 1 if
// its other prime factors are at most 17, 2 otherwise.
inline int fft_level(int64_t N) {
  if (N < 1) return 0;
  while (N % 2 == 0) N /= 2;
  while (N % 3 == 0) N /= 3;
  if (N == 1) return 0;
  for (int64_t p : {5, 7, 11, 13, 17})
    while (N % p == 0) N /= p;
  return N == 1 ? 1 : 2;
}

// True if N has a prime factor other than 2 and 3 (its plans need level 1 or
// 2 transforms).
inline bool fft_general(int64_t N) { return fft_level(N) > 0; }

// f(std::integral_constant<int, L>()) for the transform level L in 0..2.
template <typename F>
inline void with_fft_level(int L, F &&f) {
  if (L == 2)
    f(std::integral_constant<int, 2>());
  else if (L == 1)
    f(std::integral_constant<int, 1>());
  else
    f(std::integral_constant<int, 0>());
}

// lds_fft level of the run-time-size kernels NS = -1 and NS = -2.
constexpr int fft_lv(int NS) { return NS == -2 ? 2 : (NS == -1 ? 1 : 0); }

// True if N >= 1 and every prime factor of N is at most kFftMaxPrime.
inline bool fft_size_ok(int64_t N) {
  if (N < 1) return false;
  for (int64_t p = 2; p <= kFftMaxPrime; ++p)
    while (N % p == 0) N /= p;
  return N == 1;
}

}  // namespace tomo
