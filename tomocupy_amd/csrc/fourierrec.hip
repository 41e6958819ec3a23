// USFFT backprojector (tomocupy cfunc_fourierrec) and its adjoint for gfx950.
//
// Data flow of fr_backprojection for one call (nzc = ceil(nz/2) complex
// slices; the gridding treats them in groups of 4, one MFMA column pair each):
//   k_fr_prep    data (nz,nproj,n) -> pair-pack, shifted n-point DFT,
//                scale (4/n)*c0            -> G [c][p*n+k] complex
//   k_fr_grid    deterministic gather of the polar samples onto the
//                periodic 2n x 2n grid as 16x16-tile matrix products on the
//                matrix cores (tile sample lists built on the host in FP64),
//                + k_fr_grid_reduce for tiles split over several waves
//                                              -> F [c][a][b]
//   k_fr_rowfft  shifted inverse FFT along b, pruned to the n centre columns
//                -> R [c][a][v']
//   k_fr_colfft  shifted inverse FFT along a, pruned to the n rows divphi
//                reads, deapodisation (FP64 tables), circle mask, unpacking
//                -> obj (nz,n,n) f32
// which is, step for step, cfunc_fourierrec::backprojection
// (src/cuda/cfunc_fourierrec.cu:67-93): ifftshiftc/FFT1D/ifftshiftc (:78-80),
// gather (:82) + wrap (:84) as one periodic gather, fftshiftc/FFT2D^-1/
// fftshiftc (:86-89) as two pruned 1D passes, divphi (:91), circ (:92).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "common.h"
#include "fft_lds.h"
#include "internal.h"
#include "precision.h"

using namespace tomo;

namespace {

constexpr int kGridWaves = 4;   // waves (= work items) per gridding workgroup
constexpr int kGridChunk = 512; // default max list entries per work item (multiple of 16)
constexpr int kGridBatch = 64;  // list entries a gridding wave stages at a time
// device memory for the interpolation's tile partials (fr_forward): as many
// groups of 4 complex slices per launch as fit
constexpr int64_t kInterpBytes = (int64_t)512 << 20;
constexpr int kInterpPitch = 17;  // LDS row pitch (float4) of the interpolation tile
// The four 16-lane groups of a gridding wave read staged entries 8 apart in
// the same instruction; one pad int4 (entries) or two pad float2 (spectra)
// after every 8 put the four addresses on different LDS banks.
constexpr int kSeLen = kGridBatch + kGridBatch / 8;
constexpr int kSgLen = kGridBatch + kGridBatch / 4;
__device__ __forceinline__ int se_at(int e) { return e + (e >> 3); }
__device__ __forceinline__ int sg_at(int e) { return e + 2 * (e >> 3); }
constexpr int kColEPT = 32;     // FFT values per thread in the column pass
// log2 of the columns a column-pass workgroup transforms for 2n = N: as many
// as fit 512 threads at kColEPT values each (and the LDS), at most 16
constexpr int col_lg(int N) { return ilog2c(16384 / N) > 4 ? 4 : ilog2c(16384 / N); }

// Sign 1 - 2*((t+1) % 2): -1 for even t, +1 for odd t (kernels_fourierrec.cuh:121-132).
__device__ __forceinline__ float shift_sign(int t) { return (t & 1) ? 1.f : -1.f; }

// NS: n as a compile-time constant (0 = sizes known at run time only, -1 and
// -2 the same for sizes of fft_level 1 and 2), as in
// every FFT kernel of this file.
// max over the workgroup of the threads' m; thread 0 gets the result
// (LDS for the per-wave maxima; every thread must call it)
__device__ __forceinline__ float block_max(float m) {
  __shared__ float wmax[16];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wmax[w] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 1; q < nw; ++q) m = fmaxf(m, wmax[q]);
  return m;
}

// Sum over the workgroup of the threads' v; thread 0 gets the result (every
// thread must call it).
__device__ __forceinline__ float2 block_sum2(float2 v) {
  __shared__ float2 wsum[16];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wsum[w] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 1; q < nw; ++q) v = cadd(v, wsum[q]);
  return v;
}

// Writes the shifted spectrum row out[k] = sm[k] * (-1)^k * scale and, for
// float32 plans with a gmax buffer, the row's largest |re|, |im| to gmax[row]
// (the gridding scales its split half-precision factors with the maximum).
template <typename T>
__device__ __forceinline__ void store_spectrum(cpx<T> *out, const float2 *sm, int n, float scale,
                                               int tid, int nt, float *gmax, size_t row) {
  float m = 0.f;
  for (int k = tid; k < n; k += nt) {
    const float2 v = cscale(sm[lds_pad(k)], shift_sign(k) * scale);
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    stc(out + k, v);
  }
  if constexpr (std::is_same<T, float>::value) {
    if (gmax) {
      m = block_max(m);
      if (tid == 0) gmax[row] = m;
    }
  }
}

// The gmax part of store_spectrum for a thread's running maximum m (every
// thread must call it).
template <typename T>
__device__ __forceinline__ void store_max(float m, float *gmax, size_t row) {
  if constexpr (std::is_same<T, float>::value) {
    if (gmax) {
      m = block_max(m);
      if (threadIdx.x == 0) gmax[row] = m;
    }
  }
}

template <typename T, int NS>
__global__ __launch_bounds__(512) void k_fr_prep(const T *__restrict__ data, cpx<T> *__restrict__ G,
                                                 FftDesc fd, int n, int nproj, int nz, int nzc,
                                                 float scale, float *__restrict__ gmax) {
  extern __shared__ float2 sm[];
  const int p = blockIdx.x, c = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  if (c >= nzc) return;
  cpx<T> *out = G + ((size_t)c * nproj + p) * n;
  const T *r1 = data + ((size_t)c * nproj + p) * n;
  const int z2 = c + nzc;
  const T *r2 = data + ((size_t)(z2 < nz ? z2 : c) * nproj + p) * n;
  const bool has2 = z2 < nz;
  if constexpr (NS > 0) {
    float m = 0.f;
    fft_c_global<NS, -1, kFftEPT, 0>(
        sm, fd.tw, tid, nt,
        [&](int j, int) {
          return cscale(tof2(cpx<T>{r1[j], has2 ? r2[j] : T(0)}), shift_sign(j));
        },
        [&](int k, int, float2 x) {
          const float2 v = cscale(x, shift_sign(k) * scale);
          m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
          stc(out + k, v);
        });
    store_max<T>(m, gmax, (size_t)c * nproj + p);
  } else {
    stage_lds<8, cpx<T>>(
        n, tid, nt, [&](int j) { return cpx<T>{r1[j], has2 ? r2[j] : T(0)}; },
        [&](int j, cpx<T> v) { sm[lds_pad(j)] = cscale(tof2(v), shift_sign(j)); });
    __syncthreads();
    lds_fft<-1, kFftEPT, fft_lv(NS)>(sm, fd, 0, tid, nt);
    store_spectrum(out, sm, n, scale, tid, nt, gmax, (size_t)c * nproj + p);
  }
}

// Fused FBP filter + pair packing + shifted detector FFT (the fused form of
// fbp_filter_center and the first stage of the backprojection).  Per angle p
// and complex slice c: z = x1 + i*x2 zero-padded to 2n, FFT_2n, multiply by
// the two slices' wrapped filter spectra (Y = Z (H1+H2)/2 + conj(Z(-k))
// (H1-H2)/2, exact for real rows), IFFT_2n, add the edge-replication terms
// x[0] HL + x[n-1] HR per slice, then the shifted n-point DFT of k_fr_prep.
template <typename T, int NS>
__global__ __launch_bounds__(512) void k_fr_filtprep(
    const T *__restrict__ data, cpx<T> *__restrict__ G, const float2 *__restrict__ H2,
    const float *__restrict__ HL, const float *__restrict__ HR, FftDesc fd2, FftDesc fd1, int n,
    int nproj, int nz, int nzc, float scale, float *__restrict__ gmax) {
  extern __shared__ float2 sm[];
  __shared__ float edge[4];
  const int p = blockIdx.x, c = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  if (c >= nzc) return;
  cpx<T> *out = G + ((size_t)c * nproj + p) * n;
  const int n2 = 2 * n;
  const int z2 = c + nzc;
  const bool has2 = z2 < nz;
  const int zb = has2 ? z2 : c;
  const T *r1 = data + ((size_t)c * nproj + p) * n;
  const T *r2 = data + ((size_t)zb * nproj + p) * n;
  if (tid == 0) {
    edge[0] = ldr(r1);
    edge[1] = ldr(r1 + n - 1);
    edge[2] = has2 ? ldr(r2) : 0.f;
    edge[3] = has2 ? ldr(r2 + n - 1) : 0.f;
  }
  // High-pass filters (|H2[0]| under 1 % of |H2[n/2]| for both slices, the
  // ramp family) transform the differenced row d[j] = z[j] - z[j-1] with the
  // spectra H2 / (1 - W^k) (k_flt_H2_64) and the DC term from the row sum:
  // the float32 rounding of the transform is then relative to the small
  // differences of a smooth row rather than to its large low frequencies,
  // which the filter discards (at n = 2048, 24 angles, shepp: 1.1e-5 -> about
  // 1e-6 rel-L2 against the FP64 filter).
  // (float32 plans only: 16-bit storage rounds far above this error)
  const float2 *H2c = H2 + (size_t)c * n2, *H2b = H2 + (size_t)zb * n2;
  const bool diff = std::is_same<T, float>::value &&
                    cabs2(H2c[0]) <= 1e-4f * cabs2(H2c[n / 2]) &&
                    cabs2(H2b[0]) <= 1e-4f * cabs2(H2b[n / 2]);
  auto zat = [&](int j) {  // z[j], zero outside [0, n)
    return (unsigned)j < (unsigned)n ? tof2(cpx<T>{r1[j], has2 ? r2[j] : T(0)})
                                     : make_float2(0.f, 0.f);
  };
  float2 zsum = make_float2(0.f, 0.f);  // this thread's part of the row sum
  auto zdiff = [&](int j) {
    const float2 zj = zat(j);
    zsum = cadd(zsum, zj);
    return csub(zj, zat(j - 1));
  };
  // (one loader per branch: a branch inside the loader would keep the loads
  // of a thread from being issued together)
  if constexpr (NS > 0) {
    // (the barriers inside make edge[] visible before its use below)
    if (diff)
      fft_c_from_global<2 * NS, -1, kFftEPT, 0>(sm, fd2.tw, tid, nt,
                                                [&](int j, int) { return zdiff(j); });
    else
      fft_c_from_global<2 * NS, -1, kFftEPT, 0>(sm, fd2.tw, tid, nt,
                                                [&](int j, int) { return zat(j); });
  } else {
    auto st = [&](int j, float2 v) { sm[lds_pad(j)] = v; };
    if (diff)
      stage_lds<8, float2>(n2, tid, nt, [&](int j) { return zdiff(j); }, st);
    else
      stage_lds<8, float2>(n2, tid, nt, [&](int j) { return zat(j); }, st);
    __syncthreads();
    lds_fft<-1, kFftEPT, fft_lv(NS)>(sm, fd2, 0, tid, nt);
  }
  float2 z0 = make_float2(0.f, 0.f);  // Z[0] (thread 0), for the differenced form
  if (diff) z0 = block_sum2(zsum);
  const float2 *Ha = diff ? H2c + (size_t)nz * n2 : H2c;
  const float2 *Hb = diff ? H2b + (size_t)nz * n2 : H2b;
  for (int k = tid; k <= n; k += nt) {
    const int km = (n2 - k) % n2;
    float2 zk = sm[lds_pad(k)], zm = sm[lds_pad(km)];
    if (diff && k == 0) zk = zm = z0;
    const float2 ak = Ha[k], bk = Hb[k], am = Ha[km], bm = Hb[km];
    const float2 sk = cscale(cadd(ak, bk), 0.5f), dk = cscale(csub(ak, bk), 0.5f);
    const float2 sm_ = cscale(cadd(am, bm), 0.5f), dm = cscale(csub(am, bm), 0.5f);
    const float2 yk = cadd(cmul(zk, sk), cmul(cconj(zm), dk));
    const float2 ym = cadd(cmul(zm, sm_), cmul(cconj(zk), dm));
    sm[lds_pad(k)] = yk;
    if (km != k) sm[lds_pad(km)] = ym;
  }
  __syncthreads();
  if constexpr (NS > 0)
    lds_fft_c<2 * NS, 1, kFftEPT, 0>(sm, fd2.tw, tid, nt);
  else
    lds_fft<1, kFftEPT, fft_lv(NS)>(sm, fd2, 0, tid, nt);
  const float *hl1 = HL + (size_t)c * n, *hr1 = HR + (size_t)c * n;
  const float *hl2 = HL + (size_t)zb * n, *hr2 = HR + (size_t)zb * n;
  const float e0 = edge[0], e1 = edge[1], e2 = edge[2], e3 = edge[3];
  for (int j = tid; j < n; j += nt) {
    float2 y = sm[lds_pad(j)];
    y.x += e0 * hl1[j] + e1 * hr1[j];
    y.y += e2 * hl2[j] + e3 * hr2[j];
    sm[lds_pad(j)] = cscale(y, shift_sign(j));
  }
  __syncthreads();
  if constexpr (NS > 0) {
    // the n-point pass from LDS, its last stage straight to the spectrum
    constexpr int RL = StaticPlan<NS>::radix(StaticPlan<NS>::passes - 1);
    lds_fft_c<NS, -1, kFftEPT, 0, 0, 1, StaticPlan<NS>::passes - 1>(sm, fd1.tw, tid, nt);
    float m = 0.f;
    fft_pass_c_last<RL, NS / RL, NS, -1, kFftEPT, 0>(sm, fd1.tw, tid, nt,
                                                     [&](int k, int, float2 x) {
      const float2 v = cscale(x, shift_sign(k) * scale);
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
      stc(out + k, v);
    });
    store_max<T>(m, gmax, (size_t)c * nproj + p);
  } else {
    lds_fft<-1, kFftEPT, fft_lv(NS)>(sm, fd1, 0, tid, nt);
    store_spectrum(out, sm, n, scale, tid, nt, gmax, (size_t)c * nproj + p);
  }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Gaussian factor of one list entry for tile column (or row) i.  Entries are
// {sample, footprint starts, du bits, dv bits}; the footprint starts are two
// halves (columns low, rows high) holding b = start + 8, the start of the
// 9-cell footprint relative to the tile plus 8, in [0, 24); padding entries
// (sample -1) hold b = 255, outside every tile.  With x0 = i - start - 4 - du
// the tap offset, x = x0 + 0.5 - kGaussDelta = (fiA - b) - du with the lane
// constant fiA = i + 4.5 - kGaussDelta, and the footprint (start <= i <=
// start + 8, i.e. x0 in (-5, 4]) is |x| <= 4.5 - kGaussDelta up to taps with
// x0 in (-5, -5 + 2 kGaussDelta), whose weights (< 2e-8 of the centre tap)
// are dropped.  The exponent K2 x0^2 + s is evaluated as
// fma(fma(K2, x, m2), x, cs) with m2 = -2 K2 cB, cs = K2 cB^2 + s,
// cB = 0.5 - kGaussDelta.  K2 x0^2 >= -26 inside the footprint, far from
// where the hardware exponential flushes, so v_exp_f32 is used directly.
constexpr float kGaussDelta = 1.f / 1024.f;
constexpr float kGaussLim = 4.5f - kGaussDelta;
constexpr float kGaussCB = 0.5f - kGaussDelta;

struct GaussLane {
  float fiA, K2, m2, cs0, cs1;
  __device__ GaussLane(int i, float k2, float s0, float s1) {
    fiA = (float)i + 4.5f - kGaussDelta;
    K2 = k2;
    m2 = -2.f * k2 * kGaussCB;
    const double q = (double)k2 * (double)kGaussCB * (double)kGaussCB;
    cs0 = (float)(q + (double)s0);
    cs1 = (float)(q + (double)s1);
  }
  // fiA - b with b the low (HI = 0) or high (HI = 1) half of bb, widened
  // inside one fused multiply-add (v_fma_mix_f32)
  template <int HI>
  __device__ __forceinline__ float offset(int bb) const {
    float r;
    if constexpr (HI)
      asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
          : "=v"(r)
          : "v"(bb), "v"(fiA));
    else
      asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(bb), "v"(fiA));
    return r;
  }
  __device__ __forceinline__ float one(float fb, float du, float cs) const {
    const float x = fb - du;
    const float arg = fmaf(fmaf(K2, x, m2), x, cs);
    return fabsf(x) <= kGaussLim ? __builtin_amdgcn_exp2f(arg) : 0.f;
  }
  // e0: column factor (times 2^s0), e1: row factor (times 2^s1)
  __device__ __forceinline__ void pair(const int4 &e, float &e0, float &e1) const {
    e0 = one(offset<0>(e.y), __int_as_float(e.z), cs0);
    e1 = one(offset<1>(e.y), __int_as_float(e.w), cs1);
  }
};

// Footprint-start field of a list entry (see GaussLane).
inline int pack_starts(int col_start, int row_start) {
  const _Float16 c = (_Float16)(float)(col_start + 8), r = (_Float16)(float)(row_start + 8);
  unsigned short uc, ur;
  std::memcpy(&uc, &c, 2);
  std::memcpy(&ur, &r, 2);
  return (int)((unsigned)uc | ((unsigned)ur << 16));
}
inline int pad_starts() { return pack_starts(255 - 8, 255 - 8); }
constexpr int kPadStarts = 0x5BF85BF8;  // pad_starts(): 255.0 in both halves

// Gridding as a matrix product on the matrix cores.  For one 16x16 cell tile
// (rows a, columns b of the periodic 2n x 2n grid) and a batch of polar
// samples s:
//   F[a][b] += sum_s (E1_s(a) * G_s) * E0_s(b)
// with the separable Gaussian factors E1 (rows) and E0 (columns): A = E1*G
// (16 x K), B = E0 (K x 16), one matrix instruction per slice and re/im part.
// A wave walks one work item (a tile's sample list or a bounded chunk of it)
// in a fixed order; chunks of heavy tiles land in private partial buffers
// that k_fr_grid_reduce sums in chunk order, so the result is bitwise
// reproducible.
// 16-bit plans: the same tile products with 32 samples per matrix
// instruction (v_mfma_f32_16x16x32 f16 / bf16, FP32 accumulation).  Lane l
// supplies the 8 K entries of row l%16 of A and of column l%16 of B that
// belong to samples 8(l/16)..8(l/16)+7 of each batch of 32 (the same samples
// for both operands, so the order of K inside the instruction does not
// matter); the factors and the weighted spectra are rounded to the storage
// type, the sums stay in FP32.  Tile lists are padded to multiples of 32.
template <typename T>
struct Mma16;
template <>
struct Mma16<_Float16> {
  typedef _Float16 v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 pack(float a, float b, float c, float d) {
    return v4{(_Float16)a, (_Float16)b, (_Float16)c, (_Float16)d};
  }
  static __device__ __forceinline__ floatx4 mma(v4 a, v4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
  }
  typedef _Float16 v8 __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ v8 pack8(const float *x) {
    v8 r;
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = (_Float16)x[u];
    return r;
  }
  static __device__ __forceinline__ floatx4 mma8(v8 a, v8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mma16<__bf16> {
  typedef short v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ short bits(float x) {
    return __builtin_bit_cast(short, (__bf16)x);
  }
  static __device__ __forceinline__ v4 pack(float a, float b, float c, float d) {
    return v4{bits(a), bits(b), bits(c), bits(d)};
  }
  static __device__ __forceinline__ floatx4 mma(v4 a, v4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  typedef __bf16 v8 __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ v8 pack8(const float *x) {
    v8 r;
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = (__bf16)x[u];
    return r;
  }
  static __device__ __forceinline__ floatx4 mma8(v8 a, v8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

template <typename T>
__global__ __launch_bounds__(64 * kGridWaves) void k_fr_grid16(
    const int4 *__restrict__ items, int nitems, const int4 *__restrict__ ent,
    const cpx<T> *__restrict__ G, cpx<T> *__restrict__ F, float *__restrict__ part, int n,
    int ntile, int nslots, int nsamp, int nzc, float K2) {
  typedef Mma16<T> M;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int it = blockIdx.x * kGridWaves + wave;
  const int grp = blockIdx.y;
  if (it >= nitems) return;
  const int4 item = items[it];
  const int tile = item.x, ebeg = item.y, cnt = item.z, slot = item.w;
  const int g4 = lane >> 4, i = lane & 15;
  const int nlive = min(4, nzc - 4 * grp);
  const cpx<T> *Gg = G + (size_t)(4 * grp) * nsamp;
  const GaussLane gl(i, K2, 0.f, 0.f);
  floatx4 acc[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) acc[m] = floatx4{0.f, 0.f, 0.f, 0.f};
  __shared__ int4 se[kGridWaves][2][kSeLen];
  __shared__ float2 sg[kGridWaves][2][4][kSgLen];
  const int eend = ebeg + cnt;
  int4 re = make_int4(-1, kPadStarts, 0, 0);
  cpx<T> rg[4];
  auto fetch = [&](int base) {
    const int idx = base + lane;
    re = idx < eend ? ent[idx] : make_int4(-1, kPadStarts, 0, 0);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      rg[c] = (re.x >= 0 && c < nlive) ? Gg[(size_t)c * nsamp + re.x] : cpx<T>{T(0), T(0)};
  };
  if (cnt > 0) fetch(ebeg);
  int buf = 0;
  for (int bb = ebeg; bb < eend; bb += kGridBatch, buf ^= 1) {
    se[wave][buf][se_at(lane)] = re;
#pragma unroll
    for (int c = 0; c < 4; ++c) sg[wave][buf][c][sg_at(lane)] = tof2(rg[c]);
    __builtin_amdgcn_wave_barrier();
    if (bb + kGridBatch < eend) fetch(bb + kGridBatch);
    const int ngr = min(kGridBatch, eend - bb) >> 5;
    for (int q = 0; q < ngr; ++q) {
      // 32 samples per instruction: lane group g4 supplies samples
      // 32q + 8 g4 + u (u < 8) to both operands
      float e0v[8], e1v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) gl.pair(se[wave][buf][se_at(32 * q + 8 * g4 + u)], e0v[u], e1v[u]);
      const typename M::v8 b = M::pack8(e0v);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        float ar[8], ai[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const float2 v = sg[wave][buf][c][sg_at(32 * q + 8 * g4 + u)];
          ar[u] = e1v[u] * v.x;
          ai[u] = e1v[u] * v.y;
        }
        acc[2 * c] = M::mma8(M::pack8(ar), b, acc[2 * c]);
        acc[2 * c + 1] = M::mma8(M::pack8(ai), b, acc[2 * c + 1]);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const int j = lane & 15, ib = 4 * (lane >> 4);
  if (slot < 0) {
    const int n2 = 2 * n;
    const int TA = tile / ntile, TB = tile - TA * ntile;
    const int bcol = 16 * TB + j;
    const size_t plane = (size_t)n2 * n2;
    const bool whole = (n2 & 15) == 0;
    const int o0 = (16 * TA + ib) * n2 + bcol;
    if (whole && grp * 4 + 4 <= nzc) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        cpx<T> *Fc = F + (size_t)(grp * 4 + c) * plane;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          stc(Fc + (o0 + r * n2), make_float2(acc[2 * c][r], acc[2 * c + 1][r]));
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (grp * 4 + c >= nzc) continue;
        cpx<T> *Fc = F + (size_t)(grp * 4 + c) * plane;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int a = 16 * TA + ib + r;
          if (a < n2 && bcol < n2)
            stc(Fc + (o0 + r * n2), make_float2(acc[2 * c][r], acc[2 * c + 1][r]));
        }
      }
    }
  } else {
    float *P = part + ((size_t)grp * nslots + slot) * 2048;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) P[m * 256 + (ib + r) * 16 + j] = acc[m][r];
  }
}

// Scale exponent of the float32 gridding: kA = 13 - floor(log2(max |G|))
// over the per-row maxima of the spectra, so every weighted spectrum value
// times 2^kA stays below 2^14 (0 for an all-zero chunk).
__global__ __launch_bounds__(1024) void k_fr_gscale(const float *__restrict__ gmax, int count,
                                                    int *__restrict__ ksc) {
  float m = 0.f;
#pragma unroll 4
  for (int i = threadIdx.x; i < count; i += 1024) m = fmaxf(m, gmax[i]);
  m = block_max(m);
  if (threadIdx.x == 0) {
    int k = 0;
    if (m > 0.f && m <= 3.4e38f) k = 13 - ilogbf(m);
    ksc[0] = max(-120, min(120, k));
  }
}

// The float32 gridding on the half-precision matrix cores.  Same tile
// products and lane layout as k_fr_grid16 (32 samples per instruction), with
// each float32 factor split into half-precision high and low parts,
// x = hi + lo (hi: x with its 13 low mantissa bits cleared, exact in half
// precision; lo = x - hi rounded toward zero), and three products
// hi*hi + hi*lo + lo*hi per tile product (the dropped lo*lo term and the
// rounding of lo are 2^-20 relative).  To keep both parts in half precision's
// normal range the row factors carry 2^kA (k_fr_gscale: the weighted spectra
// stay below 2^14) and the column factors 2^10, folded into the exponents of
// the Gaussians; the accumulators are scaled back by 2^-(kA+10) (exact) on
// store.  On gfx950 the float32 matrix instruction costs as much as 2.4
// half-precision ones and takes a quarter of the samples, and matrix and
// vector work do not overlap.

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

// Two halves {a0*b0 - h.lo, a1*b1 - h.hi}: fused multiply-adds with the
// half operand widened and the result rounded to half (v_fma_mixlo_f16 writes
// the low half of the destination, v_fma_mixhi_f16 the high half).
__device__ __forceinline__ unsigned mix_lo2(float a0, float b0, float a1, float b1, unsigned h) {
  unsigned r;
  asm("v_fma_mixlo_f16 %0, %1, %2, -%3 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %0, %4, %5, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
      : "=&v"(r)
      : "v"(a0), "v"(b0), "v"(h), "v"(a1), "v"(b1));
  return r;
}

// x = hi + lo for 8 float32 products x[u] = f[u] * g[u]: hi is the product
// rounded toward zero to half precision, lo = f*g - hi (exact product minus
// hi, rounded once to half).
__device__ __forceinline__ void split8_prod(const float *f, const float *g, half8_t &hi,
                                           half8_t &lo) {
  unsigned h[4], l[4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    h[u] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(f[2 * u] * g[2 * u],
                                                                  f[2 * u + 1] * g[2 * u + 1]));
#pragma unroll
  for (int u = 0; u < 4; ++u) l[u] = mix_lo2(f[2 * u], g[2 * u], f[2 * u + 1], g[2 * u + 1], h[u]);
  hi = __builtin_bit_cast(half8_t, make_uint4(h[0], h[1], h[2], h[3]));
  lo = __builtin_bit_cast(half8_t, make_uint4(l[0], l[1], l[2], l[3]));
}

// The same split of 8 float32 values.
__device__ __forceinline__ void split8(const float *x, half8_t &hi, half8_t &lo) {
  unsigned h[4], l[4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    h[u] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(x[2 * u], x[2 * u + 1]));
#pragma unroll
  for (int u = 0; u < 4; ++u) l[u] = mix_lo2(x[2 * u], 1.f, x[2 * u + 1], 1.f, h[u]);
  hi = __builtin_bit_cast(half8_t, make_uint4(h[0], h[1], h[2], h[3]));
  lo = __builtin_bit_cast(half8_t, make_uint4(l[0], l[1], l[2], l[3]));
}

// One work item of the float32 gridding for one group of (up to) 4 complex
// slices: Gg/Fg/Pg are the group's spectra, grid planes and partial tiles,
// nlive of its 4 slices exist, kA is the group's scale exponent.
__device__ __forceinline__ void grid_x3_item(const int4 item, const int4 *__restrict__ ent,
                                             const float2 *__restrict__ Gg,
                                             float2 *__restrict__ Fg, float *__restrict__ Pg,
                                             int kA, int n, int ntile, int nsamp, int nlive,
                                             float K2, int wave, int lane) {
  const int tile = item.x, ebeg = item.y, cnt = item.z, slot = item.w;
  const int g4 = lane >> 4, i = lane & 15;
  const GaussLane gl(i, K2, 10.f, (float)kA);
  floatx4 acc[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) acc[m] = floatx4{0.f, 0.f, 0.f, 0.f};
  __shared__ int4 se[kGridWaves][2][kSeLen];
  __shared__ float2 sg[kGridWaves][2][4][kSgLen];
  const int eend = ebeg + cnt;
  int4 re = make_int4(-1, kPadStarts, 0, 0);
  float2 rg[4];
  auto fetch = [&](int base) {
    const int idx = base + lane;
    re = idx < eend ? ent[idx] : make_int4(-1, kPadStarts, 0, 0);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      rg[c] = (re.x >= 0 && c < nlive) ? Gg[(size_t)c * nsamp + re.x] : make_float2(0.f, 0.f);
  };
  if (cnt > 0) fetch(ebeg);
  int buf = 0;
  for (int bb = ebeg; bb < eend; bb += kGridBatch, buf ^= 1) {
    se[wave][buf][se_at(lane)] = re;
#pragma unroll
    for (int c = 0; c < 4; ++c) sg[wave][buf][c][sg_at(lane)] = rg[c];
    __builtin_amdgcn_wave_barrier();
    if (bb + kGridBatch < eend) fetch(bb + kGridBatch);
    const int ngr = min(kGridBatch, eend - bb) >> 5;
    for (int q = 0; q < ngr; ++q) {
      // lane group g4 supplies samples 32q + 8 g4 + u (u < 8) to both
      // operands, so the K order inside the instruction does not matter
      float e0v[8], e1v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) gl.pair(se[wave][buf][se_at(32 * q + 8 * g4 + u)], e0v[u], e1v[u]);
      half8_t bh, bl;
      split8(e0v, bh, bl);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        float vr[8], vi[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const float2 v = sg[wave][buf][c][sg_at(32 * q + 8 * g4 + u)];
          vr[u] = v.x;
          vi[u] = v.y;
        }
        half8_t ah, al;
        split8_prod(e1v, vr, ah, al);
        acc[2 * c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[2 * c], 0, 0, 0);
        acc[2 * c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[2 * c], 0, 0, 0);
        acc[2 * c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[2 * c], 0, 0, 0);
        split8_prod(e1v, vi, ah, al);
        acc[2 * c + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[2 * c + 1], 0, 0, 0);
        acc[2 * c + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[2 * c + 1], 0, 0, 0);
        acc[2 * c + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[2 * c + 1], 0, 0, 0);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const float unscale = ldexpf(1.f, -(kA + 10));
  const int j = lane & 15, ib = 4 * (lane >> 4);
  if (slot < 0) {
    const int n2 = 2 * n;
    const int TA = tile / ntile, TB = tile - TA * ntile;
    const int bcol = 16 * TB + j;
    const size_t plane = (size_t)n2 * n2;
    // whole tiles when 16 divides 2n: no bounds tests; 32-bit offsets from
    // each plane's base
    const bool whole = (n2 & 15) == 0;
    const int o0 = (16 * TA + ib) * n2 + bcol;
    if (whole && nlive == 4) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        float2 *Fc = Fg + (size_t)c * plane;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Fc[o0 + r * n2] = make_float2(acc[2 * c][r] * unscale, acc[2 * c + 1][r] * unscale);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= nlive) continue;
        float2 *Fc = Fg + (size_t)c * plane;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int a = 16 * TA + ib + r;
          if (a < n2 && bcol < n2)
            Fc[o0 + r * n2] = make_float2(acc[2 * c][r] * unscale, acc[2 * c + 1][r] * unscale);
        }
      }
    }
  } else {
    float *Pw = Pg + (size_t)slot * 2048;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) Pw[m * 256 + (ib + r) * 16 + j] = acc[m][r] * unscale;
  }
}

__global__ __launch_bounds__(64 * kGridWaves) void k_fr_grid_x3(
    const int4 *__restrict__ items, int nitems, const int4 *__restrict__ ent,
    const float2 *__restrict__ G, float2 *__restrict__ F, float *__restrict__ part,
    const int *__restrict__ ksc, int n, int ntile, int nslots, int nsamp, int nzc, float K2) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int it = blockIdx.x * kGridWaves + wave;
  const int grp = blockIdx.y;
  if (it >= nitems) return;
  const size_t plane = (size_t)(2 * n) * (2 * n);
  grid_x3_item(items[it], ent, G + (size_t)(4 * grp) * nsamp, F + (size_t)(4 * grp) * plane,
               part + (size_t)grp * nslots * 2048, ksc[0], n, ntile, nsamp,
               min(4, nzc - 4 * grp), K2, wave, lane);
}

// One group of 4 complex slices of one plan in a batched gridding launch.
struct GridJob {
  const int4 *ent;
  const float2 *G;  // the group's spectra
  float2 *F;        // the group's first grid plane
  float *part;      // the group's partial tiles
  const int *ksc;   // the plan's scale exponent
  int n, ntile, nsamp, nlive;
  float K2;
  int pad[3];
};

// The float32 gridding of several plans (groups) in one launch: wave w takes
// work item wi[w] of job wj[w] (the items of all jobs, longest first).
__global__ __launch_bounds__(64 * kGridWaves) void k_fr_grid_x3_batch(
    const GridJob *__restrict__ jobs, const int4 *__restrict__ wi, const int *__restrict__ wj,
    int nwork) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = blockIdx.x * kGridWaves + wave;
  if (w >= nwork) return;
  const GridJob jb = jobs[wj[w]];
  grid_x3_item(wi[w], jb.ent, jb.G, jb.F, jb.part, jb.ksc[0], jb.n, jb.ntile, jb.nsamp, jb.nlive,
               jb.K2, wave, lane);
}

// Sums the partial tiles of heavily loaded tiles in chunk order.
template <typename T>
__global__ void k_fr_grid_reduce(const int4 *__restrict__ red, const float *__restrict__ part,
                                 cpx<T> *__restrict__ F, int n, int ntile, int nslots, int nzc) {
  const int4 r = red[blockIdx.x];
  const int grp = blockIdx.y, c = blockIdx.z;  // one complex slice of the group
  if (grp * 4 + c >= nzc) return;
  const int cell = threadIdx.x;  // 256 threads = 16 x 16 cells
  const int n2 = 2 * n;
  const int TA = r.x / ntile, TB = r.x - TA * ntile;
  const int a = 16 * TA + (cell >> 4), b = 16 * TB + (cell & 15);
  float re = 0.f, im = 0.f;
  // (unrolled so that the loads of several slots are in flight together; the
  // additions stay in slot order)
  const float *P = part + ((size_t)grp * nslots + r.y) * 2048 + 2 * c * 256 + cell;
#pragma unroll 4
  for (int q = 0; q < r.z; ++q) {
    re += P[(size_t)q * 2048];
    im += P[(size_t)q * 2048 + 256];
  }
  if (a >= n2 || b >= n2) return;
  const size_t plane = (size_t)n2 * n2;
  stc(F + (size_t)(grp * 4 + c) * plane + (size_t)a * n2 + b, make_float2(re, im));
}

template <typename T, int NS>
// lgB > 0: R is stored in column blocks of 2^lgB, R[c][v'/B][a][v'%B], so that
// the column pass reads each of its blocks contiguously; the workgroups of
// neighbouring rows then share an XCD, whose L2 merges their partial-line
// writes.  lgB = 0: R[c][a][v'].
__global__ __launch_bounds__(512) void k_fr_rowfft(const cpx<T> *__restrict__ F,
                                                   cpx<T> *__restrict__ R, FftDesc fd, int n,
                                                   int nzc, int lgB,
                                                   const int2 *__restrict__ span) {
  extern __shared__ float2 sm[];
  const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int a = L % gridDim.x, c = L / gridDim.x;
  if (c >= nzc) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n2 = 2 * n;
  const cpx<T> *src = F + ((size_t)c * n2 + a) * n2;
  // only the row's occupied span was gridded; the rest of the row is zero
  const int2 sp = span[a];
  const int h = n / 2;
  // compiled sizes, float32: the first pass loads the row and the last one
  // stores R (measured at n = 2048: 161 -> 146 us per chunk of 8 slices; the
  // 16-bit column-blocked R went from 164 to 209 us that way, so 16-bit plans
  // stage through LDS)
  if constexpr (NS > 0 && std::is_same<T, float>::value) {
    fft_c_global<NS, 1, kFftEPT, 0>(
        sm, fd.tw, tid, nt,
        [&](int b, int) {
          const float2 x = (b >= sp.x && b < sp.y) ? tof2(src[b]) : make_float2(0.f, 0.f);
          return (b & 1) ? make_float2(-x.x, -x.y) : x;
        },
        [&](int o, int, float2 x) {
          const int v = o - h;
          if (v < 0 || v >= n) return;
          if (o & 1) x = make_float2(-x.x, -x.y);
          if (lgB > 0)
            stc(R + (size_t)c * n2 * n + (size_t)a * (1 << lgB) +
                    (size_t)(v >> lgB) * n2 * (1 << lgB) + (v & ((1 << lgB) - 1)),
                x);
          else
            stc(R + ((size_t)c * n2 + a) * n + v, x);
        });
  } else {
    stage_lds<8, cpx<T>>(
        n2, tid, nt,
        [&](int b) { return (b >= sp.x && b < sp.y) ? src[b] : cpx<T>{T(0), T(0)}; },
        [&](int b, cpx<T> v) {
          const float2 x = tof2(v);
          sm[lds_pad(b)] = (b & 1) ? make_float2(-x.x, -x.y) : x;
        });
    __syncthreads();
    if constexpr (NS > 0)
      lds_fft_c<NS, 1, kFftEPT, 0>(sm, fd.tw, tid, nt);
    else
      lds_fft<1, kFftEPT, fft_lv(NS)>(sm, fd, 0, tid, nt);
    if (lgB > 0) {
      const int B = 1 << lgB;
      cpx<T> *dst = R + (size_t)c * n2 * n + (size_t)a * B;
      for (int v = tid; v < n; v += nt) {
        const float2 x = sm[lds_pad(v + h)];
        stc(dst + (size_t)(v >> lgB) * n2 * B + (v & (B - 1)),
            ((v + h) & 1) ? make_float2(-x.x, -x.y) : x);
      }
    } else {
      cpx<T> *dst = R + ((size_t)c * n2 + a) * n;
      for (int v = tid; v < n; v += nt) {
        const float2 x = sm[lds_pad(v + h)];
        stc(dst + v, ((v + h) & 1) ? make_float2(-x.x, -x.y) : x);
      }
    }
  }
}

// The size-specialised column pass runs 16 values per thread (1024 threads):
// its LDS (4 columns of 2n) admits one workgroup per CU, so the extra waves
// are the only latency cover.
template <typename T, int NS>
__global__ __launch_bounds__(NS > 0 ? 1024 : 512) void k_fr_colfft(const cpx<T> *__restrict__ R, T *__restrict__ obj,
                            const float *__restrict__ phix, const float *__restrict__ phiy,
                            const int2 *__restrict__ circ, FftDesc fd, int n, int nz, int nzc,
                            int lgCB, int blocked) {
  extern __shared__ float2 sm[];  // [a][t]: CB interleaved columns
  // column blocks that share cache lines go to the same XCD
  const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int cb = L % gridDim.x, c = L / gridDim.x;
  if (c >= nzc) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int CB = 1 << lgCB;
  const int n2 = 2 * n, v0 = cb * CB;
  const cpx<T> *src = R + (size_t)c * n2 * n;
  const cpx<T> *blk = src + (size_t)v0 * n2;  // block cb of the column-blocked layout
  const int z2 = c + nzc;
  const int h = n / 2;
  // compiled sizes, float32: the first pass loads R and the last one stores
  // the slices (n = 2048: 205 -> 187 us per chunk of 8 slices; 16-bit plans,
  // 129 -> 145 us, stage through LDS)
  if constexpr (NS > 0 && std::is_same<T, float>::value) {
    constexpr int LG = col_lg(NS);  // == lgCB (the plan picks NS only then)
    fft_c_global<NS, 1, kFftEPT, LG>(
        sm, fd.tw, tid, nt,
        [&](int a, int t) {
          float2 x;
          if (blocked)
            x = tof2(blk[(a << LG) + t]);
          else
            x = (v0 + t < n) ? tof2(src[(size_t)a * n + v0 + t]) : make_float2(0.f, 0.f);
          return (a & 1) ? make_float2(-x.x, -x.y) : x;
        },
        [&](int u, int t, float2 x) {
          const int ty = u - 1 - h, tx = v0 + t;
          if (ty < 0 || ty >= n || tx >= n) return;
          const int2 ci = circ[ty];
          const float ph =
              (tx >= ci.x && tx < ci.y) ? phix[tx] * phiy[ty] * ((u & 1) ? -1.f : 1.f) : 0.f;
          str(obj + ((size_t)c * n + ty) * n + tx, x.x * ph);
          if (z2 < nz) str(obj + ((size_t)z2 * n + ty) * n + tx, x.y * ph);
        });
  } else {
    stage_lds<16, cpx<T>>(
        CB * n2, tid, nt,
        [&](int idx) {
          const int a = idx >> lgCB, t = idx & (CB - 1);
          if (blocked) return blk[idx];
          return (v0 + t < n) ? src[(size_t)a * n + v0 + t] : cpx<T>{T(0), T(0)};
        },
        [&](int idx, cpx<T> v) {
          const int a = idx >> lgCB, t = idx & (CB - 1);
          const float2 x = tof2(v);
          sm[(lds_pad(a) << lgCB) + t] = (a & 1) ? make_float2(-x.x, -x.y) : x;
        });
    __syncthreads();
    if constexpr (NS > 0)
      lds_fft_c<NS, 1, kFftEPT, col_lg(NS)>(sm, fd.tw, tid, nt);
    else
      lds_fft<1, (NS < 0 ? kFftEPT : kColEPT), fft_lv(NS)>(sm, fd, lgCB, tid, nt);
    for (int idx = tid; idx < CB * n; idx += nt) {
      const int ty = idx >> lgCB, t = idx & (CB - 1);
      const int tx = v0 + t;
      if (tx >= n) continue;
      const int u = ty + 1 + h;
      float2 x = sm[(lds_pad(u) << lgCB) + t];
      const int2 ci = circ[ty];
      const float ph = (tx >= ci.x && tx < ci.y) ? phix[tx] * phiy[ty] * ((u & 1) ? -1.f : 1.f) : 0.f;
      str(obj + ((size_t)c * n + ty) * n + tx, x.x * ph);
      if (z2 < nz) str(obj + ((size_t)z2 * n + ty) * n + tx, x.y * ph);
    }
  }
}

// ---- adjoint (forward projection) kernels --------------------------------

// NS as k_fr_colfft: a compiled 2n = 2^k or 3 * 2^k (up to 1024 threads,
// col_lg(NS) columns per workgroup) or 0 / -1 / -2 for the run-time plans of
// fft_level 0 / 1 / 2.
template <int NS>
__global__ __launch_bounds__(NS > 0 ? 1024 : 512) void k_fr_colfft_adj(const float *__restrict__ obj, float2 *__restrict__ R,
                                const float *__restrict__ phix, const float *__restrict__ phiy,
                                const int2 *__restrict__ circ, FftDesc fd, int n, int nz,
                                int nzc, int lgCB) {
  extern __shared__ float2 sm[];  // [a][t]: CB interleaved columns
  const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int cb = L % gridDim.x, c = L / gridDim.x;
  if (c >= nzc) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int CB = 1 << lgCB;
  const int n2 = 2 * n, v0 = cb * CB, h = n / 2;
  const int z2 = c + nzc;
  float2 *dst = R + (size_t)c * n2 * n;
  if constexpr (NS > 0) {
    // compiled sizes: first pass from obj (rows outside [1+h, n+h] are
    // zero), last pass straight to R
    constexpr int LG = col_lg(NS);
    fft_c_global<NS, -1, kFftEPT, LG>(
        sm, fd.tw, tid, nt,
        [&](int u, int t) {
          const int ty = u - 1 - h, tx = v0 + t;
          if (ty < 0 || ty >= n || tx >= n) return make_float2(0.f, 0.f);
          const int2 ci = circ[ty];
          if (tx < ci.x || tx >= ci.y) return make_float2(0.f, 0.f);
          const float ph = phix[tx] * phiy[ty] * ((u & 1) ? -1.f : 1.f);
          return make_float2(obj[((size_t)c * n + ty) * n + tx] * ph,
                             (z2 < nz) ? obj[((size_t)z2 * n + ty) * n + tx] * ph : 0.f);
        },
        [&](int a, int t, float2 x) {
          if (v0 + t >= n) return;
          dst[(size_t)a * n + v0 + t] = (a & 1) ? make_float2(-x.x, -x.y) : x;
        });
    return;
  }
  for (int idx = tid; idx < CB * lds_padded_len(n2); idx += nt) sm[idx] = make_float2(0.f, 0.f);
  __syncthreads();
  stage_lds<8, float2>(
      CB * n, tid, nt,
      [&](int idx) {
        const int ty = idx >> lgCB, tx = v0 + (idx & (CB - 1));
        if (tx >= n) return make_float2(0.f, 0.f);
        return make_float2(obj[((size_t)c * n + ty) * n + tx],
                           (z2 < nz) ? obj[((size_t)z2 * n + ty) * n + tx] : 0.f);
      },
      [&](int idx, float2 v) {
        const int ty = idx >> lgCB, t = idx & (CB - 1);
        const int tx = v0 + t;
        if (tx >= n) return;
        const int u = ty + 1 + h;
        const int2 ci = circ[ty];
        const float ph =
            (tx >= ci.x && tx < ci.y) ? phix[tx] * phiy[ty] * ((u & 1) ? -1.f : 1.f) : 0.f;
        sm[(lds_pad(u) << lgCB) + t] = make_float2(v.x * ph, v.y * ph);
      });
  __syncthreads();
  lds_fft<-1, (NS < 0 ? kFftEPT : kColEPT), fft_lv(NS)>(sm, fd, lgCB, tid, nt);
  for (int idx = tid; idx < CB * n2; idx += nt) {
    const int a = idx >> lgCB, t = idx & (CB - 1);
    if (v0 + t >= n) continue;
    const float2 x = sm[(lds_pad(a) << lgCB) + t];
    dst[(size_t)a * n + v0 + t] = (a & 1) ? make_float2(-x.x, -x.y) : x;
  }
}

template <int NS>
__global__ __launch_bounds__(512) void k_fr_rowfft_adj(const float2 *__restrict__ R, float2 *__restrict__ F, FftDesc fd,
                                int n, int nzc) {
  extern __shared__ float2 sm[];
  const int a = blockIdx.x, c = blockIdx.y;
  if (c >= nzc) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n2 = 2 * n, h = n / 2;
  const float2 *src = R + ((size_t)c * n2 + a) * n;
  if constexpr (NS > 0) {
    float2 *dst = F + ((size_t)c * n2 + a) * n2;
    fft_c_global<NS, -1, kFftEPT, 0>(
        sm, fd.tw, tid, nt,
        [&](int v, int) {
          const float2 x = (v >= h && v < h + n) ? src[v - h] : make_float2(0.f, 0.f);
          return (v & 1) ? make_float2(-x.x, -x.y) : x;
        },
        [&](int b, int, float2 x) { dst[b] = (b & 1) ? make_float2(-x.x, -x.y) : x; });
    return;
  }
  stage_lds<8, float2>(
      n2, tid, nt,
      [&](int v) { return (v >= h && v < h + n) ? src[v - h] : make_float2(0.f, 0.f); },
      [&](int v, float2 x) { sm[lds_pad(v)] = (v & 1) ? make_float2(-x.x, -x.y) : x; });
  __syncthreads();
  lds_fft<-1, kFftEPT, fft_lv(NS)>(sm, fd, 0, tid, nt);
  float2 *dst = F + ((size_t)c * n2 + a) * n2;
  for (int b = tid; b < n2; b += nt) {
    const float2 x = sm[lds_pad(b)];
    dst[b] = (b & 1) ? make_float2(-x.x, -x.y) : x;
  }
}

// Interpolation of the periodic grid at the polar samples (gather^T), tile
// by tile over the gridding's sample lists.  A wave takes one work item (a
// 16x16 tile's list, or a chunk of it), stages the tile's values of the 4
// complex slices of its group in LDS (16 x 16 x 4 float2, 8 KB), and each
// lane takes one list entry: the part of the sample's 9x9 footprint that lies
// in the tile, sum over taps of e1(row) e0(column) F, with the separable
// Gaussian factors e(i) = 2^(K2 (i - 4 - d)^2) (d the sample's fractional
// offset) set to zero outside the tile.  The partial sums go to ipart[entry];
// k_fr_interp_sum adds each sample's (at most 4) partials in a fixed order,
// so the result is deterministic.
__global__ __launch_bounds__(64 * kGridWaves) void k_fr_interp_tiles(
    const int4 *__restrict__ items, int nitems, const int4 *__restrict__ ent,
    const float2 *__restrict__ F, float4 *__restrict__ ipart, int64_t nent, int n, int ntile,
    int nzc, float K2) {
  // per wave: the tile's 256 cells in two planes (complex slices 0-1, 2-3),
  // cell (r, c) at r * kInterpPitch + c.  The lanes read unrelated cells; in
  // one plane, with the pad after each row, the cells of one column fall on
  // different 16-byte bank slots (with the two slice pairs of a cell side by
  // side, a read used only half of the slots: 3x the bank-conflict cycles)
  __shared__ float4 ft[kGridWaves][2][16 * kInterpPitch];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int it = blockIdx.x * kGridWaves + wave;
  const int grp = blockIdx.y;  // group of 4 complex slices; its own slab of ipart
  if (it >= nitems) return;
  ipart += (size_t)grp * 2 * nent;
  const int4 item = items[it];
  const int tile = item.x, ebeg = item.y, cnt = item.z;
  const int n2 = 2 * n;
  const int TA = tile / ntile, TB = tile - TA * ntile;
  const size_t plane = (size_t)n2 * n2;
  float4 *T0 = ft[wave][0], *T1 = ft[wave][1];
  auto at = [](int r, int c) { return r * kInterpPitch + c; };
  // stage the tile: 256 cells x 4 slices, 16 values per lane
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int cell = lane + 64 * q;
    const int r = cell >> 4, cc = cell & 15;
    const int a = 16 * TA + r, bcol = 16 * TB + cc;
    float2 v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      v[c] = make_float2(0.f, 0.f);
      if (grp * 4 + c < nzc && a < n2 && bcol < n2)
        v[c] = F[(size_t)(grp * 4 + c) * plane + (size_t)a * n2 + bcol];
    }
    T0[at(r, cc)] = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
    T1[at(r, cc)] = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
  }
  __builtin_amdgcn_wave_barrier();
  const int eend = ebeg + cnt;
  for (int e0 = ebeg; e0 < eend; e0 += 64) {
    const int e = e0 + lane;
    const int4 en = e < eend ? ent[e] : make_int4(-1, kPadStarts, 0, 0);
    if (en.x < 0) continue;  // list padding
    // footprint start relative to the tile (stored as half-precision start + 8)
    const int cs = (int)(float)__builtin_bit_cast(_Float16, (unsigned short)(en.y & 0xffff)) - 8;
    const int rs = (int)(float)__builtin_bit_cast(_Float16, (unsigned short)(en.y >> 16)) - 8;
    const float du = __int_as_float(en.z), dv = __int_as_float(en.w);
    float ex[9], ey[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float x0 = (float)(i - 4) - du, x1 = (float)(i - 4) - dv;
      ex[i] = (cs + i >= 0 && cs + i < 16) ? exp2f(K2 * x0 * x0) : 0.f;
      ey[i] = (rs + i >= 0 && rs + i < 16) ? exp2f(K2 * x1 * x1) : 0.f;
    }
    float4 acc0 = make_float4(0.f, 0.f, 0.f, 0.f), acc1 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i1 = 0; i1 < 9; ++i1) {
      const int r = min(max(rs + i1, 0), 15);
#pragma unroll
      for (int i0 = 0; i0 < 9; ++i0) {
        const int cc = min(max(cs + i0, 0), 15);
        const float w = ey[i1] * ex[i0];
        const int o = at(r, cc);
        const float4 x = T0[o], y = T1[o];
        acc0.x = fmaf(w, x.x, acc0.x);
        acc0.y = fmaf(w, x.y, acc0.y);
        acc0.z = fmaf(w, x.z, acc0.z);
        acc0.w = fmaf(w, x.w, acc0.w);
        acc1.x = fmaf(w, y.x, acc1.x);
        acc1.y = fmaf(w, y.y, acc1.y);
        acc1.z = fmaf(w, y.z, acc1.z);
        acc1.w = fmaf(w, y.w, acc1.w);
      }
    }
    ipart[2 * (size_t)e] = acc0;
    ipart[2 * (size_t)e + 1] = acc1;
  }
}

// G[c][s] = sum of sample s's tile partials (entries ient[s], -1 = none, in
// list order).
__global__ __launch_bounds__(256) void k_fr_interp_sum(const int4 *__restrict__ ient,
                                                       const float4 *__restrict__ ipart,
                                                       int64_t nent, float2 *__restrict__ G,
                                                       int nsamp, int nzc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int grp = blockIdx.y;
  if (s >= nsamp) return;
  ipart += (size_t)grp * 2 * nent;
  const int4 q = ient[s];
  const int idx[4] = {q.x, q.y, q.z, q.w};
  float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (idx[u] < 0) continue;
    const float4 x = ipart[2 * (size_t)idx[u]], y = ipart[2 * (size_t)idx[u] + 1];
    a0.x += x.x, a0.y += x.y, a0.z += x.z, a0.w += x.w;
    a1.x += y.x, a1.y += y.y, a1.z += y.z, a1.w += y.w;
  }
  const float2 v[4] = {make_float2(a0.x, a0.y), make_float2(a0.z, a0.w),
                       make_float2(a1.x, a1.y), make_float2(a1.z, a1.w)};
#pragma unroll
  for (int c = 0; c < 4; ++c)
    if (grp * 4 + c < nzc) G[(size_t)(grp * 4 + c) * nsamp + s] = v[c];
}

// Adjoint of k_fr_prep: the shifted inverse n-point DFT of each spectrum row
// back to the two real slices of the pair.  NS as k_fr_prep: compiled sizes
// read the spectrum straight into the first pass and store from the last.
template <int NS>
__global__ __launch_bounds__(512) void k_fr_prep_adj(const float2 *__restrict__ G,
                                                     float *__restrict__ data, FftDesc fd, int n,
                                                     int nproj, int nz, int nzc, float scale) {
  extern __shared__ float2 sm[];
  const int p = blockIdx.x, c = blockIdx.y;
  if (c >= nzc) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const float2 *in = G + ((size_t)c * nproj + p) * n;
  const int z2 = c + nzc;
  float *r1 = data + ((size_t)c * nproj + p) * n;
  float *r2 = data + ((size_t)z2 * nproj + p) * n;
  if constexpr (NS > 0) {
    fft_c_global<NS, 1, kFftEPT, 0>(
        sm, fd.tw, tid, nt, [&](int k, int) { return cscale(in[k], shift_sign(k) * scale); },
        [&](int j, int, float2 x) {
          x = cscale(x, shift_sign(j));
          r1[j] = x.x;
          if (z2 < nz) r2[j] = x.y;
        });
  } else {
    stage_lds<8, float2>(
        n, tid, nt, [&](int k) { return in[k]; },
        [&](int k, float2 v) { sm[lds_pad(k)] = cscale(v, shift_sign(k) * scale); });
    __syncthreads();
    lds_fft<1, kFftEPT, fft_lv(NS)>(sm, fd, 0, tid, nt);
    for (int j = tid; j < n; j += nt) {
      const float2 x = cscale(sm[lds_pad(j)], shift_sign(j));
      r1[j] = x.x;
      if (z2 < nz) r2[j] = x.y;
    }
  }
}

}  // namespace

struct tomo_fr_plan {
  int64_t n = 0, nproj = 0, nz = 0, nzc = 0, ngrp = 0;
  int dtype = TOMO_F32;
  size_t csz = sizeof(float2);  // bytes of one stored complex value
  int64_t nsamp = 0;
  FftPlan fft1, fft2;
  DevBuf ient, ent, items, red, part, phix, phiy, circ;
  DevBuf ipart;      // (list entries) x 4 complex: tile partials of fr_forward (first use)
  int64_t nent = 0;  // list entries, padding included
  DevBuf span;  // (2n,) int2: columns [x, y) of grid row a that hold samples
  std::vector<int4> items_h;  // the gridding work items (host copy, for batches)
  DevBuf gmax, gsc;  // float32 plans: per-row spectrum maxima, gridding scale exponent
  int ntile = 0, nitems = 0, nred = 0, nslots = 0;
  DevBuf G, F, R;
  int colCB = 1, lgCB = 0, colThreads = 256, prepThreads = 256, rowThreads = 256;
  int nsRow = 0, nsCol = 0, nsPrep = 0;  // compiled FFT sizes in use (0 = run-time sizes)
  int colThreadsStatic = 1024;           // column-pass workgroup of the compiled sizes
  int lgB = 0;                           // column-blocked R (see k_fr_rowfft)
  int gridChunk = kGridChunk;            // max list entries per gridding work item
  int rowLds = 0, prepLds = 0;
  size_t colLds = 0;
  float scale1 = 0.f, K2 = 0.f;
};

namespace {

int build_tables(tomo_fr_plan *P, const std::vector<float> &theta) {
  const int64_t n = P->n, nproj = P->nproj, n2 = 2 * n;
  const double mu = -std::log(1e-3) / (2.0 * (double)n * (double)n);
  const int m = 4;
  const double c0 = M_PI / (mu * 4.0 * (double)n * (double)n);
  const double C1n = -M_PI * M_PI / (mu * 4.0 * (double)n * (double)n);
  // Half-precision storage keeps the spectra and grids 2^8 above the
  // reference's scaling (well inside the normal range of IEEE half); the
  // factor is taken out again with the deapodisation.
  const double sG = P->dtype == TOMO_F16 ? 256.0 : 1.0;
  P->scale1 = (float)(4.0 / (double)n * c0 * sG);
  P->K2 = (float)(C1n / std::log(2.0));

  // polar sample geometry (kernels_fourierrec.cuh:23-36, :86-98), FP64
  const int64_t S = nproj * n;
  P->nsamp = S;
  std::vector<float4> geom(S);
  std::vector<int> b0v(S), b1v(S);
  for (int64_t p = 0; p < nproj; ++p) {
    const double th = (double)theta[p];
    const double ct = std::cos(th), st = std::sin(th);
    for (int64_t k = 0; k < n; ++k) {
      const double t = (double)(k - n / 2) / (double)n;
      double x = t * ct, y = -t * st;
      if (x >= 0.5) x = 0.5 - 1e-5;
      if (y >= 0.5) y = 0.5 - 1e-5;
      const double u = (double)(2 * n) * x, v = (double)(2 * n) * y;
      const double fu = std::floor(u), fv = std::floor(v);
      const int64_t s = p * n + k;
      b0v[s] = (int)fu - m;
      b1v[s] = (int)fv - m;
      geom[s] = make_float4((float)(u - fu), (float)(v - fv), (float)b0v[s], (float)b1v[s]);
    }
  }
  // 16x16 tile lists of the periodic grid: entries {sample, footprint start
  // relative to the tile (columns | rows << 8, each in [-8, 16))}, samples in
  // increasing order, each list padded with {-1, 0} to a multiple of the
  // samples per matrix instruction (4 for float32 plans, 16 for 16-bit plans)
  const int nt = (int)ceil_div(n2, 16);
  P->ntile = nt;
  auto tiles_of = [&](int base, int *tl, int *rel) {  // distinct tiles of rows base..base+8
    int cnt = 0;
    for (int q = 0; q < 9; ++q) {
      int r = (int)((n + base + q) % n2);
      if (r < 0) r += (int)n2;
      const int t = r / 16;
      bool seen = false;
      for (int u = 0; u < cnt; ++u) seen |= (tl[u] == t);
      if (!seen) {
        int v = (int)(((n + base - 16 * t) % n2 + n2) % n2);
        if (v >= n2 - 8) v -= (int)n2;
        tl[cnt] = t;
        rel[cnt] = v;
        cnt++;
      }
    }
    return cnt;
  };
  const size_t ntiles = (size_t)nt * nt;
  std::vector<int64_t> counts(ntiles + 1, 0);
  for (int64_t s = 0; s < S; ++s) {
    int tr[4], rr[4], tc[4], rc2[4];
    const int nr = tiles_of(b1v[s], tr, rr), nc = tiles_of(b0v[s], tc, rc2);
    for (int u = 0; u < nr; ++u)
      for (int v = 0; v < nc; ++v) counts[(size_t)tr[u] * nt + tc[v]]++;
  }
  std::vector<int64_t> off(ntiles + 1, 0);
  const int64_t gpad = 32;  // samples per half-precision matrix instruction
  for (size_t t = 0; t < ntiles; ++t) off[t + 1] = off[t] + (counts[t] + gpad - 1) / gpad * gpad;
  TOMO_REQUIRE(pad_starts() == kPadStarts, "fr_create: internal padding encoding mismatch");
  std::vector<int4> ent(std::max<int64_t>(1, off[ntiles]), make_int4(-1, pad_starts(), 0, 0));
  std::vector<int64_t> fill(off.begin(), off.end() - 1);
  // entries of each sample (at most 2 x 2 tiles), for fr_forward's partial sums
  std::vector<int4> ient(S, make_int4(-1, -1, -1, -1));
  std::vector<unsigned char> nient(S, 0);
  for (int64_t s = 0; s < S; ++s) {
    int tr[4], rr[4], tc[4], rc2[4];
    const int nr = tiles_of(b1v[s], tr, rr), nc = tiles_of(b0v[s], tc, rc2);
    for (int u = 0; u < nr; ++u)
      for (int v = 0; v < nc; ++v) {
        if (rr[u] < -8 || rr[u] >= 16 || rc2[v] < -8 || rc2[v] >= 16) {
          set_error("fr_create: internal tile offset out of range");
          return TOMO_EINVAL;
        }
        const size_t t = (size_t)tr[u] * nt + tc[v];
        float du = geom[s].x, dv = geom[s].y;
        int dub, dvb;
        std::memcpy(&dub, &du, 4);
        std::memcpy(&dvb, &dv, 4);
        const int64_t e = fill[t]++;
        ent[e] = make_int4((int)s, pack_starts(rc2[v], rr[u]), dub, dvb);
        int *slot = &ient[s].x;
        if (nient[s] >= 4) {
          set_error("fr_create: internal: a sample in more than 4 tiles");
          return TOMO_EINVAL;
        }
        slot[nient[s]++] = (int)e;
      }
  }
  TOMO_REQUIRE(off[ntiles] < ((int64_t)1 << 31), "fr_create: tile lists too large");
  // Occupied span of each tile row: the column tiles [lo, hi) between its
  // first and last tile with samples (the samples fill a disc, so the tiles
  // in between have samples too; if one does not, the row's span is the whole
  // row).  Tiles outside their row's span are never gridded (they are zero)
  // and the row pass reads only the span of each grid row.
  std::vector<int> tlo(nt, 0), thi(nt, 0);
  for (int ta = 0; ta < nt; ++ta) {
    int lo = nt, hi = -1;
    for (int tb = 0; tb < nt; ++tb)
      if (counts[(size_t)ta * nt + tb] > 0) lo = std::min(lo, tb), hi = std::max(hi, tb);
    if (hi < lo) continue;  // no samples in this tile row
    bool gap = false;
    for (int tb = lo; tb <= hi; ++tb) gap |= counts[(size_t)ta * nt + tb] == 0;
    if (gap) lo = 0, hi = nt - 1;
    tlo[ta] = lo;
    thi[ta] = hi + 1;
  }
  {
    std::vector<int2> span(n2);
    for (int64_t a = 0; a < n2; ++a) {
      const int ta = (int)(a / 16);
      span[a] = make_int2((int)std::min<int64_t>(16 * (int64_t)tlo[ta], n2),
                          (int)std::min<int64_t>(16 * (int64_t)thi[ta], n2));
    }
    int rc0 = P->span.alloc(sizeof(int2) * n2);
    if (rc0) return rc0;
    TOMO_HIP(hipMemcpy(P->span.p, span.data(), sizeof(int2) * n2, hipMemcpyHostToDevice));
  }
  // work items: whole tiles, or chunks of gridChunk entries for heavy tiles
  const int64_t kChunk = P->gridChunk;
  std::vector<int4> items, red;
  int nslots = 0;
  for (size_t t = 0; t < ntiles; ++t) {
    const int ta = (int)(t / nt), tb = (int)(t % nt);
    if (tb < tlo[ta] || tb >= thi[ta]) continue;  // outside the occupied span
    const int64_t L = off[t + 1] - off[t];
    if (L <= kChunk) {
      items.push_back(make_int4((int)t, (int)off[t], (int)L, -1));
    } else {
      const int nch = (int)ceil_div(L, kChunk);
      red.push_back(make_int4((int)t, nslots, nch, 0));
      for (int q = 0; q < nch; ++q)
        items.push_back(make_int4((int)t, (int)(off[t] + (int64_t)q * kChunk),
                                  (int)std::min<int64_t>(kChunk, L - (int64_t)q * kChunk),
                                  nslots + q));
      nslots += nch;
    }
  }
  // longest items first (results do not depend on the order); an XCD-aware
  // raster order (runs of neighbouring tiles per XCD, for L2 reuse of the
  // spectra) measured 17% slower: the load imbalance outweighs the reuse
  std::stable_sort(items.begin(), items.end(), [](const int4 &x, const int4 &y) { return x.z > y.z; });
  P->items_h = items;
  P->nitems = (int)items.size();
  P->nred = (int)red.size();
  P->nslots = nslots;
  int rc = P->ent.alloc(sizeof(int4) * ent.size());
  if (rc) return rc;
  TOMO_HIP(hipMemcpy(P->ent.p, ent.data(), sizeof(int4) * ent.size(), hipMemcpyHostToDevice));
  if ((rc = P->items.alloc(sizeof(int4) * std::max<size_t>(1, items.size())))) return rc;
  TOMO_HIP(hipMemcpy(P->items.p, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
  if (!red.empty()) {
    if ((rc = P->red.alloc(sizeof(int4) * red.size()))) return rc;
    TOMO_HIP(hipMemcpy(P->red.p, red.data(), sizeof(int4) * red.size(), hipMemcpyHostToDevice));
  }
  P->nent = off[ntiles];
  if ((rc = P->ient.alloc(sizeof(int4) * S))) return rc;
  TOMO_HIP(hipMemcpy(P->ient.p, ient.data(), sizeof(int4) * S, hipMemcpyHostToDevice));

  // deapodisation (kernels_fourierrec.cuh:14-15), FP64, separable in x and y;
  // the 1/nproj and (1 - n%4) factors ride on the y table
  std::vector<float> px(n), py(n);
  std::vector<int2> circ(n);
  for (int64_t t = 0; t < n; ++t) {
    const double d = (double)t / (double)n - 0.5;
    px[t] = (float)std::exp(mu * (double)n * (double)n * d * d);
    py[t] = (float)(std::exp(mu * (double)n * (double)n * d * d) / (double)nproj *
                    (double)(1 - n % 4) / sG);
  }
  // circle mask 4x^2 + 4y^2 < 1 (kernels_fourierrec.cuh:161-174), FP64, as a
  // per-row column interval
  for (int64_t ty = 0; ty < n; ++ty) {
    const double yy = (double)(ty - n / 2) / (double)n;
    int lo = (int)n, hi = 0;
    for (int64_t tx = 0; tx < n; ++tx) {
      const double xx = (double)(tx - n / 2) / (double)n;
      if (4.0 * xx * xx + 4.0 * yy * yy < 1.0) {
        lo = std::min<int>(lo, (int)tx);
        hi = std::max<int>(hi, (int)tx + 1);
      }
    }
    if (lo >= hi) lo = hi = 0;
    circ[ty] = make_int2(lo, hi);
  }
  if ((rc = P->phix.alloc(sizeof(float) * n))) return rc;
  if ((rc = P->phiy.alloc(sizeof(float) * n))) return rc;
  if ((rc = P->circ.alloc(sizeof(int2) * n))) return rc;
  TOMO_HIP(hipMemcpy(P->phix.p, px.data(), sizeof(float) * n, hipMemcpyHostToDevice));
  TOMO_HIP(hipMemcpy(P->phiy.p, py.data(), sizeof(float) * n, hipMemcpyHostToDevice));
  TOMO_HIP(hipMemcpy(P->circ.p, circ.data(), sizeof(int2) * n, hipMemcpyHostToDevice));
  return TOMO_OK;
}


// Workgroup shapes of the FFT passes for width n, from the radix sequences
// alone (no HIP call): the detector pass (n points), the row pass (2n) and
// the column pass (2n, one column per workgroup at least) each need at most
// 512 threads at their values per thread.  Rejects e.g. n = 4000
// (2n = 2^6 5^3: 15 values per thread in the radix-5 passes, 576 threads).
int fr_fft_shapes_ok(int64_t n) {
  FftDesc d1{}, d2{};
  int rc = FftPlan::radices((int)n, d1);
  if (!rc) rc = FftPlan::radices((int)(2 * n), d2);
  if (rc) return rc;
  const int colEPT = fft_general(n) ? kFftEPT : kColEPT;
  const struct {
    const char *pass;
    int threads;
  } need[] = {{"detector (n-point)", fft_threads(d1, n)},
              {"row (2n-point)", fft_threads(d2, 2 * n)},
              {"column (2n-point)", fft_threads(d2, 2 * n, colEPT)}};
  for (const auto &q : need)
    TOMO_REQUIRE(q.threads > 0 && q.threads <= 512,
                 "fr_create: n=%lld: the %s FFT pass needs more than 512 threads per workgroup "
                 "(its radix-3/5/7/11/13 passes cover fewer than 16 values per thread)",
                 (long long)n, q.pass);
  return TOMO_OK;
}

}  // namespace

extern "C" int fr_create(tomo_fr_plan **plan, int64_t n, int64_t nproj, int64_t nz,
                         const float *theta, int dtype, void *stream) {
  return fr_create_opts(plan, n, nproj, nz, theta, dtype, 0, 0, stream);
}

extern "C" int fr_create_opts(tomo_fr_plan **plan, int64_t n, int64_t nproj, int64_t nz,
                              const float *theta, int dtype, int grid_chunk, int fft_generic,
                              void *stream) {
  clear_error();
  TOMO_REQUIRE(plan != nullptr, "fr_create: plan pointer is NULL");
  *plan = nullptr;
  TOMO_REQUIRE(n >= 20 && n % 2 == 0, "fr_create: n=%lld must be even and >= 20", (long long)n);
  TOMO_REQUIRE(n <= 4096, "fr_create: n=%lld exceeds the supported maximum 4096", (long long)n);
  TOMO_REQUIRE(fft_size_ok(n), "fr_create: n=%lld must have no prime factor above %d",
               (long long)n, kFftMaxPrime);
  TOMO_REQUIRE(nproj >= 1 && nproj <= (1 << 20), "fr_create: nproj=%lld out of range", (long long)nproj);
  TOMO_REQUIRE(nz >= 1 && nz <= (1 << 16), "fr_create: nz=%lld out of range", (long long)nz);
  TOMO_REQUIRE(theta != nullptr, "fr_create: theta is NULL");
  TOMO_REQUIRE(dtype == TOMO_F32 || dtype == TOMO_F16 || dtype == TOMO_BF16,
               "fr_create: unknown dtype %d", dtype);
  TOMO_REQUIRE(nproj * n < (int64_t)1 << 31, "fr_create: nproj*n too large");
  TOMO_REQUIRE(grid_chunk == 0 || (grid_chunk >= 64 && grid_chunk <= 8192 && grid_chunk % 64 == 0),
               "fr_create: grid_chunk=%d must be 0 (default) or a multiple of 64 in [64, 8192]",
               grid_chunk);
  TOMO_REQUIRE(fft_generic == 0 || fft_generic == 1, "fr_create: fft_generic must be 0 or 1");
  {
    const int rc0 = fr_fft_shapes_ok(n);
    if (rc0) return rc0;
  }
  auto *P = new (std::nothrow) tomo_fr_plan();
  if (!P) {
    set_error("fr_create: host allocation failed");
    return TOMO_ENOMEM;
  }
  P->n = n;
  P->nproj = nproj;
  P->nz = nz;
  P->nzc = (nz + 1) / 2;
  P->ngrp = ceil_div(P->nzc, 4);
  P->dtype = dtype;
  if (grid_chunk > 0) P->gridChunk = grid_chunk;
  P->csz = dtype == TOMO_F32 ? sizeof(float2) : 2 * sizeof(uint16_t);
  hipStream_t st = (hipStream_t)stream;
  std::vector<float> th(nproj);
  int rc = TOMO_OK;
  auto fail = [&](int code) {
    fr_destroy(P);
    return code;
  };
  {
    hipError_t e = hipMemcpyAsync(th.data(), theta, sizeof(float) * nproj, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error("fr_create: reading theta failed: %s", hipGetErrorString(e));
      return fail(TOMO_EHIP);
    }
  }
  if ((rc = build_tables(P, th))) return fail(rc);
  if ((rc = P->fft1.init((int)n))) return fail(rc);
  if ((rc = P->fft2.init((int)(2 * n)))) return fail(rc);
  const int64_t n2 = 2 * n;
  const size_t plane = (size_t)n2 * n2;
  // G holds the float32 spectra of the adjoint as well
  if ((rc = P->G.alloc(sizeof(float2) * (size_t)P->ngrp * 4 * P->nsamp))) return fail(rc);
  if (P->dtype == TOMO_F32 &&
      ((rc = P->gmax.alloc(sizeof(float) * (size_t)P->ngrp * 4 * P->nproj)) ||
       (rc = P->gsc.alloc(256))))
    return fail(rc);
  if ((rc = P->part.alloc(sizeof(float) * 2048 * (size_t)P->ngrp * std::max(1, P->nslots))))
    return fail(rc);
  // F and R in the plan's storage precision (float32 when the adjoint may run)
  if ((rc = P->F.alloc(P->csz * (size_t)P->ngrp * 4 * plane))) return fail(rc);
  if ((rc = P->R.alloc(P->csz * (size_t)P->ngrp * 4 * n2 * n))) return fail(rc);
  // column pass: CB columns of length 2n in LDS
  const size_t maxlds = max_lds_per_block();
  const size_t col1 = (size_t)lds_padded_len((int)n2) * sizeof(float2);
  // the general-size column kernel (NS = -1) stages 16 values per thread: its
  // radix-5/7 and prime passes spill at 32
  const int colEPT = fft_general(n) ? kFftEPT : kColEPT;
  int CB = 1, lg = 0;
  while (CB * 2 <= 16 && (size_t)(CB * 2) * col1 <= maxlds &&
         fft_threads(P->fft2.d, (int64_t)(CB * 2) * n2, colEPT) > 0 &&
         fft_threads(P->fft2.d, (int64_t)(CB * 2) * n2, colEPT) <= 512)
    CB *= 2, lg += 1;
  P->colCB = CB;
  P->lgCB = lg;
  P->colLds = (size_t)CB * col1;
  TOMO_REQUIRE(P->colLds <= maxlds, "fr_create: 2n=%lld exceeds the LDS of one workgroup", (long long)n2);
  P->colThreads = fft_threads(P->fft2.d, (int64_t)CB * n2, colEPT);
  P->prepThreads = fft_threads(P->fft1.d, n);
  P->rowThreads = fft_threads(P->fft2.d, n2);
  if (P->colThreads < 0 || P->prepThreads < 0 || P->rowThreads < 0 || P->colThreads > 512 ||
      P->prepThreads > 512 || P->rowThreads > 512) {
    set_error("fr_create: n=%lld needs more than 512 threads per FFT workgroup", (long long)n);
    return fail(TOMO_EINVAL);
  }
  // size-specialised kernels where 2n is a compiled size (with_static_size):
  // the row pass, and the column pass with col_lg(2n) columns per workgroup
  // (up to 1024 threads at 16 values each, 15 with radix-3/5 passes); the
  // spectrum kernels where n is one too.  Otherwise the run-time-plan
  // kernels of the size's transform level (0, -1, -2).
  const int lev = fft_level(n);
  const int rt = lev == 2 ? -2 : (lev == 1 ? -1 : 0);
  const bool st2 = !fft_generic && static_fft_size((int)n2);
  P->nsRow = st2 ? (int)n2 : rt;
  P->nsCol = rt;
  if (st2) {
    const int lgs = col_lg((int)n2);
    const size_t lds_s = (size_t)(1 << lgs) * col1;
    const int thr = fft_threads(P->fft2.d, (int64_t)n2 << lgs, kFftEPT);
    if (lds_s <= maxlds && thr > 0 && thr <= 1024) {
      P->nsCol = (int)n2;
      CB = 1 << lgs;
      lg = lgs;
      P->colCB = CB;
      P->lgCB = lg;
      P->colLds = lds_s;
      P->colThreadsStatic = thr;
    }
  }
  // column-blocked R for 16-bit plans only: measured at n=2048 the column
  // pass gains 40-50 us per chunk of 8 slices, the row pass loses 22 us
  // (16-byte segments) in half precision and 43 us (32-byte) in float32
  P->lgB = (dtype != TOMO_F32 && n % (1 << lg) == 0) ? lg : 0;
  P->nsPrep = (st2 && static_fft_size((int)n)) ? (int)n : rt;
  hipError_t ea = hipSuccess;
  auto lds_attr = [&](const void *k, size_t bytes) {
    if (ea == hipSuccess)
      ea = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  };
  with_static_size(P->nsCol, [&](auto ns) {
    constexpr int NS = decltype(ns)::value;
    lds_attr((const void *)k_fr_colfft<float, NS>, P->colLds);
    lds_attr((const void *)k_fr_colfft<_Float16, NS>, P->colLds);
    lds_attr((const void *)k_fr_colfft<__bf16, NS>, P->colLds);
  });
  if (ea != hipSuccess) {
    set_error("fr_create: hipFuncSetAttribute failed: %s", hipGetErrorString(ea));
    return fail(TOMO_EHIP);
  }
  with_static_size(P->nsCol, [&](auto ns) {
    lds_attr((const void *)k_fr_colfft_adj<decltype(ns)::value>, P->colLds);
  });
  const int rowLds = (int)(lds_padded_len((int)n2) * sizeof(float2));
  P->rowLds = rowLds;
  P->prepLds = (int)(lds_padded_len((int)n) * sizeof(float2));
  with_static_size(P->nsRow, [&](auto ns) {
    constexpr int NS = decltype(ns)::value;
    lds_attr((const void *)k_fr_rowfft<float, NS>, rowLds);
    lds_attr((const void *)k_fr_rowfft<_Float16, NS>, rowLds);
    lds_attr((const void *)k_fr_rowfft<__bf16, NS>, rowLds);
  });
  if (ea != hipSuccess) {
    set_error("fr_create: hipFuncSetAttribute failed: %s", hipGetErrorString(ea));
    return fail(TOMO_EHIP);
  }
  with_static_size(P->nsRow, [&](auto ns) {
    lds_attr((const void *)k_fr_rowfft_adj<decltype(ns)::value>, rowLds);
  });
  if (ea != hipSuccess) {
    set_error("fr_create: hipFuncSetAttribute failed: %s", hipGetErrorString(ea));
    return fail(TOMO_EHIP);
  }
  *plan = P;
  return TOMO_OK;
}

template <typename T>
static int fr_tail_after_grid(tomo_fr_plan *P, T *obj, hipStream_t st);

// Gridding, the two pruned FFT passes and the output stage, from G; T is the
// storage type of G, F, R and obj.
template <typename T>
static int fr_tail(tomo_fr_plan *P, T *obj, hipStream_t st) {
  const int n = (int)P->n, nz = (int)P->nz, nzc = (int)P->nzc;
  const int ng = (int)P->ngrp;
  const int n2 = 2 * n;
  cpx<T> *G = P->G.as<cpx<T>>(), *F = P->F.as<cpx<T>>(), *R = P->R.as<cpx<T>>();
  {
    const dim3 gg((unsigned)ceil_div(P->nitems, kGridWaves), ng);
    if constexpr (std::is_same<T, float>::value) {
      {
        ProfScope ps1_("k_fr_gscale", st);
        hipLaunchKernelGGL(k_fr_gscale, dim3(1), dim3(1024), 0, st, P->gmax.as<float>(),
                           nzc * (int)P->nproj, P->gsc.as<int>());
      }
      ProfScope ps_("k_fr_grid_x3", st);
      hipLaunchKernelGGL(k_fr_grid_x3, gg, dim3(64 * kGridWaves), 0, st, P->items.as<int4>(),
                         P->nitems, P->ent.as<int4>(), (const float2 *)G, (float2 *)F,
                         P->part.as<float>(), P->gsc.as<int>(), n, P->ntile, P->nslots,
                         (int)P->nsamp, nzc, P->K2);
    } else {
      ProfScope ps_("k_fr_grid16", st);
      hipLaunchKernelGGL(k_fr_grid16<T>, gg, dim3(64 * kGridWaves), 0, st, P->items.as<int4>(),
                         P->nitems, P->ent.as<int4>(), G, F, P->part.as<float>(), n, P->ntile,
                         P->nslots, (int)P->nsamp, nzc, P->K2);
    }
  }
  TOMO_LAUNCH_CHECK();
  return fr_tail_after_grid<T>(P, obj, st);
}

// The stages after the gridding: the partial-tile sums, the two pruned FFT
// passes and the output stage.
template <typename T>
static int fr_tail_after_grid(tomo_fr_plan *P, T *obj, hipStream_t st) {
  const int n = (int)P->n, nz = (int)P->nz, nzc = (int)P->nzc;
  const int ng = (int)P->ngrp;
  const int n2 = 2 * n;
  cpx<T> *F = P->F.as<cpx<T>>(), *R = P->R.as<cpx<T>>();
  if (P->nred > 0) {
    ProfScope ps_("k_fr_grid_reduce", st);
    hipLaunchKernelGGL(k_fr_grid_reduce<T>, dim3(P->nred, ng, 4), dim3(256), 0, st,
                       P->red.as<int4>(), P->part.as<float>(), F, n, P->ntile, P->nslots, nzc);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_fr_rowfft", st);
    with_static_size(P->nsRow, [&](auto ns) {
      hipLaunchKernelGGL((k_fr_rowfft<T, decltype(ns)::value>), dim3(n2, ng * 4),
                         dim3(P->rowThreads), P->rowLds, st, F, R, P->fft2.d, n, nzc, P->lgB,
                         P->span.as<int2>());
    });
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_fr_colfft", st);
    with_static_size(P->nsCol, [&](auto ns) {
      hipLaunchKernelGGL((k_fr_colfft<T, decltype(ns)::value>),
                         dim3((unsigned)ceil_div(n, P->colCB), ng * 4),
                         dim3(decltype(ns)::value > 0 ? P->colThreadsStatic : P->colThreads),
                         P->colLds, st, R, obj, P->phix.as<float>(), P->phiy.as<float>(),
                         P->circ.as<int2>(), P->fft2.d, n, nz, nzc, P->lgCB, P->lgB > 0 ? 1 : 0);
    });
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

template <typename T>
static int fr_backprojection_t(tomo_fr_plan *P, T *obj, const T *data, hipStream_t st) {
  const int n = (int)P->n, nproj = (int)P->nproj, nz = (int)P->nz, nzc = (int)P->nzc;
  const int ng = (int)P->ngrp;
  {
    ProfScope ps_("k_fr_prep", st);
    with_static_size(P->nsPrep, [&](auto ns) {
      hipLaunchKernelGGL((k_fr_prep<T, decltype(ns)::value>), dim3(nproj, ng * 4),
                         dim3(P->prepThreads), P->prepLds, st, data, P->G.as<cpx<T>>(), P->fft1.d,
                         n, nproj, nz, nzc, P->scale1, P->gmax.as<float>());
    });
  }
  TOMO_LAUNCH_CHECK();
  return fr_tail<T>(P, obj, st);
}

template <typename T>
static int fr_head(tomo_fr_plan *P, tomo_flt_plan *F, const T *data, hipStream_t st) {
  const int n = (int)P->n, nproj = (int)P->nproj, nz = (int)P->nz, nzc = (int)P->nzc;
  const int ng = (int)P->ngrp;
  {
    ProfScope ps_("k_fr_filtprep", st);
    with_static_size(P->nsPrep, [&](auto ns) {
      hipLaunchKernelGGL((k_fr_filtprep<T, decltype(ns)::value>), dim3(nproj, ng * 4),
                         dim3(P->rowThreads), P->rowLds, st, data, P->G.as<cpx<T>>(),
                         F->H2.as<float2>(), F->HL.as<float>(), F->HR.as<float>(), F->fft2n.d,
                         P->fft1.d, n, nproj, nz, nzc, P->scale1, P->gmax.as<float>());
    });
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

template <typename T>
static int fr_filter_backprojection_t(tomo_fr_plan *P, tomo_flt_plan *F, T *obj, const T *data,
                                      hipStream_t st) {
  const int rc = fr_head<T>(P, F, data, st);
  return rc ? rc : fr_tail<T>(P, obj, st);
}

extern "C" int fr_backprojection(tomo_fr_plan *P, void *obj, const void *data, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "fr_backprojection: plan is NULL");
  TOMO_REQUIRE(obj != nullptr && data != nullptr, "fr_backprojection: NULL array");
  hipStream_t st = (hipStream_t)stream;
  switch (P->dtype) {
    case TOMO_F16:
      return fr_backprojection_t(P, (_Float16 *)obj, (const _Float16 *)data, st);
    case TOMO_BF16:
      return fr_backprojection_t(P, (__bf16 *)obj, (const __bf16 *)data, st);
    default:
      return fr_backprojection_t(P, (float *)obj, (const float *)data, st);
  }
}

extern "C" int fr_filter_backprojection(tomo_fr_plan *P, tomo_flt_plan *F, void *obj,
                                        const void *data, const float *wfilter, double center,
                                        const float *sht, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr && F != nullptr, "fr_filter_backprojection: plan is NULL");
  TOMO_REQUIRE(obj != nullptr && data != nullptr && wfilter != nullptr,
               "fr_filter_backprojection: NULL array");
  TOMO_REQUIRE(std::isfinite(center), "fr_filter_backprojection: center is not finite");
  TOMO_REQUIRE(F->n == P->n && F->nproj == P->nproj && F->nz == P->nz,
               "fr_filter_backprojection: filter plan (n=%lld nproj=%lld nz=%lld) does not match "
               "(n=%lld nproj=%lld nz=%lld)", (long long)F->n, (long long)F->nproj,
               (long long)F->nz, (long long)P->n, (long long)P->nproj, (long long)P->nz);
  TOMO_REQUIRE(F->H2.p != nullptr, "fr_filter_backprojection: filter plan has no fused path "
               "(needs ne >= 2n and 2n <= 8192)");
  hipStream_t st = (hipStream_t)stream;
  int rc = flt_prepare_2n(F, wfilter, center, sht, st);
  if (rc) return rc;
  return fr_filter_backprojection_prepared(P, F, obj, data, stream);
}

extern "C" int fr_filter_backprojection_prepared(tomo_fr_plan *P, tomo_flt_plan *F, void *obj,
                                                 const void *data, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr && F != nullptr, "fr_filter_backprojection_prepared: plan is NULL");
  TOMO_REQUIRE(obj != nullptr && data != nullptr, "fr_filter_backprojection_prepared: NULL array");
  TOMO_REQUIRE(F->n == P->n && F->nproj == P->nproj && F->nz == P->nz,
               "fr_filter_backprojection_prepared: filter plan does not match the plan");
  TOMO_REQUIRE(F->H2.p != nullptr && F->prepared,
               "fr_filter_backprojection_prepared: the filter plan has no prepared kernels "
               "(call flt_prepare first)");
  hipStream_t st = (hipStream_t)stream;
  switch (P->dtype) {
    case TOMO_F16:
      return fr_filter_backprojection_t(P, F, (_Float16 *)obj, (const _Float16 *)data, st);
    case TOMO_BF16:
      return fr_filter_backprojection_t(P, F, (__bf16 *)obj, (const __bf16 *)data, st);
    default:
      return fr_filter_backprojection_t(P, F, (float *)obj, (const float *)data, st);
  }
}

// A batch: the plans, the gridding work list over all of them, and
// kBatchStreams internal streams (default priority) over which the
// per-request stages fan out from the caller's stream and join back, through
// events, so that the call stays stream-ordered and capturable.
constexpr int kBatchStreams = 4;

struct tomo_fr_batch {
  std::vector<tomo_fr_plan *> plans;
  std::vector<int> lane;  // internal stream of each request (longest first, least loaded)
  DevBuf jobs, wi, wj;
  int nwork = 0;
  hipStream_t st[kBatchStreams] = {};
  hipEvent_t fork = nullptr, mid = nullptr, done[kBatchStreams] = {};
};

extern "C" void fr_batch_destroy(tomo_fr_batch *B) {
  if (!B) return;
  B->jobs.release();
  B->wi.release();
  B->wj.release();
  for (int i = 0; i < kBatchStreams; ++i) {
    if (B->st[i]) (void)hipStreamDestroy(B->st[i]);
    if (B->done[i]) (void)hipEventDestroy(B->done[i]);
  }
  if (B->fork) (void)hipEventDestroy(B->fork);
  if (B->mid) (void)hipEventDestroy(B->mid);
  delete B;
}

extern "C" int fr_batch_create(tomo_fr_batch **batch, tomo_fr_plan *const *plans, int count) {
  clear_error();
  TOMO_REQUIRE(batch != nullptr, "fr_batch_create: batch pointer is NULL");
  *batch = nullptr;
  TOMO_REQUIRE(plans != nullptr && count >= 1, "fr_batch_create: need at least one plan");
  for (int r = 0; r < count; ++r) {
    TOMO_REQUIRE(plans[r] != nullptr, "fr_batch_create: plan %d is NULL", r);
    TOMO_REQUIRE(plans[r]->dtype == TOMO_F32, "fr_batch_create: plan %d is not float32", r);
    for (int q = 0; q < r; ++q)
      TOMO_REQUIRE(plans[q] != plans[r], "fr_batch_create: plan %d appears twice", r);
  }
  auto *B = new (std::nothrow) tomo_fr_batch();
  if (!B) {
    set_error("fr_batch_create: host allocation failed");
    return TOMO_ENOMEM;
  }
  B->plans.assign(plans, plans + count);
  // requests by decreasing cost (spectra + grid planes, in elements)
  std::vector<double> cost(count);
  for (int r = 0; r < count; ++r) {
    const tomo_fr_plan *P = B->plans[r];
    cost[r] = (double)P->nz * P->nproj * P->n + (double)P->nzc * 4.0 * P->n * P->n;
  }
  std::vector<int> order(count);
  for (int r = 0; r < count; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  B->lane.assign(count, 0);
  double load[kBatchStreams] = {};
  for (int r : order) {
    int k = 0;
    for (int i = 1; i < kBatchStreams; ++i)
      if (load[i] < load[k]) k = i;
    B->lane[r] = k;
    load[k] += cost[r];
  }
  // the gridding work: each job's (plan, group's) items stay together, in the
  // plan's longest-first order, jobs of the largest requests first; so
  // consecutive workgroups share a plan's spectra in L2 as in the per-plan
  // launch, and the small requests' short items fill the tail
  std::vector<GridJob> jobs;
  std::vector<std::pair<int4, int>> work;
  for (int r : order) {
    tomo_fr_plan *P = B->plans[r];
    const size_t plane = (size_t)(2 * P->n) * (2 * P->n);
    for (int g = 0; g < (int)P->ngrp; ++g) {
      GridJob jb{};
      jb.ent = P->ent.as<int4>();
      jb.G = P->G.as<float2>() + (size_t)(4 * g) * P->nsamp;
      jb.F = P->F.as<float2>() + (size_t)(4 * g) * plane;
      jb.part = P->part.as<float>() + (size_t)g * P->nslots * 2048;
      jb.ksc = P->gsc.as<int>();
      jb.n = (int)P->n;
      jb.ntile = P->ntile;
      jb.nsamp = (int)P->nsamp;
      jb.nlive = (int)std::min<int64_t>(4, P->nzc - 4 * g);
      jb.K2 = P->K2;
      for (const int4 &it : P->items_h) work.push_back({it, (int)jobs.size()});
      jobs.push_back(jb);
    }
  }
  std::vector<int4> wi(work.size());
  std::vector<int> wj(work.size());
  for (size_t k = 0; k < work.size(); ++k) wi[k] = work[k].first, wj[k] = work[k].second;
  B->nwork = (int)work.size();
  int rc = B->jobs.alloc(sizeof(GridJob) * jobs.size());
  if (!rc) rc = B->wi.alloc(sizeof(int4) * std::max<size_t>(1, wi.size()));
  if (!rc) rc = B->wj.alloc(sizeof(int) * std::max<size_t>(1, wj.size()));
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpy(B->jobs.p, jobs.data(), sizeof(GridJob) * jobs.size(), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess && !wi.empty())
    e = hipMemcpy(B->wi.p, wi.data(), sizeof(int4) * wi.size(), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess && !wj.empty())
    e = hipMemcpy(B->wj.p, wj.data(), sizeof(int) * wj.size(), hipMemcpyHostToDevice);
  if (!rc && e != hipSuccess) {
    set_error("fr_batch_create: copy failed: %s", hipGetErrorString(e));
    rc = TOMO_EHIP;
  }
  for (int i = 0; i < kBatchStreams && !rc; ++i) {
    e = hipStreamCreateWithFlags(&B->st[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B->done[i], hipEventDisableTiming);
    if (e != hipSuccess) {
      set_error("fr_batch_create: stream or event creation failed: %s", hipGetErrorString(e));
      rc = TOMO_EHIP;
    }
  }
  if (!rc) {
    e = hipEventCreateWithFlags(&B->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B->mid, hipEventDisableTiming);
    if (e != hipSuccess) {
      set_error("fr_batch_create: event creation failed: %s", hipGetErrorString(e));
      rc = TOMO_EHIP;
    }
  }
  if (rc) {
    fr_batch_destroy(B);
    return rc;
  }
  *batch = B;
  return TOMO_OK;
}

// The caller's stream `st` -> every internal stream (after `ev` on st).
static int batch_fork(tomo_fr_batch *B, hipStream_t st, hipEvent_t ev) {
  hipError_t e = hipEventRecord(ev, st);
  for (int i = 0; i < kBatchStreams && e == hipSuccess; ++i) e = hipStreamWaitEvent(B->st[i], ev, 0);
  if (e != hipSuccess) {
    set_error("fr_batch_filter_backprojection: fork failed: %s", hipGetErrorString(e));
    return TOMO_EHIP;
  }
  return TOMO_OK;
}

// Every internal stream -> the caller's stream.
static int batch_join(tomo_fr_batch *B, hipStream_t st) {
  hipError_t e = hipSuccess;
  for (int i = 0; i < kBatchStreams && e == hipSuccess; ++i) {
    e = hipEventRecord(B->done[i], B->st[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, B->done[i], 0);
  }
  if (e != hipSuccess) {
    set_error("fr_batch_filter_backprojection: join failed: %s", hipGetErrorString(e));
    return TOMO_EHIP;
  }
  return TOMO_OK;
}

extern "C" int fr_batch_filter_backprojection(tomo_fr_batch *B, tomo_flt_plan *const *flts,
                                              void *const *objs, const void *const *datas,
                                              void *stream) {
  clear_error();
  TOMO_REQUIRE(B != nullptr && flts != nullptr && objs != nullptr && datas != nullptr,
               "fr_batch_filter_backprojection: NULL argument");
  const int count = (int)B->plans.size();
  for (int r = 0; r < count; ++r) {
    const tomo_fr_plan *P = B->plans[r];
    const tomo_flt_plan *F = flts[r];
    TOMO_REQUIRE(F != nullptr && objs[r] != nullptr && datas[r] != nullptr,
                 "fr_batch_filter_backprojection: NULL array or filter plan for request %d", r);
    TOMO_REQUIRE(F->n == P->n && F->nproj == P->nproj && F->nz == P->nz,
                 "fr_batch_filter_backprojection: filter plan %d does not match its plan", r);
    TOMO_REQUIRE(F->H2.p != nullptr && F->prepared,
                 "fr_batch_filter_backprojection: filter plan %d is not prepared (flt_prepare)",
                 r);
  }
  hipStream_t st = (hipStream_t)stream;
  // filter + spectra of every request, fanned out over the internal streams
  int rc = batch_fork(B, st, B->fork);
  for (int r = 0; r < count && !rc; ++r) {
    tomo_fr_plan *P = B->plans[r];
    hipStream_t s = B->st[B->lane[r]];
    rc = fr_head<float>(P, flts[r], (const float *)datas[r], s);
    if (rc) break;
    ProfScope ps_("k_fr_gscale", s);
    hipLaunchKernelGGL(k_fr_gscale, dim3(1), dim3(1024), 0, s, P->gmax.as<float>(),
                       (int)(P->nzc * P->nproj), P->gsc.as<int>());
  }
  if (!rc) rc = batch_join(B, st);
  if (rc) return rc;
  TOMO_LAUNCH_CHECK();
  // one gridding launch for all requests, on the caller's stream
  if (B->nwork > 0) {
    ProfScope ps_("k_fr_grid_x3_batch", st);
    hipLaunchKernelGGL(k_fr_grid_x3_batch, dim3((unsigned)ceil_div(B->nwork, kGridWaves)),
                       dim3(64 * kGridWaves), 0, st, B->jobs.as<GridJob>(), B->wi.as<int4>(),
                       B->wj.as<int>(), B->nwork);
  }
  TOMO_LAUNCH_CHECK();
  // partial-tile sums, FFT passes and output stage, fanned out again
  rc = batch_fork(B, st, B->mid);
  for (int r = 0; r < count && !rc; ++r)
    rc = fr_tail_after_grid<float>(B->plans[r], (float *)objs[r], B->st[B->lane[r]]);
  if (!rc) rc = batch_join(B, st);
  return rc;
}

extern "C" int fr_forward(tomo_fr_plan *P, void *data_, const void *obj_, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "fr_forward: plan is NULL");
  TOMO_REQUIRE(obj_ != nullptr && data_ != nullptr, "fr_forward: NULL array");
  if (P->dtype != TOMO_F32) {
    set_error("fr_forward: the forward projection is implemented for float32 plans only");
    return TOMO_EUNSUPPORTED;
  }
  // tile partials of the interpolation, allocated on the first call: one slab
  // per group of 4 complex slices, for up to kInterpGroups groups per launch
  const int64_t nentp = std::max<int64_t>(1, P->nent);
  const int gb = (int)std::max<int64_t>(
      1, std::min<int64_t>(P->ngrp, kInterpBytes / (2 * (int64_t)sizeof(float4) * nentp)));
  if (!P->ipart.p) {
    const int rc0 = P->ipart.alloc(2 * sizeof(float4) * (size_t)nentp * gb);
    if (rc0) return rc0;
  }
  float *data = (float *)data_;
  const float *obj = (const float *)obj_;
  hipStream_t st = (hipStream_t)stream;
  const int n = (int)P->n, nproj = (int)P->nproj, nz = (int)P->nz, nzc = (int)P->nzc;
  const int ng = (int)P->ngrp;
  const int n2 = 2 * n;
  {
    ProfScope ps_("k_fr_colfft_adj", st);
    with_static_size(P->nsCol, [&](auto ns) {
      constexpr int NS = decltype(ns)::value;
      hipLaunchKernelGGL(k_fr_colfft_adj<NS>, dim3((unsigned)ceil_div(n, P->colCB), ng * 4),
                         dim3(NS > 0 ? P->colThreadsStatic : P->colThreads), P->colLds, st, obj,
                         P->R.as<float2>(),
                         P->phix.as<float>(), P->phiy.as<float>(), P->circ.as<int2>(),
                         P->fft2.d, n, nz, nzc, P->lgCB);
    });
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_fr_rowfft_adj", st);
    with_static_size(P->nsRow, [&](auto ns) {
      hipLaunchKernelGGL(k_fr_rowfft_adj<decltype(ns)::value>, dim3(n2, ng * 4),
                         dim3(P->rowThreads), P->rowLds,
                         st, P->R.as<float2>(), P->F.as<float2>(), P->fft2.d, n, nzc);
    });
  }
  TOMO_LAUNCH_CHECK();
  for (int g = 0; g < ng; g += gb) {
    const int ngl = std::min(gb, ng - g);  // groups in this launch
    {
      ProfScope ps_("k_fr_interp_tiles", st);
      hipLaunchKernelGGL(k_fr_interp_tiles,
                         dim3((unsigned)ceil_div(P->nitems, kGridWaves), (unsigned)ngl, 1),
                         dim3(64 * kGridWaves), 0, st, P->items.as<int4>(), P->nitems,
                         P->ent.as<int4>(), P->F.as<float2>() + (size_t)g * 4 * n2 * n2,
                         P->ipart.as<float4>(), nentp, n, P->ntile, nzc - 4 * g, P->K2);
    }
    TOMO_LAUNCH_CHECK();
    {
      ProfScope ps_("k_fr_interp_sum", st);
      hipLaunchKernelGGL(k_fr_interp_sum, dim3((unsigned)ceil_div(P->nsamp, 256), (unsigned)ngl),
                         dim3(256), 0, st, P->ient.as<int4>(), P->ipart.as<float4>(), nentp,
                         P->G.as<float2>() + (size_t)g * 4 * P->nsamp, (int)P->nsamp,
                         nzc - 4 * g);
    }
    TOMO_LAUNCH_CHECK();
  }
  {
    ProfScope ps_("k_fr_prep_adj", st);
    with_static_size(P->nsPrep, [&](auto ns) {
      hipLaunchKernelGGL(k_fr_prep_adj<decltype(ns)::value>, dim3(nproj, ng * 4),
                         dim3(P->prepThreads), P->prepLds, st, P->G.as<float2>(), data,
                         P->fft1.d, n, nproj, nz, nzc, P->scale1);
    });
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}

extern "C" void fr_destroy(tomo_fr_plan *P) {
  if (!P) return;
  P->ient.release();
  P->ipart.release();
  P->span.release();
  P->ent.release();
  P->items.release();
  P->red.release();
  P->part.release();
  P->gmax.release();
  P->gsc.release();
  P->phix.release();
  P->phiy.release();
  P->circ.release();
  P->G.release();
  P->F.release();
  P->R.release();
  P->fft1.tw.release();
  P->fft2.tw.release();
  delete P;
}

extern "C" int fr_get_sizes(const tomo_fr_plan *P, int64_t *n, int64_t *nproj, int64_t *nz) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "fr_get_sizes: plan is NULL");
  if (n) *n = P->n;
  if (nproj) *nproj = P->nproj;
  if (nz) *nz = P->nz;
  return TOMO_OK;
}
