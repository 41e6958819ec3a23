// "vo-all" ring removal (tomocupy remove_all_stripe, src/tomocupy/processing/
// remove_stripe.py:311-497; algorithms 3-6 of Vo, Atwood and Drakopoulou,
// Opt. Express 26 (2018)) for gfx950, on the float32 transmission chunk
// res (nproj, nz, n) in place:
//   dead stripes  (_rs_dead3, :438-465): roughness of every detector column
//                 (sum over angles of |t - 10-angle running mean|) over its
//                 detector median -> stripe test -> linear interpolation of
//                 the located columns between their unmasked neighbours;
//   large stripes (_rs_large3, :404-435): columns sorted over the angles, the
//                 sorted image median-filtered along the detector; the ratio
//                 of their central-rank means locates stripes; every column
//                 is divided by that ratio and located ones take the smoothed
//                 sorted values in their own angle order;
//   small stripes (_rs_sort3, :389-401): sort, median along the detector
//                 (dim 1) or over a window of detector and rank (dim 2),
//                 unsort, for every column.
// The stripe test (_detect_stripe_batch, :352-375) fits a line through the
// middle half of the sorted ratios of a slice, in FP64 here (float32 in the
// reference).
//
// Layout: the chunk is transposed once into T[z][x][p] (each detector column
// a contiguous row of nproj angles), worked on there, and transposed back.
// Column sorts are bitonic networks in LDS over 64-bit keys (order-preserving
// float bits, angle index), i.e. stable as the reference's argsort; the
// detector medians keep, per lane, its sorted window in LDS and slide it one
// insert/remove at a time.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "common.h"
#include "preproc.h"

using namespace tomo;

namespace {

__device__ __forceinline__ uint32_t fkey(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fback(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// half-sample symmetric reflection into [0, N) (scipy/cupyx 'reflect')
__device__ __forceinline__ int refl(int m, int N) {
  if ((unsigned)m < (unsigned)N) return m;
  const int P = 2 * N;
  int mm = m % P;
  if (mm < 0) mm += P;
  return mm >= N ? P - 1 - mm : mm;
}

// FP64 sum over the workgroup (blockDim.x a multiple of 64, <= 1024);
// every thread gets the total.
__device__ double block_sum(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// T[z][x][p] <-> res[p][z][x] through 64 x 64 LDS tiles (coalesced on both
// sides).  TO_COLS: res -> T; else T -> res.
template <bool TO_COLS>
__global__ __launch_bounds__(256) void k_vo_transpose(const float *__restrict__ src,
                                                      float *__restrict__ dst, int nproj,
                                                      int nz, int n) {
  __shared__ float tile[64][65];
  const int z = blockIdx.z, p0 = blockIdx.y * 64, x0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (TO_COLS) {
    for (int i = ty; i < 64; i += 4) {
      const int p = p0 + i, x = x0 + tx;
      if (p < nproj && x < n) tile[i][tx] = src[((size_t)p * nz + z) * n + x];
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
      const int x = x0 + i, p = p0 + tx;
      if (p < nproj && x < n) dst[((size_t)z * n + x) * nproj + p] = tile[tx][i];
    }
  } else {
    for (int i = ty; i < 64; i += 4) {
      const int x = x0 + i, p = p0 + tx;
      if (p < nproj && x < n) tile[tx][i] = src[((size_t)z * n + x) * nproj + p];
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
      const int p = p0 + i, x = x0 + tx;
      if (p < nproj && x < n) dst[((size_t)p * nz + z) * n + x] = tile[i][tx];
    }
  }
}

// diff[row] = sum_p |t[p] - mean(t[p-5 .. p+4])| ('reflect'), one workgroup
// per column row of T (remove_stripe.py:442-443: uniform_filter1d(tomo, 10,
// axis=0), then the sum over angles).  The running mean is formed in float32
// as the reference's float32 filter output; the sum is FP64.
__global__ __launch_bounds__(256) void k_vo_roughness(const float *__restrict__ T,
                                                      float *__restrict__ diff, int nproj) {
  extern __shared__ float row[];
  __shared__ double red[4];
  const size_t ri = blockIdx.x;
  const float *src = T + ri * nproj;
  for (int p = threadIdx.x; p < nproj; p += 256) row[p] = src[p];
  __syncthreads();
  double acc = 0.0;
  for (int p = threadIdx.x; p < nproj; p += 256) {
    float s = 0.f;
#pragma unroll
    for (int j = -5; j < 5; ++j) s += row[refl(p + j, nproj)];
    s = s / 10.f;
    acc += fabs((double)(row[p] - s));
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) diff[ri] = (float)acc;
}

// first index i in [0, w] of the lane's sorted window with key >= k
__device__ __forceinline__ int lower_bound_win(const uint32_t *wv, int w, uint32_t k) {
  int lo = 0, hi = w;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (wv[mid * 64] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// window[pn] = key after moving window[po+1 .. pn] down by one (pn >= po)
__device__ __forceinline__ void shift_down(uint32_t *wv, int po, int pn, uint32_t key) {
  int i = po;
  for (; i + 4 <= pn; i += 4) {  // four independent loads, then the stores
    const uint32_t a0 = wv[(i + 1) * 64], a1 = wv[(i + 2) * 64], a2 = wv[(i + 3) * 64],
                   a3 = wv[(i + 4) * 64];
    wv[i * 64] = a0;
    wv[(i + 1) * 64] = a1;
    wv[(i + 2) * 64] = a2;
    wv[(i + 3) * 64] = a3;
  }
  for (; i < pn; ++i) wv[i * 64] = wv[(i + 1) * 64];
  wv[pn * 64] = key;
}

// window[pn] = key after moving window[pn .. po-1] up by one (pn <= po)
__device__ __forceinline__ void shift_up(uint32_t *wv, int po, int pn, uint32_t key) {
  int i = po;
  for (; i - 4 >= pn; i -= 4) {
    const uint32_t a0 = wv[(i - 1) * 64], a1 = wv[(i - 2) * 64], a2 = wv[(i - 3) * 64],
                   a3 = wv[(i - 4) * 64];
    wv[i * 64] = a0;
    wv[(i - 1) * 64] = a1;
    wv[(i - 2) * 64] = a2;
    wv[(i - 3) * 64] = a3;
  }
  for (; i > pn; --i) wv[i * 64] = wv[(i - 1) * 64];
  wv[pn * 64] = key;
}

// Median along the detector: out[(z n + x) R + r] = rank-(w/2) value of
// in[(z n + x') R + r], x' = x - w/2 .. x - w/2 + w - 1 reflected
// (ndimage.median_filter with a window of w along the detector).  One lane
// per r (64 per workgroup), xs consecutive outputs per workgroup: the lane
// keeps its window sorted in LDS (column `lane`, stride 64) and slides it by
// one removal and one insertion per output, both positions found by binary
// search first so that the moves between them are independent loads and
// stores.  The keys leaving and entering are loaded one step ahead.
__global__ __launch_bounds__(64) void k_vo_xmedian(const float *__restrict__ in,
                                                   float *__restrict__ out, int R, int n, int w,
                                                   int xs) {
  extern __shared__ uint32_t win[];
  const int lane = threadIdx.x;
  const int r = blockIdx.x * 64 + lane;
  const int x0 = blockIdx.y * xs;
  const int z = blockIdx.z;
  if (r >= R || x0 >= n) return;  // no barriers below
  const float *base = in + (size_t)z * n * R + r;
  uint32_t *wv = win + lane;
  const int h = w / 2;
  for (int i = 0; i < w; ++i) {  // binary insertion sort of the first window
    const uint32_t k = fkey(base[(size_t)refl(x0 - h + i, n) * R]);
    shift_up(wv, i, lower_bound_win(wv, i, k), k);
  }
  const int x1 = min(x0 + xs, n);
  float *dst = out + (size_t)z * n * R + r;
  float vo = base[(size_t)refl(x0 - h, n) * R];      // leaves after output x0
  float vn = base[(size_t)refl(x0 - h + w, n) * R];  // enters after output x0
  for (int x = x0; x < x1; ++x) {
    dst[(size_t)x * R] = fback(wv[h * 64]);
    if (x + 1 >= x1) break;
    const uint32_t ko = fkey(vo), kn = fkey(vn);
    if (x + 2 < x1) {
      vo = base[(size_t)refl(x + 1 - h, n) * R];
      vn = base[(size_t)refl(x + 1 - h + w, n) * R];
    }
    if (ko == kn) continue;
    const int po = lower_bound_win(wv, w, ko);  // holds ko (present by construction)
    const int lb = lower_bound_win(wv, w, kn);
    if (kn > ko) shift_down(wv, po, lb - 1, kn);  // lb > po
    else shift_up(wv, po, lb, kn);                // lb <= po
  }
}

// #{col[l .. u-1] < t} for a sorted non-empty run (u - l <= 127): branch-free
// binary search in seven fixed steps, so that the searches of several
// columns overlap
__device__ __forceinline__ int count_below(const uint32_t *col, int l, int u, uint32_t t) {
  int pos = l;
#pragma unroll
  for (int step = 64; step >= 1; step >>= 1) {
    const int q = pos + step;
    const bool ok = q <= u && col[min(q, u) - 1] < t;
    pos = ok ? q : pos;
  }
  return pos - l;
}

// 2-D median of the sorted image (dim 2 of _rs_sort3, remove_stripe.py:
// 397-398: median_filter with a w x w window over detector and rank):
// out[(z n + x) R + r] = rank-(w*w/2) key of S over x' in the reflected
// detector window of x and r' in the reflected rank window of r.  Every
// column of S is sorted over the ranks, so the reflected rank window of a
// column is at most three sorted runs ([a, b] clipped, its reflection below 0
// and above R-1, w <= R): #{key < t} over the window is a sum of binary
// searches, and the rank-(w*w/2) key follows bit by bit (the largest t with
// #{key < t} <= w*w/2).  Workgroup: 64 ranks x 4 detector positions, the
// (4 + w - 1) columns x (64 + w - 1) ranks they need staged in LDS as keys.
__global__ __launch_bounds__(256) void k_vo_median2d(const float *__restrict__ in,
                                                     float *__restrict__ out, int R, int n,
                                                     int w) {
  extern __shared__ uint32_t tile[];
  const int lane = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int r0 = blockIdx.x * 64, x0 = blockIdx.y * 4, z = blockIdx.z;
  const int h = w / 2;
  const int t0 = max(0, r0 - h), t1 = min(R - 1, r0 + 63 - h + w - 1);
  const int rows = t1 - t0 + 1, cols = 4 + w - 1;
  const float *src = in + (size_t)z * n * R;
  for (int idx = threadIdx.x; idx < rows * cols; idx += 256) {
    const int c = idx / rows, q = idx - c * rows;
    tile[c * rows + q] = fkey(src[(size_t)refl(x0 - h + c, n) * R + t0 + q]);
  }
  __syncthreads();
  const int r = r0 + lane, x = x0 + ty;
  if (r >= R || x >= n) return;
  // rank ranges (absolute) of the reflected window of r
  const int a = r - h, b = r - h + w - 1;
  int lo[3], hi[3], nr = 0;
  lo[nr] = max(a, 0) - t0;
  hi[nr] = min(b, R - 1) - t0 + 1;
  ++nr;
  if (a < 0) {
    lo[nr] = 0 - t0;
    hi[nr] = -a - t0;
    ++nr;
  }
  if (b > R - 1) {
    lo[nr] = 2 * R - 1 - b - t0;
    hi[nr] = R - t0;
    ++nr;
  }
  const int K = (w * w) / 2;
  const uint32_t *col0 = tile + ty * rows;
  uint32_t m = 0;
  for (int bit = 31; bit >= 0; --bit) {
    const uint32_t t = m | (1u << bit);
    int cnt = 0;
#pragma unroll 4
    for (int i = 0; i < w; ++i) cnt += count_below(col0 + i * rows, lo[0], hi[0], t);
    for (int g = 1; g < nr; ++g)
      for (int i = 0; i < w; ++i) cnt += count_below(col0 + i * rows, lo[g], hi[g], t);
    if (cnt <= K) m = t;
  }
  out[((size_t)z * n + x) * R + r] = fback(m);
}

// Median along the detector of one value per column (R = 1, the roughness
// of the dead-stripe test): one thread per output, bitwise selection of the
// rank-(w/2) key over the window staged in LDS (the largest t with
// #{key < t} <= w/2, over 32 bits).
__global__ __launch_bounds__(256) void k_vo_rowmedian(const float *__restrict__ in,
                                                      float *__restrict__ out, int n, int w) {
  extern __shared__ uint32_t seg[];  // 256 + w - 1 keys
  const int z = blockIdx.y, x0 = blockIdx.x * 256, h = w / 2;
  const float *row = in + (size_t)z * n;
  for (int i = threadIdx.x; i < 256 + w - 1; i += 256) seg[i] = fkey(row[refl(x0 - h + i, n)]);
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= n) return;
  const uint32_t *s = seg + threadIdx.x;
  uint32_t m = 0;
  for (int bit = 31; bit >= 0; --bit) {
    const uint32_t t = m | (1u << bit);
    int c = 0;
    for (int i = 0; i < w; ++i) c += s[i] < t;
    if (c <= h) m = t;
  }
  out[(size_t)z * n + x] = fback(m);
}

// Stable sort of one column row of T over the angles: S = sorted values,
// I = their angle indices (argsort + take_along_axis, remove_stripe.py:
// 392-393, 413-414).  Bitonic network over P2 = 256 E 64-bit keys
// (fkey(value) << 32 | angle), E consecutive keys per thread in registers:
// compare-exchange steps of stride < E stay in registers, larger strides
// exchange whole register sets with the partner thread through LDS
// (element i of thread t at i * 256 + t, conflict-free).  With
// mean1 != NULL also the mean of the sorted ranks [r0, r1) (list1, :417),
// FP64 sum, float32 result.
template <int E>
__global__ __launch_bounds__(256) void k_vo_sort(const float *__restrict__ T,
                                                 float *__restrict__ S,
                                                 uint16_t *__restrict__ I,
                                                 float *__restrict__ mean1, int nproj, int r0,
                                                 int r1) {
  extern __shared__ unsigned long long key[];
  __shared__ double red[4];
  constexpr int P2 = 256 * E;
  const size_t ri = blockIdx.x;
  const float *src = T + ri * nproj;
  const int t = threadIdx.x;
  unsigned long long v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int g = t * E + i;
    v[i] = g < nproj ? ((unsigned long long)fkey(src[g]) << 32) | (unsigned)g : ~0ull;
  }
  for (int k = 2; k <= P2; k <<= 1) {
    for (int j = k >> 1; j >= E; j >>= 1) {  // strides across threads
      const int tp = t ^ (j / E);
      const bool lower = t < tp;
#pragma unroll
      for (int i = 0; i < E; ++i) key[i * 256 + t] = v[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const unsigned long long o = key[i * 256 + tp];
        const bool asc = ((t * E + i) & k) == 0;
        v[i] = (lower == asc) ? (v[i] < o ? v[i] : o) : (v[i] < o ? o : v[i]);
      }
      __syncthreads();
    }
#pragma unroll
    for (int jj = E / 2; jj >= 1; jj >>= 1) {  // strides within a thread
      if (jj < k) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
          if ((i & jj) == 0) {
            const bool asc = ((t * E + i) & k) == 0;
            const unsigned long long a = v[i], b = v[i + jj];
            if ((a > b) == asc) {
              v[i] = b;
              v[i + jj] = a;
            }
          }
        }
      }
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int q = t * E + i;
    if (q < nproj) {
      const float val = fback((uint32_t)(v[i] >> 32));
      S[ri * nproj + q] = val;
      I[ri * nproj + q] = (uint16_t)(v[i] & 0xffffu);
      if (q >= r0 && q < r1) acc += (double)val;
    }
  }
  if (mean1) {
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) mean1[ri] = (float)(acc / (double)(r1 - r0));
  }
}

// mean of ranks [r0, r1) of every row of M (list2, remove_stripe.py:419)
__global__ __launch_bounds__(256) void k_vo_mean(const float *__restrict__ M,
                                                 float *__restrict__ mean2, int nproj, int r0,
                                                 int r1) {
  __shared__ double red[4];
  const size_t ri = blockIdx.x;
  double acc = 0.0;
  for (int q = r0 + threadIdx.x; q < r1; q += 256) acc += (double)M[ri * nproj + q];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) mean2[ri] = (float)(acc / (double)(r1 - r0));
}

// Stripe test of one slice (one workgroup of 1024 per z): f = num / den
// (float32, the reference's listfact), written to fact; the sorted f
// (bitonic, LDS) gives the FP64 least-squares line through the descending
// ranks ndrop .. n-ndrop-2; mask = dilated (1 x 3) test result, and with
// zero_edges the first and last two columns cleared (remove_stripe.py:
// 445-451, 420-423).
__global__ __launch_bounds__(1024) void k_vo_detect(const float *__restrict__ num,
                                                    const float *__restrict__ den,
                                                    float *__restrict__ fact,
                                                    uint8_t *__restrict__ mask, int n, int P2,
                                                    double snr, int zero_edges) {
  extern __shared__ uint32_t sk[];
  __shared__ double red[16];
  __shared__ double par[4];
  const int z = blockIdx.x;
  const float *nz_ = num + (size_t)z * n;
  const float *dz = den + (size_t)z * n;
  for (int x = threadIdx.x; x < P2; x += 1024) {
    if (x < n) {
      const float f = nz_[x] / dz[x];
      fact[(size_t)z * n + x] = f;
      sk[x] = fkey(f);
    } else {
      sk[x] = 0xffffffffu;
    }
  }
  __syncthreads();
  const int half = P2 >> 1;
  for (int k = 2; k <= P2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < half; t += 1024) {
        const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        const int l = i + j;
        const uint32_t a = sk[i], b = sk[l];
        if ((a > b) == ((i & k) == 0)) {
          sk[i] = b;
          sk[l] = a;
        }
      }
      __syncthreads();
    }
  }
  // descending rank i holds ascending entry n-1-i
  const int nd = n / 4;
  const int i0 = nd, i1 = n - nd - 1;  // ranks [i0, i1)
  double sy = 0.0, sxy = 0.0, sx = 0.0, sxx = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 1024) {
    const double y = (double)fback(sk[n - 1 - i]);
    sy += y;
    sxy += (double)i * y;
    sx += (double)i;
    sxx += (double)i * (double)i;
  }
  sy = block_sum(sy, red);
  sxy = block_sum(sxy, red);
  sx = block_sum(sx, red);
  sxx = block_sum(sxx, red);
  if (threadIdx.x == 0) {
    const double m = (double)(i1 - i0);
    const double xm = sx / m, ym = sy / m;
    const double slope = (sxy - m * ym * xm) / (sxx - m * xm * xm);
    const double icpt = ym - slope * xm;
    const double numt1 = icpt + slope * (double)(n - 1);
    const double noise = fmax(fabs(numt1 - icpt), 1e-6);
    const double vmax = (double)fback(sk[n - 1]), vmin = (double)fback(sk[0]);
    const double val1 = fabs(vmax - icpt) / noise, val2 = fabs(vmin - numt1) / noise;
    par[0] = val1 >= snr ? icpt + noise * snr * 0.5 : INFINITY;    // upper (or none)
    par[1] = val2 >= snr ? numt1 - noise * snr * 0.5 : -INFINITY;  // lower (or none)
  }
  __syncthreads();
  const double upper = par[0], lower = par[1];
  uint8_t *flag = reinterpret_cast<uint8_t *>(sk);  // the keys are no longer needed
  for (int x = threadIdx.x; x < n; x += 1024) {
    const double f = (double)fact[(size_t)z * n + x];
    flag[x] = (f > upper || f <= lower) ? 1 : 0;
  }
  __syncthreads();
  for (int x = threadIdx.x; x < n; x += 1024) {
    uint8_t m = flag[x] | (x > 0 ? flag[x - 1] : 0) | (x + 1 < n ? flag[x + 1] : 0);
    if (zero_edges && (x < 2 || x >= n - 2)) m = 0;
    mask[(size_t)z * n + x] = m;
  }
}

// Located dead columns: t[x] = t[l] + (x - l) (t[r] - t[l]) / (r - l) with l, r
// the nearest unmasked columns (remove_stripe.py:452-461; the difference in
// float32, the rest in FP64 as the reference's int64 x float32 arithmetic).
// One workgroup per column row; rows l and r are unmasked, so never written.
__global__ __launch_bounds__(256) void k_vo_interp(float *__restrict__ T,
                                                   const uint8_t *__restrict__ mask, int n,
                                                   int nproj) {
  const size_t ri = blockIdx.x;
  const int z = (int)(ri / n), x = (int)(ri % n);
  const uint8_t *mz = mask + (size_t)z * n;
  if (!mz[x]) return;
  int l = x - 1, r = x + 1;
  while (l > 0 && mz[l]) --l;
  while (r < n - 1 && mz[r]) ++r;
  if (l < 0 || r >= n || mz[l] || mz[r]) return;  // edge columns are never masked here
  const float *tl = T + ((size_t)z * n + l) * nproj;
  const float *tr = T + ((size_t)z * n + r) * nproj;
  float *t = T + ri * nproj;
  const double dx = (double)(x - l), span = (double)(r - l);
  for (int p = threadIdx.x; p < nproj; p += 256) {
    const float d = tr[p] - tl[p];
    t[p] = (float)((double)tl[p] + dx * (double)d / span);
  }
}

// LARGE: located rows take the smoothed sorted values in their angle order,
// the others are divided by the ratio (remove_stripe.py:424-435).
// SORT: every row takes them (:400-401).
template <bool LARGE>
__global__ __launch_bounds__(256) void k_vo_scatter(float *__restrict__ T,
                                                    const float *__restrict__ M,
                                                    const uint16_t *__restrict__ I,
                                                    const uint8_t *__restrict__ mask,
                                                    const float *__restrict__ fact, int nproj) {
  const size_t ri = blockIdx.x;
  float *t = T + ri * nproj;
  if (LARGE && !mask[ri]) {
    const float f = fact[ri];
    for (int p = threadIdx.x; p < nproj; p += 256) t[p] = t[p] / f;
    return;
  }
  const float *m = M + ri * nproj;
  const uint16_t *id = I + ri * nproj;
  for (int q = threadIdx.x; q < nproj; q += 256) t[id[q]] = m[q];
}

int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// The sorts and medians take up to 64 KB of dynamic LDS (kVoMaxProj 64-bit
// keys, kVoMaxN 32-bit keys, kVoMaxWin x 64 window keys), above the default
// limit once their static LDS is added: raise it once per process.
hipError_t vo_lds_setup() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const int bytes = 65536;
    const void *sorts[] = {(const void *)k_vo_sort<1>,  (const void *)k_vo_sort<2>,
                           (const void *)k_vo_sort<4>,  (const void *)k_vo_sort<8>,
                           (const void *)k_vo_sort<16>, (const void *)k_vo_sort<32>};
    for (const void *f : sorts)
      if (err == hipSuccess)
        err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (err == hipSuccess)
      err = hipFuncSetAttribute((const void *)k_vo_detect,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (err == hipSuccess)
      err = hipFuncSetAttribute((const void *)k_vo_xmedian,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  });
  return err;
}

}  // namespace

extern "C" int pre_remove_all_stripe(tomo_pre_plan *P, float *res, double snr, int la_size,
                                     int sm_size, int dim, void *stream) {
  clear_error();
  TOMO_REQUIRE(P != nullptr, "pre_remove_all_stripe: plan is NULL");
  TOMO_REQUIRE(res != nullptr, "pre_remove_all_stripe: res is NULL");
  TOMO_REQUIRE(std::isfinite(snr), "pre_remove_all_stripe: snr must be finite");
  TOMO_REQUIRE(la_size >= 1 && la_size <= kVoMaxWin && sm_size >= 1 && sm_size <= kVoMaxWin,
               "pre_remove_all_stripe: window sizes must be in 1..%d (la %d, sm %d)", kVoMaxWin,
               la_size, sm_size);
  TOMO_REQUIRE(dim == 1 || dim == 2, "pre_remove_all_stripe: dim must be 1 or 2");
  const int nproj = (int)P->nproj, nz = (int)P->nz, n = (int)P->n;
  TOMO_REQUIRE(nproj <= kVoMaxProj, "pre_remove_all_stripe: nproj=%d > %d", nproj, kVoMaxProj);
  TOMO_REQUIRE(n >= 8 && n <= kVoMaxN, "pre_remove_all_stripe: n=%d not in 8..%d", n, kVoMaxN);
  TOMO_REQUIRE(dim == 1 || (sm_size <= 64 && sm_size <= nproj),
               "pre_remove_all_stripe: dim=2 needs sm_size <= 64 and <= nproj (sm %d)", sm_size);
  const size_t cells = (size_t)nz * n * nproj, rows = (size_t)nz * n;
  TOMO_REQUIRE(rows <= 0x7fffffff, "pre_remove_all_stripe: too many columns");
  TOMO_REQUIRE(max_lds_per_block() >= 65536 + 1024,
               "pre_remove_all_stripe: needs 65 KB of LDS per workgroup");
  {
    const hipError_t e = vo_lds_setup();
    if (e != hipSuccess) {
      set_error("pre_remove_all_stripe: hipFuncSetAttribute failed: %s", hipGetErrorString(e));
      return TOMO_EHIP;
    }
  }
  auto &V = P->vo;
  if (!V.T.p) {
    int rc = V.T.alloc(sizeof(float) * cells);
    if (!rc) rc = V.S.alloc(sizeof(float) * cells);
    if (!rc) rc = V.M.alloc(sizeof(float) * cells);
    if (!rc) rc = V.I.alloc(sizeof(uint16_t) * cells);
    if (!rc) rc = V.rowf.alloc(sizeof(float) * rows * 5);
    if (!rc) rc = V.mask.alloc(rows);
    if (rc) {
      V.release();
      return rc;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  float *T = V.T.as<float>(), *S = V.S.as<float>(), *M = V.M.as<float>();
  uint16_t *I = V.I.as<uint16_t>();
  float *diff = V.rowf.as<float>(), *bck = diff + rows, *mean1 = diff + 2 * rows,
        *mean2 = diff + 3 * rows, *fact = diff + 4 * rows;
  uint8_t *mask = V.mask.as<uint8_t>();
  const dim3 tgrid((unsigned)ceil_div(n, 64), (unsigned)ceil_div(nproj, 64), (unsigned)nz);
  const int P2 = pow2_at_least(nproj), N2 = pow2_at_least(n);
  // rs_large: ndrop = int(0.5 * drop_ratio * nproj), drop_ratio 0.1
  const int ndrop = (int)(0.5 * 0.1 * (double)nproj);
  const int r0 = ndrop, r1 = nproj - ndrop;
  auto xmedian = [&](const float *in, float *out, int R, int w) {
    if (R == 1) {
      hipLaunchKernelGGL(k_vo_rowmedian, dim3((unsigned)ceil_div(n, 256), (unsigned)nz),
                         dim3(256), sizeof(uint32_t) * (256 + w - 1), st, in, out, n, w);
      return;
    }
    const int xs = 128;
    const dim3 g((unsigned)ceil_div(R, 64), (unsigned)ceil_div(n, xs), (unsigned)nz);
    hipLaunchKernelGGL(k_vo_xmedian, g, dim3(64), sizeof(uint32_t) * 64 * w, st, in, out, R, n,
                       w, xs);
  };
  auto detect = [&](const float *num, const float *den, int zero_edges) {
    hipLaunchKernelGGL(k_vo_detect, dim3((unsigned)nz), dim3(1024), sizeof(uint32_t) * N2, st,
                       num, den, fact, mask, n, N2, snr, zero_edges);
  };
  const int E = P2 <= 256 ? 1 : P2 / 256;  // keys per thread of the column sort
  auto sort = [&](float *m1) {
    const dim3 g((unsigned)rows), b(256);
    const size_t lds = sizeof(unsigned long long) * 256 * E;
    switch (E) {
      case 1: hipLaunchKernelGGL(k_vo_sort<1>, g, b, lds, st, T, S, I, m1, nproj, r0, r1); break;
      case 2: hipLaunchKernelGGL(k_vo_sort<2>, g, b, lds, st, T, S, I, m1, nproj, r0, r1); break;
      case 4: hipLaunchKernelGGL(k_vo_sort<4>, g, b, lds, st, T, S, I, m1, nproj, r0, r1); break;
      case 8: hipLaunchKernelGGL(k_vo_sort<8>, g, b, lds, st, T, S, I, m1, nproj, r0, r1); break;
      case 16:
        hipLaunchKernelGGL(k_vo_sort<16>, g, b, lds, st, T, S, I, m1, nproj, r0, r1);
        break;
      default:
        hipLaunchKernelGGL(k_vo_sort<32>, g, b, lds, st, T, S, I, m1, nproj, r0, r1);
        break;
    }
  };
  {
    ProfScope ps_("k_vo_transpose", st);
    hipLaunchKernelGGL(k_vo_transpose<true>, tgrid, dim3(256), 0, st, res, T, nproj, nz, n);
  }
  TOMO_LAUNCH_CHECK();
  {
    // dead stripes
    ProfScope ps_("k_vo_dead", st);
    hipLaunchKernelGGL(k_vo_roughness, dim3((unsigned)rows), dim3(256), sizeof(float) * nproj,
                       st, T, diff, nproj);
    xmedian(diff, bck, 1, la_size);
    detect(diff, bck, 1);
    hipLaunchKernelGGL(k_vo_interp, dim3((unsigned)rows), dim3(256), 0, st, T, mask, n, nproj);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_vo_sort", st);
    sort(mean1);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_vo_xmedian", st);
    xmedian(S, M, nproj, la_size);
  }
  TOMO_LAUNCH_CHECK();
  {
    // large stripes: ratio, test, normalisation and replacement
    ProfScope ps_("k_vo_large", st);
    hipLaunchKernelGGL(k_vo_mean, dim3((unsigned)rows), dim3(256), 0, st, M, mean2, nproj, r0,
                       r1);
    detect(mean1, mean2, 0);
    hipLaunchKernelGGL(k_vo_scatter<true>, dim3((unsigned)rows), dim3(256), 0, st, T, M, I, mask,
                       fact, nproj);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_vo_sort", st);
    sort(nullptr);
  }
  TOMO_LAUNCH_CHECK();
  if (dim == 1) {
    ProfScope ps_("k_vo_xmedian", st);
    xmedian(S, M, nproj, sm_size);
  } else {
    ProfScope ps_("k_vo_median2d", st);
    const int w = sm_size, h = w / 2;
    const size_t lds = sizeof(uint32_t) * (4 + w - 1) * (size_t)std::min(nproj, 64 - h + w - 1 + h);
    hipLaunchKernelGGL(k_vo_median2d,
                       dim3((unsigned)ceil_div(nproj, 64), (unsigned)ceil_div(n, 4), (unsigned)nz),
                       dim3(256), lds, st, S, M, nproj, n, w);
  }
  TOMO_LAUNCH_CHECK();
  {
    ProfScope ps_("k_vo_unsort", st);
    hipLaunchKernelGGL(k_vo_scatter<false>, dim3((unsigned)rows), dim3(256), 0, st, T, M, I,
                       mask, fact, nproj);
    hipLaunchKernelGGL(k_vo_transpose<false>, tgrid, dim3(256), 0, st, T, res, nproj, nz, n);
  }
  TOMO_LAUNCH_CHECK();
  return TOMO_OK;
}
