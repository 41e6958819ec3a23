// Internal: the preprocessing plan shared by preproc.hip and stripe.hip.
#pragma once

#include <string>
#include <vector>

#include "common.h"

struct tomo_pre_plan;

namespace tomo {

constexpr int kMaxDezinger = 9;  // largest median window of pre_remove_outliers
constexpr int kVoMaxProj = 8192;  // vo-all: angles of a column sorted in LDS
constexpr int kVoMaxWin = 255;    // vo-all: largest detector median window
constexpr int kVoMaxN = 16384;    // vo-all: detector width of a slice's stripe test

// Scratch of pre_remove_all_stripe (allocated on the plan's first call)
struct VoBufs {
  DevBuf T, S, M;  // (nz, n, nproj) float: columns, sorted columns, smoothed
  DevBuf I;        // (nz, n, nproj) uint16: angle of each sorted rank
  DevBuf rowf;     // 5 x (nz, n) float: roughness, its median, two means, ratio
  DevBuf mask;     // (nz, n) uint8: located columns
  void release() {
    T.release();
    S.release();
    M.release();
    I.release();
    rowf.release();
    mask.release();
  }
};

// One level of the Fourier-wavelet stripe removal: analysis input H x W,
// bands Ho x Wo; the notch keeps nk angular frequencies of the stripe band.
struct FwLevel {
  int H = 0, W = 0, Ho = 0, Wo = 0;
  int nk = 0;
  int KM = 8;   // notch frequencies a kernel instance holds (8, 16 or 32 >= nk)
  int R = 8;    // output rows per wave of the column analysis
  int nrb = 0;  // row blocks of the column analysis
  DevBuf tw;    // (nk, Ho) float2: cos/sin(2 pi k t / Ho)
  DevBuf gk;    // (nk,) float: c_k g_k / Ho
  DevBuf ll;    // (nz, Ho, Wo): low band = input of the next level
  DevBuf part;  // (nz, nrb, KM, Wo) float2: partial notch DFTs of the stripe band
  DevBuf X;     // (nz, KM, Wo) float2: notch coefficients, kept until the synthesis
};

// dark/flat correction of one count (proc_functions.py:67-74): flat0 =
// flat mean (or the linear interpolation v of the two halves' means,
// :62-66) * float32(1 + 1e-5) - dark0; result (x - dark0) / flat0
__device__ __forceinline__ float darkflat(float x, float d0, float fa, float fb, float v,
                                          int flat_linear) {
  const float eps = 1.0f + 1e-5f;  // float32(1 + 1e-5), proc_functions.py:71
  float f0 = flat_linear ? (1.0f - v) * fa + v * fb : fa;
  f0 = f0 * eps - d0;
  return (x - d0) / f0;
}

__device__ __forceinline__ float minus_log(float x) {
  // x <= 0 -> 1; -log; NaN -> 6; +-inf -> 0 (proc_functions.py:76-83)
  if (x <= 0.f) x = 1.f;
  float r = -logf(x);
  if (isnan(r)) r = 6.0f;
  if (isinf(r)) r = 0.0f;
  return r;
}

// dark0 = mean(dark), flat means, in the plan's scratch (k_pre_means)
int pre_launch_means(tomo_pre_plan *P, const uint16_t *dark, const uint16_t *flat,
                     double bright_ratio, int flat_linear, hipStream_t st);

}  // namespace tomo

struct tomo_pre_plan {
  int64_t nproj = 0, nz = 0, n = 0, ndark = 0, nflat = 0;
  tomo::DevBuf dark0, flatA, flatB;
  // Fourier-wavelet scratch, (re)built when the settings change
  std::string fw_name;
  int fw_levels = 0;
  double fw_sigma = 0.0;
  int fw_L = 0;
  std::vector<tomo::FwLevel> lv;
  tomo::DevBuf rowlo, rowhi, synlo, synhi, syn[2];
  tomo::DevBuf outl;  // medians of pre_remove_outliers, allocated on first use
  tomo::DevBuf ti;    // column means and corrections of pre_remove_stripe_ti (first use)
  tomo::VoBufs vo;    // scratch of pre_remove_all_stripe (first use)
};
