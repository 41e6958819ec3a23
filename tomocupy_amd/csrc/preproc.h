// Internal: the preprocessing plan shared by preproc.hip and stripe.hip.
#pragma once

#include <string>
#include <vector>

#include "common.h"

namespace tomo {

// One level of the Fourier-wavelet stripe removal: analysis input H x W,
// bands Ho x Wo; the notch keeps nk angular frequencies of the stripe band.
struct FwLevel {
  int H = 0, W = 0, Ho = 0, Wo = 0;
  int nk = 0;
  size_t xoff = 0;  // offset of this level's notch coefficients in notchX (float2)
  DevBuf tw;   // (nk, Ho) float2: cos/sin(2 pi k t / Ho)
  DevBuf gk;   // (nk,) float: c_k g_k / Ho
  DevBuf ll, lh, hl, hh;  // (nz, Ho, Wo) each
};

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
  tomo::DevBuf rowlo, rowhi, synlo, synhi, syn[2], notch, notchX;
};
