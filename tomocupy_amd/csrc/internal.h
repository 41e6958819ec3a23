// Internal: the filter plan, shared by filter.hip and fourierrec.hip (the
// fused filter + backprojection entry point reads its derived kernels).
#pragma once

#include "common.h"
#include "fft_lds.h"

struct tomo_flt_plan {
  int64_t n = 0, ne = 0, nproj = 0, nz = 0;
  int dtype = 0;           // storage of the filtered rows: TOMO_F32, TOMO_F16 or TOMO_BF16
  tomo::FftPlan fft;       // ne
  tomo::FftPlan fft2n;     // 2n (fused path)
  tomo::DevBuf w;          // (nz, ne/2+1) complex, filled by flt_filter_center
  int threads = 256;
  size_t lds = 0;
  // fused path: per slice, the spectrum of the impulse response wrapped to
  // length 2n, and the edge-replication responses HL/HR (see filter.hip)
  // H2: (2, nz, 2n) complex, the wrapped filter spectra and the same for a
  // differenced input; Hd: FP64 impulse response (nz, ne)
  tomo::DevBuf H2, HL, HR, P, Hd;
  int prepared = 0;  // H2/HL/HR hold the kernels of the last flt_prepare
};

namespace tomo {
// Builds H2/HL/HR of `plan` for filter `wfilter` (device float32 (ne/2+1,)),
// rotation centre and per-slice shifts `sht` (device float32 (nz,) or NULL).
int flt_prepare_2n(tomo_flt_plan *plan, const float *wfilter, double center, const float *sht,
                   hipStream_t st);
}  // namespace tomo
