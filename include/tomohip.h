/*
 * tomohip.h -- C ABI of libtomohip.so, the MI355X (gfx950) implementation of
 * tomocupy's fourierrec filtered-backprojection path.
 *
 * Conventions
 *   - Every entry point returns TOMO_OK (0) or a negative TOMO_E* status and
 *     never throws.  tomo_last_error() returns a thread-local message for the
 *     last failure on the calling thread.
 *   - Arguments are validated before any HIP call, so invalid sizes, dtypes
 *     and null pointers are reported without touching a GPU.
 *   - Array arguments are device pointers (hipMalloc'ed or torch tensors),
 *     row-major and contiguous, unless a name says otherwise.  `stream` is a
 *     hipStream_t passed as void* (NULL = the legacy default stream).
 *   - A plan owns its scratch buffers and tables; one call per plan may be in
 *     flight at a time (calls on one stream are ordered).  Distinct plans may
 *     be used from distinct threads.
 *   - dtype: TOMO_F32 (float32 storage and arithmetic), TOMO_F16 / TOMO_BF16
 *     (IEEE half / bfloat16 storage of the sinograms, spectra, grids and
 *     images; float32 arithmetic and accumulation; FP64-derived tables).
 *     Arrays documented as "dtype" have the element type of the plan.
 */
#ifndef TOMOHIP_H
#define TOMOHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TOMO_OK = 0,
  TOMO_EINVAL = -1,     /* invalid argument (checked before any HIP call) */
  TOMO_EHIP = -2,       /* a HIP runtime call or kernel launch failed */
  TOMO_ENOMEM = -3,     /* device or host allocation failed */
  TOMO_EUNSUPPORTED = -4 /* valid request this build does not implement */
};

enum { TOMO_F32 = 0, TOMO_F16 = 1, TOMO_BF16 = 2 };

/* Thread-local message describing the last failure on this thread ("" if none). */
const char *tomo_last_error(void);
/* Library version string. */
const char *tomo_version(void);

/*
 * Kernel timing for measurement (no counterpart in the reference, which has no
 * profiling hooks).  While enabled, each kernel launch is bracketed by two HIP
 * events on its own stream; tomo_prof_read() waits for the recorded events and
 * returns the accumulated milliseconds and launch count of one kernel by name
 * (e.g. "k_fr_grid").  tomo_prof_only(name) restricts the timing to one
 * kernel name (NULL or "" = every kernel), so that a measured run records two
 * events per launch of that kernel only.  tomo_prof_reset() clears the
 * accumulators.
 */
int tomo_prof_enable(int on);
int tomo_prof_only(const char *kernel);
int tomo_prof_reset(void);
int tomo_prof_read(const char *kernel, double *total_ms, int64_t *launches);

/*
 * Inner product for the iterative solvers of the differentiation form (no
 * counterpart in the reference): out[0] = sum_i a[i] b[i] over n float32
 * values, accumulated in FP64 in a fixed order (deterministic).  out is a
 * device array of at least 1025 doubles: out[1..1024] hold the per-block
 * partial sums.  Asynchronous on `stream`.
 */
int tomo_dot_f64(const float *a, const float *b, int64_t n, double *out, void *stream);

/*
 * The vector updates of one CG iteration of the differentiation form's
 * least-squares solver (no counterpart in the reference), each one pass over
 * device memory with the step sizes read on the device (no host wait):
 *
 * tomo_cg_step: alpha = rr[0] / pq[0] (0 when pq[0] <= 0), rounded to
 * float32; x += alpha p; r -= alpha q; then out[0] = sum_i r[i]^2 in FP64 in
 * exactly the order of tomo_dot_f64 (so out[0] equals tomo_dot_f64(r, r)
 * bit for bit).  out: at least 1025 doubles, as for tomo_dot_f64; it must not
 * alias rr or pq.
 *
 * tomo_cg_direction: beta = rr_new[0] / rr[0] (0 when rr[0] <= 0), rounded
 * to float32; p = r + beta p.
 */
int tomo_cg_step(float *x, float *r, const float *p, const float *q, int64_t n,
                 const double *rr, const double *pq, double *out, void *stream);
int tomo_cg_direction(float *p, const float *r, int64_t n, const double *rr,
                      const double *rr_new, void *stream);

/* ------------------------------------------------------------------------ */
/* USFFT backprojector                                                      */
/* ------------------------------------------------------------------------ */
typedef struct tomo_fr_plan tomo_fr_plan;
typedef struct tomo_flt_plan tomo_flt_plan;

/*
 * Replaces cfunc_fourierrec::cfunc_fourierrec(nproj, nz/2, n, theta)
 * (src/cuda/cfunc_fourierrec.cu:4-50) together with FourierRec.__init__
 * (src/tomocupy/reconstruction/fourierrec.py:49-59).
 *   n        detector width = image size; even, 20..4096, with no prime
 *            factor above 127 (the reference's cuFFT takes any length; e.g.
 *            2448 = 2^4 3^2 17 is accepted)
 *   nproj    number of angles, >= 1
 *   nz       number of REAL slices per call (the reference's ncz), >= 1;
 *            odd nz is accepted and paired with an implicit zero slice
 *   theta    device float32 (nproj,) angles in radians; copied at creation
 *            (the reference keeps a borrowed pointer)
 *   dtype    TOMO_F32, TOMO_F16 or TOMO_BF16 (the reference's float16 module,
 *            cfunc_fourierrecfp16.i); 16-bit plans store their spectra 2^8
 *            above the float32 scaling, so half storage stays in its normal
 *            range for sinograms of the usual (post -log, filtered) magnitude
 *   stream   stream used for the creation-time uploads (synchronised)
 */
int fr_create(tomo_fr_plan **plan, int64_t n, int64_t nproj, int64_t nz,
              const float *theta, int dtype, void *stream);

/*
 * fr_create with tuning options (no counterpart in the reference; used by the
 * plan autotuner, tomocupy_amd/autotune.py).  Both options change only the
 * rounding of the results; each configuration is deterministic.
 *   grid_chunk   max sample-list entries per gridding work item: 0 = default
 *                (512), else a multiple of 64 in [64, 8192]
 *   fft_generic  1 = run-time radix plans for every FFT, 0 = compiled
 *                size-specialised transforms where 2n is 1024..8192
 */
int fr_create_opts(tomo_fr_plan **plan, int64_t n, int64_t nproj, int64_t nz,
                   const float *theta, int dtype, int grid_chunk, int fft_generic,
                   void *stream);

/*
 * Replaces FourierRec.backprojection(obj, data, stream)
 * (fourierrec.py:61-68) and cfunc_fourierrec::backprojection
 * (cfunc_fourierrec.cu:67-93).
 *   obj   output, device dtype (nz, n, n)
 *   data  input filtered sinograms, device dtype (nz, nproj, n); NOT modified
 *         (the reference's Python layer protects it with a packed copy)
 * The real-pair packing data[c] + i*data[c+nz/2] and the unpacking are done
 * inside the kernels.  Deterministic: no floating-point atomics.
 */
int fr_backprojection(tomo_fr_plan *plan, void *obj, const void *data, void *stream);

/*
 * fbp_filter_center followed by backprojection in one pass over the data
 * (backproj_functions.py:93-109 then fourierrec.py:61-68, as rec.py:175-176
 * calls them), with `flt` created for the same n, nproj, nz.  The
 * edge-replicated filter of length ne is applied exactly as a length-2n
 * circular convolution plus two edge terms; the filtered sinograms are never
 * written to memory.  Unlike flt_filter_center, `data` is NOT modified.
 *   obj      output, device dtype (nz, n, n)
 *   data     device dtype (nz, nproj, n) sinograms (after -log)
 *   wfilter, center, sht  as flt_filter_center
 */
int fr_filter_backprojection(tomo_fr_plan *plan, tomo_flt_plan *flt, void *obj,
                             const void *data, const float *wfilter, double center,
                             const float *sht, void *stream);

/*
 * fr_filter_backprojection in two steps, so that a caller filtering many
 * chunks with one filter, centre and shift computes the filter kernels once
 * (no counterpart in the reference): flt_prepare builds them in `flt`
 * (same arguments as flt_filter_center), fr_filter_backprojection_prepared
 * uses the last prepared ones.  Stream-ordered: prepare on the stream (or
 * before the streams) the backprojections run on.
 */
int flt_prepare(tomo_flt_plan *flt, const float *wfilter, double center, const float *sht,
                void *stream);
int fr_filter_backprojection_prepared(tomo_fr_plan *plan, tomo_flt_plan *flt, void *obj,
                                      const void *data, void *stream);

/*
 * Batches of independent requests (no counterpart in the reference, which
 * reconstructs one chunk at a time: rec.py:167-189).  A batch groups the
 * float32 plans of `count` requests (sizes may differ, each plan in at most
 * one live batch at a time); fr_batch_filter_backprojection then does
 * fr_filter_backprojection_prepared for every request r,
 *   objs[r] <- filtered backprojection of datas[r] with flts[r],
 * but grids all requests in ONE gridding launch, whose work items are sorted
 * longest first over the whole batch, so that the small requests of a
 * mixed batch fill the machine beside the large ones.  Each result equals
 * that of fr_filter_backprojection_prepared on the same plan bit for bit.
 * The call is ordered on `stream` (it starts after the work queued there
 * before it, and work queued after it waits for all of it; capturable into a
 * graph); inside, the per-request stages fan out over 4 streams the batch
 * owns.  The plans' scratch is used; plans must not run elsewhere meanwhile.
 *   batch  out: the batch handle (fr_batch_destroy frees it, not the plans)
 *   plans  host array of `count` fr plans, float32, distinct
 *   flts   host array of prepared filter plans (flt_prepare), one per plan
 *   objs   host array of device float32 (nz, n, n) outputs
 *   datas  host array of device float32 (nz, nproj, n) inputs, NOT modified
 */
typedef struct tomo_fr_batch tomo_fr_batch;
int fr_batch_create(tomo_fr_batch **batch, tomo_fr_plan *const *plans, int count);
int fr_batch_filter_backprojection(tomo_fr_batch *batch, tomo_flt_plan *const *flts,
                                   void *const *objs, const void *const *datas, void *stream);
void fr_batch_destroy(tomo_fr_batch *batch);

/*
 * Adjoint of fr_backprojection (USFFT forward projection); no counterpart in
 * the reference's tomography path, structured like cfunc_usfft2d::fwd
 * (src/cuda/cfunc_usfft2d.cu:53-70).  For real arrays
 * sum(fr_backprojection(d) * o) == sum(d * fr_forward(o)) to rounding.
 *   data  output, device float32 (nz, nproj, n)
 *   obj   input, device float32 (nz, n, n); NOT modified
 * float32 plans only (TOMO_EUNSUPPORTED for 16-bit plans); every n that
 * fr_create accepts.
 */
int fr_forward(tomo_fr_plan *plan, void *data, const void *obj, void *stream);

/* Replaces cfunc_fourierrec::free (cfunc_fourierrec.cu:56-65). NULL is a no-op. */
void fr_destroy(tomo_fr_plan *plan);

/* Plan sizes (the SWIG read-only attributes n, nproj, nz, cfunc_fourierrec.i:11-14). */
int fr_get_sizes(const tomo_fr_plan *plan, int64_t *n, int64_t *nproj, int64_t *nz);

/* ------------------------------------------------------------------------ */
/* FBP filter                                                               */
/* ------------------------------------------------------------------------ */

/*
 * Replaces cfunc_filter::cfunc_filter(nproj, nz, ne) (src/cuda/cfunc_filter.cu:4-34)
 * and FBPFilter.__init__ (src/tomocupy/reconstruction/fbp_filter.py:47-52).
 *   n      unpadded detector width; even, >= 20
 *   ne     padded FFT length (4n in float32, backproj_functions.py:50-54); ne >= n, even,
 *          <= 16384, with no prime factor above 127
 *   nproj, nz  rows per call
 *   dtype  storage of the filtered rows: TOMO_F32, TOMO_F16 (cfunc_filterfp16.i) or
 *          TOMO_BF16; the arithmetic is float32
 */
int flt_create(tomo_flt_plan **plan, int64_t n, int64_t ne, int64_t nproj, int64_t nz,
               int dtype);

/*
 * Replaces cfunc_filter::filter(g, w, stream) (cfunc_filter.cu:49-61, mulw
 * kernels_filter.cuh:4-18) as called by FBPFilter.filter (fbp_filter.py:54-58).
 *   g  device dtype (nz, nproj, ne), filtered in place: C2R(R2C(g) * w[z]),
 *      unnormalised (the 1/ne is folded into the filter by calc_filter)
 *   w  device complex64 (nz, ne/2+1) as interleaved float pairs
 * The imaginary part of w at the Nyquist bin is ignored, as a C2R does.
 */
int flt_filter(tomo_flt_plan *plan, void *g, const float *w, void *stream);

/*
 * Replaces BackprojFunctions.fbp_filter_center(data, sht)
 * (src/tomocupy/reconstruction/backproj_functions.py:93-109) in one pass:
 * edge-replicate padding to ne, per-slice centre phase ramp
 * w[z,k] = wfilter[k] * exp(-2 pi i (-center + sht[z] + n/2) k/ne) computed in
 * FP64, filtering and cropping, without the (nz, nproj, ne) temporary.
 *   data     device dtype (nz, nproj, n), filtered in place
 *   wfilter  device float32 (ne/2+1,) from flt_calc_filter
 *   center   rotation centre (pixels)
 *   sht      device float32 (nz,) per-slice extra shift, or NULL for zeros
 */
int flt_filter_center(tomo_flt_plan *plan, void *data, const float *wfilter,
                      double center, const float *sht, void *stream);

void flt_destroy(tomo_flt_plan *plan);

/*
 * Replaces FBPFilter.calc_filter(name) (fbp_filter.py:60-88) with _wint
 * (fbp_filter.py:90-118), evaluated on the host in FP64.  Host function; needs
 * no GPU.  name: none, ramp, shepp, cosine, cosine2, hamming, hann, parzen.
 *   out  host float64 (ne/2+1,) (the reference casts to float32 afterwards)
 */
int flt_calc_filter(int64_t ne, const char *name, double *out);

/* ------------------------------------------------------------------------ */
/* Chunk preprocessing                                                      */
/* ------------------------------------------------------------------------ */
typedef struct tomo_pre_plan tomo_pre_plan;

/*
 * Chunk preprocessing plan for raw uint16 chunks in tomocupy's layout:
 * data (nproj, nz, n), dark (ndark, nz, n), flat (nflat, nz, n)
 * (rec.py:80-84).  Owns the dark/flat mean scratch.
 */
int pre_create(tomo_pre_plan **plan, int64_t nproj, int64_t nz, int64_t n,
               int64_t ndark, int64_t nflat);

/*
 * ProcFunctions.proc_sino + proc_proj + the swap to sinogram order
 * (src/tomocupy/processing/proc_functions.py:55-83, 140-198; rec.py:172-174)
 * without stripe removal:
 *   out[z, p, x] = -log((data[p,z,x] - dark0[z,x]) / flat0[z,x])   (minus_log != 0)
 * with dark0 = mean(dark), flat0 = mean(flat)/bright_ratio*(1+1e-5) - dark0,
 * and minus_log's clamp/NaN rules.  flat_linear != 0 interpolates the flat per
 * projection as proc_functions.py:62-66.
 * Arithmetic is float32 (the reference computes in its dtype, float16 for
 * its half-precision runs); out_dtype only sets the storage of `out`.
 *   out        device (nz, nproj, n) of out_dtype (TOMO_F32, TOMO_F16, TOMO_BF16)
 */
int pre_process(tomo_pre_plan *plan, const uint16_t *data, const uint16_t *dark,
                const uint16_t *flat, void *out, int out_dtype, double bright_ratio,
                int flat_linear, int minus_log, void *stream);

/*
 * pre_process without a plan and without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * dark/flat means live in a static device array of each GPU (12 MiB, held
 * until the process ends), so nz*n <= 1048576 (larger sizes return -1); the
 * sizes of pre_create and the arguments and results of pre_process
 * otherwise.  Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (different streams, different threads) race.
 */
int pre_process_noalloc(int64_t nproj, int64_t nz, int64_t n, int64_t ndark, int64_t nflat,
                        const uint16_t *data, const uint16_t *dark, const uint16_t *flat,
                        void *out, int out_dtype, double bright_ratio, int flat_linear,
                        int minus_log, void *stream);

/*
 * Dark/flat correction only, in tomocupy's projection layout:
 *   res (nproj, nz, n) float32 = (data - dark0)/flat0  (proc_functions.py:55-74)
 */
int pre_darkflat(tomo_pre_plan *plan, const uint16_t *data, const uint16_t *dark,
                 const uint16_t *flat, float *res, double bright_ratio, int flat_linear,
                 void *stream);

/*
 * pre_process and pre_darkflat on float32 counts: the binned frames of
 * pre_bin (with binning > 0 the reference reads its frames as float and sums
 * them, dataio/reader.py:96-99, 350-355; the dark/flat correction is then
 * computed on those sums, proc_functions.py:55-74).  Same arguments and
 * results otherwise.
 */
int pre_process_f32(tomo_pre_plan *plan, const float *data, const float *dark,
                    const float *flat, void *out, int out_dtype, double bright_ratio,
                    int flat_linear, int minus_log, void *stream);
int pre_darkflat_f32(tomo_pre_plan *plan, const float *data, const float *dark,
                     const float *flat, float *res, double bright_ratio, int flat_linear,
                     void *stream);

/*
 * Binning, utils.downsample(data, binning) (src/tomocupy/utils.py:157-167) as
 * the reader applies it to projections, darks and flats
 * (dataio/reader.py:350-389): dst[r, z, x] = sum of the 2^b x 2^b block
 * src[r, z 2^b + a, x 2^b + c], a, c in [0, 2^b).
 *   src  device uint16 (nrows, nz << binning, n << binning), nz and n of the plan
 *   dst  device float32 (nrows, nz, n); exact sums (< 2^24 for binning <= 3)
 * binning in 0..3 (config.py:154-158); 0 converts the counts to float32.
 */
int pre_bin(tomo_pre_plan *plan, const uint16_t *src, float *dst, int64_t nrows, int binning,
            void *stream);

/*
 * minus_log (proc_functions.py:76-83) fused with the swap to sinogram order:
 *   out[z,p,x] = -log(res[p,z,x]) with x<=0 -> 0, NaN -> 6, +-inf -> 0
 * (minus_log != 0), or a plain transpose (minus_log == 0); `out` is stored
 * as out_dtype (TOMO_F32, TOMO_F16, TOMO_BF16).
 */
int pre_minus_log_swap(tomo_pre_plan *plan, const float *res, void *out, int out_dtype,
                       int minus_log, void *stream);

/*
 * pre_minus_log_swap reading rows z0 .. z0+lz-1 of every projection of a
 * resident volume vol (nproj, ld, n) float32 (the recon_steps layout,
 * rec_steps.py:114-125 and backproj_parallel.py:338-341: -log of proc_proj,
 * then the swap of the z-chunk to sinograms); rows lz .. nz-1 of out
 * (nz, nproj, n) are zero (the tail of a partial chunk).
 */
int pre_minus_log_swap_rows(tomo_pre_plan *plan, const float *vol, int64_t ld, int64_t z0,
                            int64_t lz, void *out, int out_dtype, int minus_log, void *stream);

/*
 * ProcFunctions.remove_outliers (src/tomocupy/processing/proc_functions.py:96-107),
 * applied by proc_sino to the projections, darks and flats (:146-148):
 *   f = median over the width x 1 x width window (angle, slice, detector)
 *   data = (data > f && data - f > threshold) ? f : data
 * in place on `data` (nrows, nz, n) uint16 of the plan's nz and n, bit-exact
 * with ndimage.median_filter(data, [w, 1, w]) ('reflect' borders, rank
 * w*w/2).  width 0 or 1 does nothing; widths up to 9 are supported.  The
 * median scratch is allocated on the plan's first call.
 */
int pre_remove_outliers(tomo_pre_plan *plan, uint16_t *data, int64_t nrows, int width,
                        int threshold, void *stream);

/*
 * pre_remove_outliers on float32 counts (binned frames, see pre_bin): the
 * same median (exact selection) and replacement rule, data - f compared with
 * threshold in float32 as the reference does for float frames.
 */
int pre_remove_outliers_f32(tomo_pre_plan *plan, float *data, int64_t nrows, int width,
                            int threshold, void *stream);

/*
 * remove_stripe_ti (src/tomocupy/processing/remove_stripe.py:296-309), the
 * Titarenko ring filter, in place on the float32 chunk res (nproj, nz, n):
 * v = mean over angles minus its first column, c = v circularly convolved with
 * gamma_k = beta ((1-beta)/(1+beta))^|k| (minus 1 at k = 0), res += c on the
 * central floor(mask*n) columns.  FP64 means and convolution; even n only
 * (as the reference); beta in (0, 1), mask in [0, 1].  Scratch allocated on
 * the plan's first call.
 */
int pre_remove_stripe_ti(tomo_pre_plan *plan, float *res, double beta, double mask,
                         void *stream);

/*
 * remove_all_stripe (src/tomocupy/processing/remove_stripe.py:311-497), the
 * "vo-all" method of proc_sino (proc_functions.py:157-159; Vo et al. 2018,
 * algorithms 3-6), in place on the float32 chunk res (nproj, nz, n):
 * dead stripes (roughness over its detector median, stripe test, linear
 * interpolation), large stripes (sorted columns, detector median of window
 * la_size, ratio test, normalisation, replacement) and small stripes (sort,
 * detector median of window sm_size, unsort).  Sorts are stable as the
 * reference's argsort; the stripe test's line fit is FP64.
 *   snr        stripe test ratio (config.py vo-all-snr, default 3)
 *   la_size    window of the large-stripe median, 1..255 (default 61)
 *   sm_size    window of the small-stripe median, 1..255 (default 21)
 *   dim        1 (the reference's default): small-stripe median along the
 *              detector; 2: over a sm_size x sm_size window of detector and
 *              rank (sm_size <= 64 and <= nproj)
 * nproj <= 8192, 8 <= n <= 16384.  Scratch (about 14 bytes per sample of the
 * chunk) is allocated on the plan's first call.
 */
int pre_remove_all_stripe(tomo_pre_plan *plan, float *res, double snr, int la_size,
                          int sm_size, int dim, void *stream);

/*
 * ProcFunctions.rotate_proj (src/tomocupy/processing/proc_functions.py:126-138),
 * applied by proc_proj (:182-184): ndimage.rotate(data, angle, reshape=False,
 * order, axes=(2, 1), mode='nearest') on projections src (nproj, ny, nx)
 * float32 into dst (same shape, not in place): dst[p,y,x] = src[p] sampled
 * at R (y, x) + centre - R centre, R = [[cos, sin], [-sin, cos]] of the angle
 * in degrees, centre ((ny-1)/2, (nx-1)/2), neighbours clamped.  order 0
 * (nearest), 1 (bilinear, the reference CLI's default), 2 to 5 (B-spline of
 * that order on the prefiltered plane, edge-padded by 12 as SciPy's spline
 * path).
 *   work  orders 2-5: device float32 of nproj (ny + 24) (nx + 24) values
 *         (the spline coefficients); may be NULL for orders 0, 1
 */
int pre_rotate_proj(const float *src, float *dst, int64_t nproj, int64_t ny, int64_t nx,
                    double angle, int order, float *work, void *stream);

/* ---------------------------------------------------------------------------
 * Phase retrieval (ProcFunctions.proc_proj, proc_functions.py:169-181)
 * ------------------------------------------------------------------------- */
typedef struct tomo_ph_plan tomo_ph_plan;

/*
 * Replaces the set-up of retrieve_phase.paganin_filter / farago_filter
 * (src/tomocupy/processing/retrieve_phase.py:59-141): pad widths
 * (_calc_pad, :166-200: to the next power of two of size + ceil(pi lambda
 * dist / ps^2), split on both sides), reciprocal grid (:236-304) and the
 * filter's maximum, for projections of dy x dz (both even).
 *   method      0 paganin, 1 Gpaganin, 2 farago (config.py retrieve-phase-method)
 *   pixel_size  [cm]; dist [cm]; energy [keV] (proc_functions.py:171-174
 *               converts the CLI's micrometres and millimetres)
 *   alpha       paganin regularisation; db delta/beta (Gpaganin, farago);
 *               W [cm] (Gpaganin)
 *   batch       projections transformed per launch (work buffer
 *               batch * Ny * Nx complex)
 * Padded sizes Ny, Nx must be powers of two <= 8192.
 */
int ph_create(tomo_ph_plan **plan, int64_t dy, int64_t dz, int64_t batch, int method,
              double pixel_size, double dist, double energy, double alpha, double db,
              double W);
/* Pad widths and padded sizes of the plan. */
int ph_get_pads(const tomo_ph_plan *plan, int64_t *py, int64_t *pz, int64_t *Ny, int64_t *Nx);
/*
 * _retrieve_phase (retrieve_phase.py:143-161) in place on data (nproj, dy, dz)
 * float32: edge-replicated padding, fft2, times the normalised filter, ifft2,
 * real part, crop.  Transforms and filter in float32.
 */
int ph_apply(tomo_ph_plan *plan, float *data, int64_t nproj, void *stream);
void ph_destroy(tomo_ph_plan *plan);

/*
 * ProcFunctions.pad360 (src/tomocupy/processing/proc_functions.py:109-124),
 * the double field of view of 360-degree scans (file_type double_fov,
 * :193-194), on sinograms in the plan's layout: sino (nz, nproj, n) ->
 * out (nz, nproj, 2n), both of dtype.  The detector axis is flipped when
 * flip != 0 (rotation axis in the left half, reader.py:143-148), the last
 * w = max(1, int(2 (n - center))) columns are tapered by
 * t^5 (126 - 420 t + 540 t^2 - 315 t^3 + 70 t^4), t = 1 - i/w (FP64), and the
 * right half is zero.  `center` is the axis after the flip rule (>= n/2).
 */
int pre_pad360(tomo_pre_plan *plan, const void *sino, void *out, int dtype, double center,
               int flip, void *stream);

/*
 * Fourier-wavelet stripe removal, remove_stripe_fw
 * (src/tomocupy/processing/remove_stripe.py:253-293), in place on
 *   res (nproj, nz, n) float32.
 * wname: haar, db5, sym5, sym16 (config.py:178-182); level >= 1; sigma > 0.
 */
int pre_remove_stripe_fw(tomo_pre_plan *plan, float *res, double sigma, const char *wname,
                         int level, void *stream);

/*
 * The whole chunk preprocessing with Fourier-wavelet stripe removal, fused:
 * proc_sino (dark/flat, remove_stripe_fw) + proc_proj (minus_log) + the swap
 * to sinogram order (proc_functions.py:140-198, rec.py:172-174).  The dark/flat
 * correction happens in the first wavelet pass and minus_log + swap in the
 * last, so the transmission chunk is never written.
 *   data, dark, flat   uint16 chunk as pre_process
 *   out                device (nz, nproj, n) of out_dtype
 *   sigma, wname, level  as pre_remove_stripe_fw
 */
int pre_process_fw(tomo_pre_plan *plan, const uint16_t *data, const uint16_t *dark,
                   const uint16_t *flat, void *out, int out_dtype, double bright_ratio,
                   int flat_linear, int minus_log, double sigma, const char *wname, int level,
                   void *stream);

void pre_destroy(tomo_pre_plan *plan);

#ifdef __cplusplus
}
#endif

#endif /* TOMOHIP_H */
