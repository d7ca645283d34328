/*
 * profit_amd: C-ABI of the MI355X model-image path (rendering with adaptive
 * sub-pixel recursion, PSF convolution, fused likelihood and its adjoint).
 *
 * Plain pointers and sizes, no C++ types.  Every entry point returns 0 on
 * success and a non-zero status on failure; the message is available from
 * profit_amd_last_error() (thread-local).  No exception crosses the ABI.
 * Images are row-major with x fastest (index i + j*width), exactly as
 * libprofit's Image and an R matrix with nrow = width.
 *
 * Reference interfaces replaced (paths relative to ICRAR/ProFit):
 *   src/r_profit.cpp:549-681  R_profit_make_model      -> profit_amd_make_model
 *   src/r_profit.cpp:684-706  R_profit_convolve        -> profit_amd_convolve
 *   src/r_profit.cpp:504-547  R_profit_make_convolver  -> convolver argument of
 *                                                          profit_amd_convolve
 *   src/r_profit.cpp:757-772  R_profit_adjust_mask     -> profit_amd_adjust_mask
 *   src/r_profit.cpp:483-502  R_profit_convolvers      -> profit_amd_convolvers
 *   src/r_profit.cpp:708-755  R_profit_downsample/upsample -> profit_amd_downsample/upsample
 *   src/libprofit/src/profit/model.cpp:271-338  Model::evaluate
 *   R/profitLikeModel.R:244-332  the likelihood (R code upstream) ->
 *                                profit_amd_engine_like / _like_grad (new:
 *                                libprofit has no C/C++ likelihood or gradient)
 */
#ifndef PROFIT_AMD_H
#define PROFIT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* radial profile families (src/libprofit/src/profit/model.cpp:103-135) */
enum {
	PROFIT_AMD_SERSIC = 0,
	PROFIT_AMD_MOFFAT = 1,
	PROFIT_AMD_FERRER = 2,
	PROFIT_AMD_KING = 3,
	PROFIT_AMD_CORESERSIC = 4,
	PROFIT_AMD_BROKENEXP = 5,
	PROFIT_AMD_SKY = 100,  /* sky.cpp: bg in shape[0] */
	PROFIT_AMD_NULL = 101,
	PROFIT_AMD_PSF = 102   /* psf.cpp: a PSF point source, xcen, ycen, mag */
};

/* convolvers (src/libprofit/src/profit/convolve.h:68-80) */
enum { PROFIT_AMD_CONV_BRUTE_OLD = 0, PROFIT_AMD_CONV_BRUTE = 1, PROFIT_AMD_CONV_FFT = 3, PROFIT_AMD_CONV_NULL = 4 };
/* profit_amd's own: the brute-force sum on the matrix cores with split-BF16x3
 * products (bf16 hi + lo operands, hi*hi + hi*lo + lo*hi, FP32 accumulation
 * flushed to FP64 every 8 PSF rows); it has no libprofit counterpart. */
enum { PROFIT_AMD_CONV_BRUTE_BF16X3 = 16 };

/* likelihood functions (R/profitLikeModel.R:263-329) */
enum { PROFIT_AMD_LIKE_NORM = 0, PROFIT_AMD_LIKE_CHISQ = 1, PROFIT_AMD_LIKE_T = 2, PROFIT_AMD_LIKE_POIS = 3 };

/* gradient slots per component: d logL / d {xcen, ycen, mag, ang, axrat, box,
 * shape[0..4]} with shape as in profit_amd_profile */
#define PROFIT_AMD_NGRAD 11

/*
 * One profile with libprofit's parameter names (radial.cpp:724-752 and each
 * profile's constructor).  shape[]:
 *   sersic re, nser | moffat fwhm, con | ferrer rout, a, b | king rc, rt, a |
 *   coresersic re, rb, nser, a, b | brokenexp h1, h2, rb, a | sky bg | psf
 *   (point source): xcen, ycen, mag only.
 * profit_amd_profile_defaults() fills libprofit's constructor defaults.
 */
typedef struct profit_amd_profile {
	int32_t type;
	int32_t convolve, rough, adjust, rescale_flux;
	int32_t reserved;
	double xcen, ycen, mag, ang, axrat, box;
	double acc, rscale_switch, rscale_max;
	uint32_t resolution, max_recursions;
	double shape[5];
} profit_amd_profile;

/* Model settings (src/libprofit/src/profit/model.h:142-343) */
typedef struct profit_amd_model_desc {
	uint32_t width, height;
	double scale_x, scale_y;
	double magzero;
	uint32_t finesampling;
	int32_t return_finesampled, crop, adjust_mask, dry_run;
	const double *psf; /* psf_h rows of psf_w; normalised as Model::set_psf does; may be NULL */
	uint32_t psf_w, psf_h;
	const uint8_t *mask; /* calc region (nonzero = calculate); may be NULL */
	uint32_t mask_w, mask_h;
	int32_t convolver;   /* PROFIT_AMD_CONV_* */
	int32_t device;      /* HIP device ordinal */
	double psf_scale_x, psf_scale_y; /* Model::set_psf_pixel_scale; 0 = libprofit's default 1 */
} profit_amd_model_desc;

const char *profit_amd_last_error(void);
const char *profit_amd_version(void);

/* libprofit constructor defaults for a profile type. */
int profit_amd_profile_defaults(int32_t type, profit_amd_profile *p);

/* Model::evaluate: writes out_w*out_h doubles to out (capacity out_cap) and
 * the offset of the returned image.  Replaces R_profit_make_model
 * (src/r_profit.cpp:549-681) and Model::evaluate (model.cpp:271-338). */
int profit_amd_make_model(const profit_amd_model_desc *model, const profit_amd_profile *profiles, int32_t nprofiles,
                          double *out, size_t out_cap, uint32_t *out_w, uint32_t *out_h, int32_t *offset_x,
                          int32_t *offset_y);

/* Model::get_drawing_dimensions (model.cpp:221-225). */
int profit_amd_drawing_dims(const profit_amd_model_desc *model, const profit_amd_profile *profiles,
                            int32_t nprofiles, uint32_t *width, uint32_t *height);

/* The dimensions of the image profit_amd_make_model returns for these inputs:
 * the requested ones, or larger for an uncropped FFT frame (crop = 0), or the
 * drawing dimensions for a dry run.  No counterpart in the reference, whose
 * Model::evaluate (model.cpp:271-338) allocates its result; it lets a caller
 * size `out` (R_profit_make_model reuses previous_image when the dimensions
 * match, src/r_profit.cpp:662-670). */
int profit_amd_output_dims(const profit_amd_model_desc *model, const profit_amd_profile *profiles, int32_t nprofiles,
                           uint32_t *width, uint32_t *height);

/* Releases the device set-ups (buffers, tile lists, FFT spectra) that
 * repeated profit_amd_make_model calls with the same image geometry, PSF and
 * calc region reuse.  Replaces R_profit_clear_cache (src/r_profit.cpp:822-825),
 * which clears libprofit's OpenCL kernel cache. */
int profit_amd_clear_cache(void);

/* Model::adjust(mask, dims, psf, finesampling) (model.cpp:248-257); replaces
 * R_profit_adjust_mask (src/r_profit.cpp:757-772). */
int profit_amd_adjust_mask(const uint8_t *mask, uint32_t mask_w, uint32_t mask_h, uint32_t img_w, uint32_t img_h,
                           const double *psf, uint32_t psf_w, uint32_t psf_h, uint32_t finesampling, uint8_t *out,
                           size_t out_cap, uint32_t *out_w, uint32_t *out_h);

/* Convolver::convolve(src, krn, mask) for the given convolver; replaces
 * R_profit_convolve (src/r_profit.cpp:684-706). out has src_w*src_h doubles. */
int profit_amd_convolve(int32_t convolver, const double *src, uint32_t src_w, uint32_t src_h, const double *krn,
                        uint32_t krn_w, uint32_t krn_h, const uint8_t *mask, double *out, int32_t device);

/* Allocation-free variant of profit_amd_convolve.  No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  src, krn, mask (nullable) and out
 * are DEVICE pointers owned by the caller; the work is enqueued on `stream`
 * (a hipStream_t; NULL = the default stream) and the call returns after
 * enqueueing, calling no hipMalloc, hipFree or other allocating routine.  The
 * staging lives in a fixed device array of about 8.7 MB per GPU, held until
 * the process ends, which limits: src_w, src_h <= 4096 and src_w * src_h <=
 * 2^20; krn_w, krn_h <= 255 and no larger than the source (the allocating
 * entry point extends a smaller source).  Larger inputs are rejected with an
 * error.  Convolvers: PROFIT_AMD_CONV_BRUTE and _BRUTE_OLD (the same kernel as
 * profit_amd_convolve, bit-identical results) and _NULL (a plain copy of the
 * source, which uses no scratch and so takes any size).  Every call on one
 * GPU shares that array: calls that overlap on one GPU race. */
int profit_amd_convolve_noalloc(int32_t convolver, const double *src, uint32_t src_w, uint32_t src_h, const double *krn,
                                uint32_t krn_w, uint32_t krn_h, const uint8_t *mask, double *out, int32_t device,
                                void *stream);

/* Names of the available convolvers, comma separated (R_profit_convolvers). */
const char *profit_amd_convolvers(void);

/* Image::downsample(SUM|SAMPLE|AVERAGE) and Image::upsample (image.cpp:246-367);
 * replace R_profit_downsample / R_profit_upsample.  mode: 0 sample, 1 sum,
 * 2 average for downsampling; 0 copy, 1 scale for upsampling. */
int profit_amd_downsample(const double *img, uint32_t w, uint32_t h, uint32_t factor, int32_t mode, double *out,
                          size_t out_cap, uint32_t *out_w, uint32_t *out_h);
int profit_amd_upsample(const double *img, uint32_t w, uint32_t h, uint32_t factor, int32_t mode, double *out,
                        size_t out_cap, uint32_t *out_w, uint32_t *out_h);

/* ---------------------------------------------------------------- likelihood engine
 *
 * A galaxy is what profitSetupData prepares (R/profitSetupData.R:151-344):
 * image size, PSF, pre-adjusted calc region, likelihood region, data, sigma
 * and the likelihood function.  A batch is a list of instances, each a
 * galaxy plus a component count; one call evaluates logL (and its gradient)
 * of every instance in one launch sequence (a HIP graph after the first call).
 * Every radial component of a galaxy with a PSF is convolved (as ProFit's R
 * code sets); sky and point-source (PROFIT_AMD_PSF) components follow their
 * convolve flag, as libprofit's Model::produce_image does.
 */
typedef struct profit_amd_engine profit_amd_engine;

typedef struct profit_amd_galaxy_desc {
	uint32_t width, height;
	double scale_x, scale_y;
	double magzero;
	const double *psf;        /* may be NULL (no convolution) */
	uint32_t psf_w, psf_h;
	const uint8_t *calcregion; /* pre-adjusted calc region (drawing size) or NULL */
	uint32_t calc_w, calc_h;
	const uint8_t *region;    /* likelihood region (width*height) or NULL = every pixel */
	const double *data;       /* width*height */
	const double *sigma;      /* width*height */
	int32_t likefunc;         /* PROFIT_AMD_LIKE_* */
	double psf_scale_x, psf_scale_y; /* PSF pixel scale for point sources; 0 = 1 */
} profit_amd_galaxy_desc;

int profit_amd_engine_create(int32_t device, profit_amd_engine **engine);
void profit_amd_engine_destroy(profit_amd_engine *engine);
int profit_amd_engine_add_galaxy(profit_amd_engine *engine, const profit_amd_galaxy_desc *galaxy, int32_t *galaxy_id);
int profit_amd_engine_set_batch(profit_amd_engine *engine, const int32_t *inst_galaxy, const int32_t *inst_ncomp,
                                int32_t ninst);
/* Tile height of the convolution / rendering tiles: 4*rows (rows = 2, 4, 8, 16). */
int profit_amd_engine_set_tile_rows(profit_amd_engine *engine, int32_t rows);

/* Precision of the forward model: 0 = FP32 profile values (default), 1 =
 * FP64 profile values, for the directly evaluated pixels and the leaf samples
 * of the sub-pixel recursion alike; 2 = FP64 profile values summed into an
 * FP64 canvas that the FP64 FFT convolves (every galaxy with a PSF), so the
 * model image is FP64 end to end: the mode for bright sources whose logL the
 * FP32 canvas moves by more than 1e-3.  In every mode: classification is FP64,
 * refinement decisions are FP32 with an FP64 re-test at libprofit's own cell
 * position inside a guard band (so they equal the FP64 decisions), the brute
 * convolution (modes 0, 1) is FP32 with FP64 accumulation across PSF columns,
 * likelihood sums are FP64 and the adjoint's
 * per-sample derivatives are FP32, summed in FP64 in the recursion walk and,
 * for steep cores, in the direct pixels. */
int profit_amd_engine_set_render_fp64(profit_amd_engine *engine, int32_t on);

/* Gradient slots profit_amd_engine_like_grad computes: bit s of `mask` asks
 * for slot s (0 xcen .. 10 shape[4]; default all).  Slots left out come back
 * as 0; leaving out the box slot (bit 5) spares the per-sample logarithms of
 * d / d box.  No counterpart in the reference, whose R layer has no gradient
 * (R/profitLikeModel.R:244-342); it serves callers that fit a subset. */
int profit_amd_engine_set_grad_slots(profit_amd_engine *engine, uint32_t mask);

/* Convolution method of the engine's galaxies with a PSF, the choice that
 * Model::set_convolver (src/libprofit/src/profit/model.h:173-176) makes with a
 * BruteForceConvolver or an FFTConvolver (convolve.h:120-209):
 * PROFIT_AMD_CONV_BRUTE (default; tiled FP32 with FP64 accumulation),
 * PROFIT_AMD_CONV_FFT (FP64 2-D FFTs over a power-of-two frame) or
 * PROFIT_AMD_CONV_BRUTE_BF16X3 (matrix cores, split-BF16x3).  Rebuilds the
 * current batch. */
int profit_amd_engine_set_conv_method(profit_amd_engine *engine, int32_t convolver);

/* Autotuner (no libprofit counterpart; libprofit's convolver choice is the
 * caller's): times the FP32 brute force at 4, 8 and 16 rows per thread, the
 * FFT and (with PROFIT_AMD_TUNE_ALLOW_BF16X3 in flags) the split-BF16x3
 * kernels on the current batch with `params`, then gives each galaxy the
 * kernel of least predicted cost and keeps that mix only if it beats the best
 * single kernel.  kinds (nullable, one per galaxy): the chosen
 * PROFIT_AMD_CONV_* value.  ms (nullable, 6 doubles): forward + adjoint
 * convolution time per evaluation for brute x 3, fft, bf16x3 (0 when not
 * tried), chosen. */
enum { PROFIT_AMD_TUNE_ALLOW_BF16X3 = 1 };
int profit_amd_engine_autotune(profit_amd_engine *engine, const profit_amd_profile *params, int32_t reps,
                               int32_t flags, int32_t *kinds, double *ms);

/* logL per instance.  params: one profit_amd_profile per component,
 * instance-major, in the batch's component order. */
int profit_amd_engine_like(profit_amd_engine *engine, const profit_amd_profile *params, double *loglike);

/* logL per instance and d logL / d slot (PROFIT_AMD_NGRAD per component): the
 * gradient callback for R's optimisers (optim(gr=)) and HMC samplers.  The
 * sub-pixel recursion tree and the adjusted sub-pixel settings of the given
 * parameters are held fixed (they are piecewise constant in the parameters). */
int profit_amd_engine_like_grad(profit_amd_engine *engine, const profit_amd_profile *params, double *loglike,
                                double *grad);

/* The model image of every instance (width*height floats each, instance-major):
 * the convolved model the likelihood sees. */
int profit_amd_engine_model_images(profit_amd_engine *engine, const profit_amd_profile *params, float *out);

/* Rendering diagnostics for a batch of one instance: per component the pixel
 * class on the drawing canvas (0 skipped by the calc region, 1 beyond
 * rscale_max, 2 direct, 3 sub-pixel integration) and the number of nodes of
 * each pixel's recursion tree. */
int profit_amd_engine_diagnostics(profit_amd_engine *engine, const profit_amd_profile *params, int32_t *cls,
                                  int32_t *nodes);

/* The sub-pixel settings each component of the batch is drawn with, after
 * the adjustment RadialProfile::initial_calculations makes when adjust is on
 * (radial.cpp:158-185, which writes them back into the profile's acc,
 * rscale_switch, resolution and rscale_max): one value per component in
 * each array (nullable); 0 for point sources, sky and null profiles. */
int profit_amd_engine_adjusted_settings(profit_amd_engine *engine, const profit_amd_profile *params, int32_t *resolution,
                                        int32_t *max_recursions, double *acc, double *rscale_switch,
                                        double *rscale_max);

/* Drawing-canvas dimensions and PSF padding of a registered galaxy. */
int profit_amd_engine_galaxy_geometry(profit_amd_engine *engine, int32_t galaxy_id, uint32_t *draw_w,
                                      uint32_t *draw_h, uint32_t *pad_x, uint32_t *pad_y);

/* Per-stage device timing with HIP events on the engine's stream (evaluations
 * then run as an eager launch sequence rather than a graph).  Stages: 0 set-up,
 * 1 render, 2 sub-pixel walk, 3 convolution + likelihood, 4 likelihood
 * finalisation, 5 adjoint correlation, 6 adjoint render, 7 adjoint sub-pixel
 * walk, 8 gradient finalisation. */
int profit_amd_engine_set_timing(profit_amd_engine *engine, int32_t on);

/* Accumulated milliseconds and launch counts per stage (9 each), the number of
 * evaluations, the last sub-pixel queue length, and the algorithmic FLOPs of
 * the forward convolution and of the adjoint correlation of one evaluation of
 * the batch (2 * psf_w * psf_h * output pixels summed over instances). */
int profit_amd_engine_stats(profit_amd_engine *engine, double *stage_ms, int64_t *stage_calls, int64_t *evaluations,
                            int32_t *queue_length, double *conv_flops, double *adjoint_flops);
/* The FFT stages kernel by kernel (HIP events between the kernels while
 * timing is on): accumulated milliseconds of the forward rows, columns and
 * inverse rows of the likelihood stage (kernel_ms[0..2]) and of the adjoint
 * stage (kernel_ms[3..5]), and the number of evaluations each was timed over
 * (calls[0..1]).  No counterpart in the reference (measurement). */
int profit_amd_engine_fft_stats(profit_amd_engine *engine, double *kernel_ms, int64_t *calls);
int profit_amd_engine_reset_stats(profit_amd_engine *engine);

#ifdef __cplusplus
}
#endif

#endif /* PROFIT_AMD_H */
