// Device-side data layout shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "profile_math.h"

namespace profit_amd {

// likelihood kinds (R/profitLikeModel.R:263-329)
enum LikeKind : int32_t { LIKE_NORM = 0, LIKE_CHISQ = 1, LIKE_T = 2, LIKE_POIS = 3 };

constexpr int kTileW = 64;       // tile width: one wavefront spans a tile row
constexpr int kRenderRows = 4;   // render tiles are 64 x 4 (256 threads)
constexpr int kMaxLevels = 32;   // recursion walker depth (max_recursions <= 30)
constexpr int kMfmaTile = 128;   // output tile side of the split-BF16x3 convolution

// output of a convolution: fused likelihood, model image, adjoint correlation
enum ConvMode : int { CONV_LIKE = 0, CONV_MODEL = 1, CONV_ADJ = 2 };

// convolution kernel of a galaxy
enum ConvKind { CK_VALU = 0, CK_FFT = 1, CK_MFMA = 2 };

// Row pitch (in complex entries) of an FFT half-spectrum frame of width
// n1 / 2 + 1: rounded up to 8 entries (128 B) so that a column block's
// eight adjacent entries of a row fill one cache line.
__host__ __device__ constexpr int fft_pitch(int n1) { return ((n1 / 2 + 1) + 7) & ~7; }

// Row pairs per block of the FFT row passes (one pair per wavefront while
// the pairs and the twiddles fit the LDS).
__host__ __device__ constexpr int fft_pairs_per_block(int n1) { return n1 <= 1024 ? 4 : (n1 <= 2048 ? 2 : 1); }

// One image pixel's likelihood inputs, fetched with one 16-byte load: the
// data (FP64), 1/sigma^2, and flags (bit 0: in the likelihood region; bit 1:
// its canvas pixel in the calc region).
struct alignas(16) PixelIn {
	double data;
	float ivar;
	uint32_t flags;
};

// One galaxy: image geometry, PSF, masks and data.  Several instances (parameter
// sets) may share a galaxy.
struct GalaxyDev {
	int32_t W, H;          // image (likelihood) dimensions
	int32_t DW, DH;        // drawing canvas dimensions
	int32_t padx, pady;    // PSF padding (canvas shift)
	int32_t Kx, Ky;        // PSF dimensions (1 x 1 unit kernel without a PSF)
	int32_t has_psf;       // 0: no PSF (the unit kernel stands in)
	int32_t conv_kind;     // CK_VALU, CK_FFT or CK_MFMA
	int32_t likefunc;
	int32_t has_calc;
	int32_t n_rtiles, n_ctiles, n_atiles;
	int32_t conv_rows;     // rows per thread of the convolution tiles
	int32_t render_fp64;   // profile values (direct pixels, recursion leaves) in FP64 instead of FP32
	int32_t canvas64;      // the canvas is FP64 (InstDev::canvas64) and every PSF convolution the FFT
	int32_t grad_box;      // the adjoint computes d / d box (the engine's gradient-slot mask)
	double sx, sy;         // profile pixel scale
	double img_sx, img_sy; // image pixel scale
	double magzero;
	double n_region;       // number of likelihood pixels
	double n_calc;         // number of calc-region pixels of the canvas
	double pois_scale;     // Poisson rescale factor (1 when not rescaled)
	const float *psf_fwd;  // flipped PSF, weights of the forward convolution
	const float *psf_adj;  // PSF as given, weights of the adjoint correlation
	const uint8_t *calc;   // DW*DH drawing mask (null: every pixel)
	const uint8_t *region; // W*H likelihood mask (already intersected with calc)
	const double *data;    // W*H, FP64: the Poisson term d log m needs the data's own precision
	const float *ivar;     // W*H, 1/sigma^2
	const PixelIn *pix;    // W*H: data, ivar, region and calc packed
	int32_t fft_n1, fft_n2;   // FFT frame (powers of two) when the FFT path is on
	int32_t fft_l1, fft_l2;   // their base-2 logarithms
	const double2 *fft_spec;  // half-spectrum of the PSF (N2 rows at fft_pitch(N1)), wrapped about the origin
	int32_t mfma_kc;          // 16-wide K chunks of one Toeplitz row block (split-BF16x3 path)
	int32_t n_mtl, n_mtc;     // 128 x 128 tiles: likelihood (image) and canvas
	const uint4 *mfrag_fwd;   // B fragments of the forward weights, [ky][kc][hi|lo][lane]
	const uint4 *mfrag_adj;   // and of the adjoint weights
	const int2 *mtiles_like;  // 128 x 128 tiles on the image holding region pixels
	const int2 *mtiles_canvas;// 128 x 128 tiles on the canvas holding calc pixels
	// the model's normalised PSF for point sources (null without a PSF), its
	// size and pixel scale (Model::set_psf_pixel_scale, default 1 x 1)
	const float *psf_pt;
	int32_t pkx, pky;
	double psf_sx, psf_sy;
	const int2 *rtiles;    // render / adjoint-render tiles (64 x 4) on the canvas
	const int2 *ctiles;    // forward convolution tiles on the image
	const int2 *atiles;    // adjoint convolution tiles on the canvas
};

struct InstDev {
	int32_t galaxy, comp0, ncomp;
	int32_t has_aux;       // the likelihood path draws unconvolved point sources / sky into aux
	float *canvas;         // DW*DH rendered (to be convolved) image
	float *w;              // W*H  d logL / d model (up to a per-instance scalar)
	float *G;              // DW*DH  correlated gradient image / convolved model
	float *aux;            // DW*DH  profiles drawn without convolution (or null)
	double2 *fbuf;         // N2 x N1 FFT frame (FFT path only)
	// per component, canvas row and 64-pixel tile column: bit set where the
	// forward render evaluated the pixel directly (ncomp x DH x ceil(DW/64))
	uint64_t *dmask;
	double *canvas64;      // DW*DH FP64 canvas (GalaxyDev::canvas64), else null
};

struct QEntry {
	int32_t comp;
	int32_t pix;
};

constexpr int kFamilies = 6;  // radial profile families (SERSIC .. BROKENEXP)

// Sub-pixel work lists: the ordinary entries, one list per profile family
// (so that each family's walker reads only its own entries), and the deep
// centre-pixel entries (resolution 8, depth 10).  Family f's list is
// light[lbase[f], lbase[f+1]).
struct WorkQueue {
	QEntry *light, *heavy;
	int *count;
	int hcap;
	int lbase[kFamilies + 1];
	// per entry, kJacStride floats: the pixel's derivatives sum_s w_s df_s /
	// d slot (slot order of GradSlot) and, last, its value sum_s w_s f_s,
	// written by the forward walk of a gradient evaluation (null otherwise)
	float *ljac, *hjac;
	// the light entries whose root node has a cell to refine (indices into
	// the family's light list, in the same segments): the root pass, one lane
	// per entry, hands these to the wavefront walker
	int *deep;
};
constexpr int kJacStride = 12;
// count[1]: length of the heavy list; then a (heavy, deep) pair of hand-out
// counters per walker launch (one per profile family); the length of each
// family's light list; the root pass's hand-out counter per family; the
// length of each family's deep list (count[0] is not used)
constexpr int kQueueCounters = 32;
__host__ __device__ constexpr int fwd_counter(int type) { return 2 + 2 * type; }
__host__ __device__ constexpr int light_count(int type) { return 14 + type; }
__host__ __device__ constexpr int root_counter(int type) { return 20 + type; }
__host__ __device__ constexpr int deep_count(int type) { return 26 + type; }


}  // namespace profit_amd
