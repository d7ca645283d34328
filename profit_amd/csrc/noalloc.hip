// Allocation-free brute-force convolution: Convolver::convolve of the
// brute-force convolver (convolve.cpp:101-231) on device buffers the caller
// owns, enqueued on the caller's stream.  It has no counterpart in the
// reference; it serves callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops).
//
// The staging that the allocating entry point (convolve_image in engine.cpp)
// allocates per call -- the FP32 copy of the source, the flipped FP32 kernel,
// the tile list, the galaxy and instance descriptors and the FP32 result --
// lives here in one file-scope device array, `scratch`, of fixed size.  Every
// GPU the library runs on holds its own copy until the process ends.  Host
// code never uses the array's name as an address: its device address is taken
// once per device with hipGetSymbolAddress and kept in g_scratch_ptr.
//
// The convolution itself is k_conv<4, CONV_MODEL> with the launch parameters
// of convolve_image (rows 4, every 64 x 16 tile in row-major order, PSF rows in
// chunks of 64), so both entry points give bit-identical results.
//
// Nothing in the array needs clearing between calls: each call writes every
// part it reads (the descriptors, tiles, source and kernel in k_noalloc_prep,
// then every output pixel, masked ones as 0, in k_conv) before reading it.
// The array keeps the last call's contents, and calls that overlap on one
// device race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

#include "device.h"
#include "engine.h"
#include "noalloc.h"

namespace profit_amd {

namespace {

constexpr int kRows = 4;                    // k_conv rows per thread, as convolve_image
constexpr int kTileH = 4 * kRows;           // 64 x 16 output tiles
constexpr size_t kAlign = 256;

constexpr size_t round_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// scratch layout (byte offsets)
constexpr size_t kOffGal = 0;
constexpr size_t kOffInst = kOffGal + round_up(sizeof(GalaxyDev));
constexpr size_t kOffTileOff = kOffInst + round_up(sizeof(InstDev));
constexpr size_t kOffTiles = kOffTileOff + round_up(2 * sizeof(int));
constexpr size_t kOffKernel = kOffTiles + round_up(kNoAllocMaxTiles * sizeof(int2));
constexpr size_t kOffSrc = kOffKernel + round_up((size_t)kNoAllocMaxKernel * kNoAllocMaxKernel * sizeof(float));
constexpr size_t kOffDst = kOffSrc + round_up(kNoAllocMaxPixels * sizeof(float));
constexpr size_t kScratchBytes = kOffDst + round_up(kNoAllocMaxPixels * sizeof(float));

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// the device address of this device's scratch array (the current device)
unsigned char *scratch_on(int device)
{
	if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device ordinal out of range");
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_ptr[device]) {
		void *p = nullptr;
		hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
		g_scratch_ptr[device] = static_cast<unsigned char *>(p);
	}
	return g_scratch_ptr[device];
}

// Fills the scratch array for one call: descriptors (from the arguments), the
// tile list, the source in FP32 and the kernel flipped in FP32 (the forward
// weights k_conv reads), one grid-stride pass.
__global__ void k_noalloc_prep(GalaxyDev g, InstDev inst, const double *__restrict__ src,
                               const double *__restrict__ krn, int ntx, unsigned char *base)
{
	const long n = (long)g.W * g.H, nk = (long)g.Kx * g.Ky, nt = g.n_atiles;
	const long stride = (long)gridDim.x * blockDim.x;
	const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t0 == 0) {
		*reinterpret_cast<GalaxyDev *>(base + kOffGal) = g;
		*reinterpret_cast<InstDev *>(base + kOffInst) = inst;
		int *toff = reinterpret_cast<int *>(base + kOffTileOff);
		toff[0] = 0;
		toff[1] = g.n_atiles;
	}
	float *canvas = reinterpret_cast<float *>(base + kOffSrc);
	float *fwd = reinterpret_cast<float *>(base + kOffKernel);
	int2 *tiles = reinterpret_cast<int2 *>(base + kOffTiles);
	for (long k = t0; k < n; k += stride) canvas[k] = (float)src[k];
	for (long k = t0; k < nk; k += stride) {
		const long ky = k / g.Kx, kx = k - ky * g.Kx;
		fwd[k] = (float)krn[(g.Ky - 1 - ky) * (long)g.Kx + (g.Kx - 1 - kx)];
	}
	for (long k = t0; k < nt; k += stride) tiles[k] = make_int2((int)(k % ntx), (int)(k / ntx));
}

__global__ void k_noalloc_out(const float *__restrict__ res, double *__restrict__ out, long n)
{
	const long stride = (long)gridDim.x * blockDim.x;
	for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) out[k] = (double)res[k];
}

__global__ void k_noalloc_copy(const double *__restrict__ src, double *__restrict__ out, long n)
{
	const long stride = (long)gridDim.x * blockDim.x;
	for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) out[k] = src[k];
}

int grid_for(long n) { return (int)std::min<long>(std::max<long>(1, (n + 255) / 256), 2048); }

}  // namespace

void convolve_noalloc(int method, const double *src, int w, int h, const double *krn, int kw, int kh,
                      const uint8_t *mask, double *out, int device, hipStream_t st)
{
	if (w <= 0 || h <= 0) throw std::invalid_argument("source dimensions must be positive");
	if (!src || !out) throw std::invalid_argument("null source or output pointer");
	const long n = (long)w * h;
	if (method == NOALLOC_NULL) {
		// NullConvolver (convolve.cpp:521-524): the source unchanged, a plain
		// device copy that uses no scratch (so any size)
		hip_check(hipSetDevice(device), "hipSetDevice");
		k_noalloc_copy<<<grid_for(n), 256, 0, st>>>(src, out, n);
		hip_check(hipGetLastError(), "copy launch");
		return;
	}
	if (w > kNoAllocMaxSide || h > kNoAllocMaxSide || (size_t)w * h > kNoAllocMaxPixels)
		throw std::invalid_argument("source too large for the allocation-free convolution: at most " +
		                            std::to_string(kNoAllocMaxSide) + " pixels per side and " +
		                            std::to_string(kNoAllocMaxPixels) + " pixels");
	hip_check(hipSetDevice(device), "hipSetDevice");
	if (method != NOALLOC_BRUTE) throw std::invalid_argument("the allocation-free convolution is brute force only");
	if (!krn) throw std::invalid_argument("null kernel pointer");
	if (kw <= 0 || kh <= 0 || kw > kNoAllocMaxKernel || kh > kNoAllocMaxKernel)
		throw std::invalid_argument("kernel dimensions must be 1 to " + std::to_string(kNoAllocMaxKernel));
	// Convolver::convolve extends a source smaller than the kernel (convolve.cpp:58-74),
	// which needs a buffer of the extended size
	if (w < kw || h < kh)
		throw std::invalid_argument("the allocation-free convolution needs a source at least as large as the kernel");
	const int ntx = (w + kTileW - 1) / kTileW, nty = (h + kTileH - 1) / kTileH;
	if ((long)ntx * nty > (long)kNoAllocMaxTiles) throw std::invalid_argument("too many tiles");
	unsigned char *base = scratch_on(device);

	GalaxyDev g{};
	g.W = g.DW = w;
	g.H = g.DH = h;
	g.Kx = kw;
	g.Ky = kh;
	g.has_calc = mask != nullptr;
	g.calc = mask;
	g.psf_fwd = reinterpret_cast<const float *>(base + kOffKernel);
	g.atiles = reinterpret_cast<const int2 *>(base + kOffTiles);
	g.n_atiles = ntx * nty;
	g.conv_rows = kRows;
	InstDev I{};
	I.canvas = reinterpret_cast<float *>(base + kOffSrc);
	I.G = reinterpret_cast<float *>(base + kOffDst);
	const long nprep = std::max<long>(std::max<long>(n, (long)kw * kh), g.n_atiles);
	k_noalloc_prep<<<grid_for(nprep), 256, 0, st>>>(g, I, src, krn, ntx, base);
	hip_check(hipGetLastError(), "prep launch");
	hip_check(launch_conv(CONV_MODEL, kRows, g.n_atiles, kw, std::min(kh, 64),
	                      reinterpret_cast<const GalaxyDev *>(base + kOffGal),
	                      reinterpret_cast<const InstDev *>(base + kOffInst),
	                      reinterpret_cast<const int *>(base + kOffTileOff), 1, nullptr, nullptr, st),
	          "convolution launch");
	k_noalloc_out<<<grid_for(n), 256, 0, st>>>(reinterpret_cast<const float *>(base + kOffDst), out, n);
	hip_check(hipGetLastError(), "output launch");
}

}  // namespace profit_amd
