// Brute-force convolution on the matrix cores with split-BF16x3 products
// (AssociativeBruteForceConvolver, src/libprofit/src/profit/convolve.cpp:158-231,
// computed with a different summation order).
//
// A PSF row ky acts on an input row as a Toeplitz matrix, so a 32 x 32 block
// of outputs is
//   C[y, x] = sum_ky sum_xi In[y + ky, xi] * T_ky[xi, x],   T_ky[xi, x] = w[ky][xi - x]
// which is a chain of 32 x 32 x 16 MFMAs over K chunks of 16 input columns.
// Every FP32 operand is split into bf16 hi + lo and the product is taken as
// hi*hi + hi*lo + lo*hi (about 16 significant bits per operand), accumulated
// in FP32 by the matrix core and flushed to FP64 every few PSF rows.
//
// Block: 256 threads = 2 x 2 wavefronts over a 128 x 128 output tile, each
// wavefront 64 x 64 outputs = 2 x 2 MFMA blocks.  The input tile (bf16 hi and
// lo planes) sits in LDS for a chunk of PSF rows; the B operands (Toeplitz
// fragments of the weights, identical for every tile) are precomputed per
// galaxy in the exact lane layout and read through the vector cache.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "device.h"

namespace profit_amd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kFlush = 8;  // PSF rows accumulated in FP32 between FP64 flushes

__device__ inline int owner_m(const int *__restrict__ off, int n, int b)
{
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (off[mid] <= b) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

__device__ inline f32x16 mfma(u32x4 a, u32x4 b, f32x16 c)
{
	return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
	                                               0, 0, 0);
}

__device__ inline f32x16 mfma3(u32x4 ah, u32x4 al, u32x4 bh, u32x4 bl, f32x16 c)
{
	c = mfma(ah, bh, c);
	c = mfma(ah, bl, c);
	return mfma(al, bh, c);
}

__device__ inline unsigned short bf16_bits(float v)
{
	return __builtin_bit_cast(unsigned short, (__bf16)v);
}

}  // namespace

// tile_off: per-instance prefix sums of 128 x 128 tiles.  pitch: LDS row pitch
// (bf16 elements) of every galaxy of the launch, kch: PSF rows per LDS chunk.
template <int MODE>
__global__ __launch_bounds__(256) void k_conv_mfma(const GalaxyDev *__restrict__ gals,
                                                   const InstDev *__restrict__ insts, const int *__restrict__ tile_off,
                                                   int ninst, int pitch, int kch, const int *__restrict__ lk_base,
                                                   double *partial)
{
	extern __shared__ __attribute__((aligned(16))) unsigned short smem[];
	const int b = blockIdx.x;
	const int inst = owner_m(tile_off, ninst, b);
	const InstDev I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int2 t = (MODE == CONV_LIKE) ? g.mtiles_like[b - tile_off[inst]] : g.mtiles_canvas[b - tile_off[inst]];
	const int Kx = g.Kx, Ky = g.Ky, KC = g.mfma_kc;
	const int x0 = t.x * kMfmaTile, y0 = t.y * kMfmaTile;
	const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int wx = wv & 1, wy = wv >> 1;
	const int r32 = lane & 31, g32 = lane >> 5;

	const float *__restrict__ in;
	const u32x4 *__restrict__ frag;
	int IW, IH, ox, oy, OW, OH;
	if (MODE == CONV_ADJ) {
		in = I.w;
		IW = g.W;
		IH = g.H;
		OW = g.DW;
		OH = g.DH;
		ox = -g.padx + Kx / 2 - Kx + 1;
		oy = -g.pady + Ky / 2 - Ky + 1;
		frag = reinterpret_cast<const u32x4 *>(g.mfrag_adj);
	}
	else {
		in = I.canvas;
		IW = g.DW;
		IH = g.DH;
		if (MODE == CONV_LIKE) {
			OW = g.W;
			OH = g.H;
			ox = g.padx - Kx / 2;
			oy = g.pady - Ky / 2;
		}
		else {
			OW = g.DW;
			OH = g.DH;
			ox = -(Kx / 2);
			oy = -(Ky / 2);
		}
		frag = reinterpret_cast<const u32x4 *>(g.mfrag_fwd);
	}
	const int LWc = kMfmaTile + Kx - 1;  // input columns with nonzero weight
	unsigned short *hi = smem;
	unsigned short *lo = smem + (size_t)(kMfmaTile + kch - 1) * pitch;

	double accd[4][16];
#pragma unroll
	for (int q = 0; q < 4; ++q)
#pragma unroll
		for (int r = 0; r < 16; ++r) accd[q][r] = 0.0;
	f32x16 c00 = {}, c01 = {}, c10 = {}, c11 = {};

	const int row0 = wy * 64 + r32;  // local output row of MFMA block yb = 0 (yb = 1: + 32)
	const int colw = wx * 64 + 8 * g32;
	int pending = 0;
	for (int ky0 = 0; ky0 < Ky; ky0 += kch) {
		const int n = min(kch, Ky - ky0);
		const int rows = kMfmaTile + n - 1;
		__syncthreads();
		const int gy0 = y0 + oy + ky0, gx0 = x0 + ox;
		for (int idx = threadIdx.x; idx < rows * pitch; idx += 256) {
			const int rr = idx / pitch, cc = idx - rr * pitch;
			const int gy = gy0 + rr, gx = gx0 + cc;
			float v = 0.f;
			if (cc < LWc && gy >= 0 && gy < IH && gx >= 0 && gx < IW) v = in[(long)gy * IW + gx];
			const __bf16 h = (__bf16)v;
			hi[idx] = __builtin_bit_cast(unsigned short, h);
			lo[idx] = bf16_bits(v - (float)h);
		}
		__syncthreads();
		for (int kk = 0; kk < n; ++kk) {
			const u32x4 *__restrict__ fr = frag + (size_t)(ky0 + kk) * KC * 128 + lane;
			const unsigned short *ah = hi + (size_t)(row0 + kk) * pitch + colw;
			const unsigned short *al = lo + (size_t)(row0 + kk) * pitch + colw;
			const size_t rstep = (size_t)32 * pitch;
			u32x4 b0h = {}, b0l = {}, b1h = {}, b1l = {}, b2h = {}, b2l = {};
			u32x4 nh = fr[0], nl = fr[64];
			for (int kca = 0; kca < KC + 2; ++kca) {
				b2h = b1h;
				b2l = b1l;
				b1h = b0h;
				b1l = b0l;
				b0h = nh;
				b0l = nl;
				if (kca + 1 < KC) {
					nh = fr[(kca + 1) * 128];
					nl = fr[(kca + 1) * 128 + 64];
				}
				else {
					nh = u32x4{0u, 0u, 0u, 0u};
					nl = nh;
				}
				const int c = 16 * kca;
				const u32x4 a0h = *reinterpret_cast<const u32x4 *>(ah + c);
				const u32x4 a0l = *reinterpret_cast<const u32x4 *>(al + c);
				const u32x4 a1h = *reinterpret_cast<const u32x4 *>(ah + rstep + c);
				const u32x4 a1l = *reinterpret_cast<const u32x4 *>(al + rstep + c);
				if (kca < KC) {
					c00 = mfma3(a0h, a0l, b0h, b0l, c00);
					c10 = mfma3(a1h, a1l, b0h, b0l, c10);
				}
				if (kca >= 2) {
					c01 = mfma3(a0h, a0l, b2h, b2l, c01);
					c11 = mfma3(a1h, a1l, b2h, b2l, c11);
				}
			}
			if (++pending == kFlush || kk == n - 1) {
#pragma unroll
				for (int r = 0; r < 16; ++r) {
					accd[0][r] += (double)c00[r];
					accd[1][r] += (double)c01[r];
					accd[2][r] += (double)c10[r];
					accd[3][r] += (double)c11[r];
				}
				c00 = f32x16{};
				c01 = f32x16{};
				c10 = f32x16{};
				c11 = f32x16{};
				pending = 0;
			}
		}
	}

	// outputs: block q = 2*yb + xb, item r -> x = r32, y = (r&3) + 8*(r>>2) + 4*g32
	if (MODE == CONV_LIKE) {
		double part = 0;
		const bool pois = g.likefunc == LIKE_POIS, tl = g.likefunc == LIKE_T;
		const double s = g.pois_scale;
#pragma unroll
		for (int q = 0; q < 4; ++q) {
			const int xx = x0 + wx * 64 + (q & 1) * 32 + r32;
#pragma unroll
			for (int r = 0; r < 16; ++r) {
				const int yy = y0 + wy * 64 + (q >> 1) * 32 + (r & 3) + 8 * (r >> 2) + 4 * g32;
				if (xx < OW && yy < OH) {
					const long o = (long)yy * OW + xx;
					float wvv = 0.f;
					if (g.region[o]) {
						const long qc = (long)(yy + g.pady) * g.DW + xx + g.padx;
						const bool in_calc = !g.has_calc || g.calc[qc];
						const double m = in_calc ? accd[q][r] + (I.has_aux ? (double)I.aux[qc] : 0.0) : 0.0;
						const double d = g.data[o];
						if (pois) {
							const double mm = s * m, dd = s * d;
							part += mm - dd * log(mm);
							wvv = (float)(-s * (1.0 - d / m));
						}
						else {
							const double diff = d - m;
							const double iv = g.ivar[o];
							part += diff * diff * iv;
							wvv = (float)(diff * iv);
							// t: the standardised residual, for k_like_t
							if (tl) wvv = (float)(diff * sqrt(iv));
						}
						if (!in_calc && !tl) wvv = 0.f;
					}
					I.w[o] = wvv;
				}
			}
		}
		__shared__ double red[4];
#pragma unroll
		for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
		if (lane == 0) red[wv] = part;
		__syncthreads();
		if (threadIdx.x == 0) partial[lk_base[inst] + (b - tile_off[inst])] = (red[0] + red[1]) + (red[2] + red[3]);
	}
	else {
#pragma unroll
		for (int q = 0; q < 4; ++q) {
			const int xx = x0 + wx * 64 + (q & 1) * 32 + r32;
#pragma unroll
			for (int r = 0; r < 16; ++r) {
				const int yy = y0 + wy * 64 + (q >> 1) * 32 + (r & 3) + 8 * (r >> 2) + 4 * g32;
				if (xx < OW && yy < OH) {
					const long o = (long)yy * OW + xx;
					float v = (float)accd[q][r];
					if (MODE == CONV_MODEL && g.has_calc && !g.calc[o]) v = 0.f;
					if (MODE == CONV_MODEL && I.has_aux) v += I.aux[o];
					I.G[o] = v;
				}
			}
		}
	}
}

// ------------------------------------------------------------------ host side

int mfma_kc(int kx) { return (kx + 31 + 15) / 16; }

int mfma_pitch(int kc) { return 64 + 16 * (kc + 2) + 8; }

// PSF rows per LDS chunk for the given pitch within `budget` bytes (0: none fit)
int mfma_rows_per_chunk(int pitch, int ky, size_t budget)
{
	const long rows = (long)(budget / ((size_t)pitch * 4));
	const long k = rows - (kMfmaTile - 1);
	if (k < 1) return 0;
	return (int)(k < ky ? k : ky);
}

static uint16_t host_bf16(float v)
{
	uint32_t u;
	std::memcpy(&u, &v, 4);
	if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}

static float host_bf16_to_float(uint16_t h)
{
	const uint32_t u = (uint32_t)h << 16;
	float v;
	std::memcpy(&v, &u, 4);
	return v;
}

// B fragments of a Ky x Kx weight array (out(p) = sum in(p + o + k) wt[k]):
// for lane l, item i of K chunk kc the weight wt[ky][16 kc + 8 (l / 32) + i - l % 32]
// (0 outside [0, Kx)), split into bf16 hi and lo; layout [ky][kc][hi|lo][lane][8].
void mfma_fragments(const float *wt, int kx, int ky, std::vector<uint16_t> &out)
{
	const int kc_n = mfma_kc(kx);
	out.assign((size_t)ky * kc_n * 2 * 64 * 8, 0);
	for (int y = 0; y < ky; ++y)
		for (int kc = 0; kc < kc_n; ++kc)
			for (int l = 0; l < 64; ++l)
				for (int i = 0; i < 8; ++i) {
					const int d = 16 * kc + 8 * (l / 32) + i - l % 32;
					const float w = (d >= 0 && d < kx) ? wt[(size_t)y * kx + d] : 0.f;
					const uint16_t h = host_bf16(w);
					const uint16_t lw = host_bf16(w - host_bf16_to_float(h));
					const size_t base = (((size_t)y * kc_n + kc) * 2) * 64 * 8;
					out[base + (size_t)l * 8 + i] = h;
					out[base + 64 * 8 + (size_t)l * 8 + i] = lw;
				}
}

template <int MODE>
static int mfma_launch(int nblocks, size_t lds, const GalaxyDev *gals, const InstDev *insts, const int *tile_off,
                       int ninst, int pitch, int kch, const int *lk_base, double *partial, hipStream_t st)
{
	static size_t done[64];
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return (int)hipErrorInvalidDevice;
	if (lds > 64 * 1024 && done[dev] < lds) {
		hipError_t e = hipFuncSetAttribute((const void *)k_conv_mfma<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   (int)lds);
		if (e != hipSuccess) return (int)e;
		done[dev] = lds;
	}
	k_conv_mfma<MODE><<<nblocks, 256, lds, st>>>(gals, insts, tile_off, ninst, pitch, kch, lk_base, partial);
	return (int)hipGetLastError();
}

int launch_conv_mfma(int mode, int nblocks, int pitch, int kch, const GalaxyDev *gals, const InstDev *insts,
                     const int *tile_off, int ninst, const int *lk_base, double *partial, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	if (kch < 1 || pitch % 8) return (int)hipErrorInvalidValue;
	const size_t lds = (size_t)2 * (kMfmaTile + kch - 1) * pitch * sizeof(unsigned short);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
	if (mode == CONV_LIKE)
		return mfma_launch<CONV_LIKE>(nblocks, lds, gals, insts, tile_off, ninst, pitch, kch, lk_base, partial, st);
	if (mode == CONV_MODEL)
		return mfma_launch<CONV_MODEL>(nblocks, lds, gals, insts, tile_off, ninst, pitch, kch, lk_base, partial, st);
	return mfma_launch<CONV_ADJ>(nblocks, lds, gals, insts, tile_off, ninst, pitch, kch, lk_base, partial, st);
}

}  // namespace profit_amd
