// FFT convolution path (FFTConvolver, src/libprofit/src/profit/convolve.cpp:233-312).
//
// libprofit computes the 2-D linear convolution through a 1-D real FFT of the
// flattened, zero-extended image (FFTW).  Here the same linear convolution is
// computed with 2-D power-of-two FFTs of a real frame in FP64 on the GPU, in
// three passes over an N1 x N2 frame per instance that keeps only the
// half-spectrum (N1/2 + 1 columns, the rest follows from Hermitian symmetry):
//   k_fft_rows_fwd  two real rows packed as one complex FFT, split into
//                   their half-spectra
//   k_fft_cols      column FFT of the half-spectrum, product with the PSF
//                   spectrum (its conjugate for the adjoint correlation),
//                   inverse column FFT
//   k_fft_rows_inv  two half-spectrum rows packed as one complex inverse FFT,
//                   with the epilogue of the brute-force path: fused
//                   likelihood, masked model image, or the adjoint image G
// Every transform runs in LDS: decimation in time from bit-reversed input for
// the forward transforms, decimation in frequency to bit-reversed output for
// the inverse ones, three radix-2 stages per pass in registers.  Element e
// of a sequence sits at swz(e) in LDS, an XOR of its low bits with the next
// ones, so the strided accesses of every pass fall on distinct banks.  A
// pass's twiddles come from one entry per thread of the transform's table
// in LDS (Tw).  The row passes run one pair of
// rows per wavefront, several pairs per block.  FP64 keeps the error relative
// to the image maximum near 1e-15, so faint pixels keep their relative
// accuracy.
#include <hip/hip_runtime.h>

#include "device.h"

namespace profit_amd {

namespace {

__device__ inline int owner(const int *__restrict__ off, int n, int b)
{
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (off[mid] <= b) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

__device__ inline unsigned bitrev(unsigned i, int logn) { return __brev(i) >> (32 - logn); }

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// complex products with explicit FMAs (the library is built without
// contraction, which the classification's libprofit rounding needs): two
// multiplications and two FMAs instead of four and two additions
__device__ inline double2 cmul(double2 a, double2 b)
{
	return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ inline double2 conjg(double2 a) { return make_double2(a.x, -a.y); }

__device__ inline double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
	return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -(a.x * b.y)));
}

// LDS position of element e of a sequence: bits 0-2 ^= bits 3-5 and bit 3 ^=
// bit 6, a bijection on every power-of-two range.  A pass reads elements at
// spacing 1, 8 or 64 (for 2^9 points); after the swizzle the 16 lanes that
// share a bank cycle address distinct 16-byte slots in each case.
// Global loads a thread issues before it waits on the first one: each load
// loop gathers kBatch elements into registers, then stores them to LDS.
constexpr int kBatch = 8;

// Threads of a column-pass block: its LDS tile limits a CU to two blocks, so
// the block is wide enough to keep four waves per SIMD loading (and its
// registers are held to the 128 per lane that four waves allow).
constexpr int kColThreads = 512;

__device__ __forceinline__ int swz(int e) { return e ^ ((e >> 3) & 7) ^ (((e >> 6) & 1) << 3); }

// The twiddle table of an n = 2^lt point transform, tab[t] = exp(-2 pi i t /
// n) for t < n/2, staged in LDS from the global table of the largest
// transform: a radix-8 pass reads one entry per thread and derives the
// other twiddles from it (csq, rot8).
struct Tw {
	const double2 *tab;
	int lt;
	// exp(-2 pi i k / 2^S), k < 2^(S-1)
	__device__ __forceinline__ double2 at(int S, int k) const { return tab[k << (lt - S)]; }
};

// Stages the n/2 entries of an n = 2^logn point table (the caller's next
// barrier orders them before their first use).
__device__ __forceinline__ Tw stage_table(double2 *wl, int logn, const double2 *__restrict__ tw, int twn)
{
	const int sh = __ffs(twn) - 1 - logn;
	for (int t = threadIdx.x; t < (1 << (logn - 1)); t += blockDim.x) wl[t] = tw[t << sh];
	return Tw{wl, logn};
}

__device__ __forceinline__ double2 csq(double2 a) { return make_double2(fma(a.x, a.x, -(a.y * a.y)), 2.0 * a.x * a.y); }

// a * exp(-2 pi i m / 8), m = 0..3 (a compile-time constant once unrolled)
__device__ __forceinline__ double2 rot8(double2 a, int m)
{
	constexpr double c = 0.70710678118654752440;
	switch (m) {
	case 0: return a;
	case 1: return make_double2(c * (a.x + a.y), c * (a.y - a.x));
	case 2: return make_double2(a.y, -a.x);
	default: return make_double2(c * (a.y - a.x), -c * (a.x + a.y));
	}
}

// The twiddles of stages s+1 .. s+K at offset lo < 2^s: w[t-1] =
// exp(-2 pi i lo / 2^(s+t)), from one table entry and K-1 squarings.  Stage
// t's butterfly j (j mod 2^(t-1) = m) takes w[t-1] exp(-2 pi i m / 2^t).
template <int K>
__device__ __forceinline__ void pass_twiddles(const Tw &T, int s, int lo, double2 *w)
{
	w[K - 1] = T.at(s + K, lo);
#pragma unroll
	for (int t = K - 2; t >= 0; --t) w[t] = csq(w[t + 1]);
}

// Radix-2 stages s+1 .. s+K of a forward DIT on the 2^K elements a thread
// holds in registers: x[j] is element e0 + (j << s), lo = e0 mod 2^s.  Stage
// S combines i and i + 2^(S-1) with twiddle exp(-2 pi i (i mod 2^(S-1)) / 2^S).
template <int K>
__device__ __forceinline__ void dit_regs(double2 *x, int s, int lo, const Tw &T)
{
	constexpr int R = 1 << K;
	double2 w[K];
	pass_twiddles<K>(T, s, lo, w);
#pragma unroll
	for (int t = 1; t <= K; ++t) {
		const int h = 1 << (t - 1);
#pragma unroll
		for (int j = 0; j < R; ++j) {
			if (j & h) continue;
			const double2 tw = rot8(w[t - 1], (j & (h - 1)) << (3 - t));
			const double2 u = x[j], v = cmul(x[j + h], tw);
			x[j] = cadd(u, v);
			x[j + h] = csub(u, v);
		}
	}
}

// Stages 1-3 of a forward DIT on elements e0 .. e0 + 7 (s = 0): their
// twiddles are 1, -i and (+-1 - i) / sqrt(2), applied as constants.
__device__ __forceinline__ void dit8_first(double2 *x)
{
	constexpr double c = 0.70710678118654752440;
#pragma unroll
	for (int j = 0; j < 8; j += 2) {
		const double2 u = x[j], v = x[j + 1];
		x[j] = cadd(u, v);
		x[j + 1] = csub(u, v);
	}
#pragma unroll
	for (int j = 0; j < 8; j += 4) {
		double2 u = x[j], v = x[j + 2];
		x[j] = cadd(u, v);
		x[j + 2] = csub(u, v);
		u = x[j + 1];
		v = make_double2(x[j + 3].y, -x[j + 3].x);  // * -i
		x[j + 1] = cadd(u, v);
		x[j + 3] = csub(u, v);
	}
	const double2 v1 = make_double2(c * (x[5].x + x[5].y), c * (x[5].y - x[5].x));   // * (1 - i) / sqrt 2
	const double2 v2 = make_double2(x[6].y, -x[6].x);                                 // * -i
	const double2 v3 = make_double2(c * (x[7].y - x[7].x), -c * (x[7].x + x[7].y));  // * (-1 - i) / sqrt 2
	const double2 u0 = x[0], u1 = x[1], u2 = x[2], u3 = x[3];
	x[0] = cadd(u0, x[4]);
	x[4] = csub(u0, x[4]);
	x[1] = cadd(u1, v1);
	x[5] = csub(u1, v1);
	x[2] = cadd(u2, v2);
	x[6] = csub(u2, v2);
	x[3] = cadd(u3, v3);
	x[7] = csub(u3, v3);
}

// Stages 3-1 of an inverse DIF on elements e0 .. e0 + 7 (s = 0), with the
// conjugate twiddles 1, i and (+-1 + i) / sqrt(2) as constants.
__device__ __forceinline__ void dif8_last(double2 *x)
{
	constexpr double c = 0.70710678118654752440;
	{
		const double2 d0 = csub(x[0], x[4]), d1 = csub(x[1], x[5]), d2 = csub(x[2], x[6]), d3 = csub(x[3], x[7]);
#pragma unroll
		for (int j = 0; j < 4; ++j) x[j] = cadd(x[j], x[j + 4]);
		x[4] = d0;
		x[5] = make_double2(c * (d1.x - d1.y), c * (d1.x + d1.y));    // * (1 + i) / sqrt 2
		x[6] = make_double2(-d2.y, d2.x);                             // * i
		x[7] = make_double2(-c * (d3.x + d3.y), c * (d3.x - d3.y));   // * (-1 + i) / sqrt 2
	}
#pragma unroll
	for (int j = 0; j < 8; j += 4) {
		const double2 d0 = csub(x[j], x[j + 2]), d1 = csub(x[j + 1], x[j + 3]);
		x[j] = cadd(x[j], x[j + 2]);
		x[j + 1] = cadd(x[j + 1], x[j + 3]);
		x[j + 2] = d0;
		x[j + 3] = make_double2(-d1.y, d1.x);  // * i
	}
#pragma unroll
	for (int j = 0; j < 8; j += 2) {
		const double2 u = x[j], v = x[j + 1];
		x[j] = cadd(u, v);
		x[j + 1] = csub(u, v);
	}
}

// Radix-2 stages s+K down to s+1 of an inverse DIF (conjugate twiddles,
// unscaled) in registers: stage S maps (u, v) at i and i + 2^(S-1) to
// (u + v, (u - v) conj w).
template <int K>
__device__ __forceinline__ void dif_regs(double2 *x, int s, int lo, const Tw &T)
{
	constexpr int R = 1 << K;
	double2 w[K];
	pass_twiddles<K>(T, s, lo, w);
#pragma unroll
	for (int t = K; t >= 1; --t) {
		const int h = 1 << (t - 1);
#pragma unroll
		for (int j = 0; j < R; ++j) {
			if (j & h) continue;
			const double2 tw = rot8(w[t - 1], (j & (h - 1)) << (3 - t));
			const double2 u = x[j], v = x[j + h];
			x[j] = cadd(u, v);
			x[j + h] = cmul(csub(u, v), conjg(tw));
		}
	}
}

// One pass of K radix-2 stages (s+1 .. s+K) over `ncol` sequences of length
// n in LDS (sequence c at data + c*stride, element e at swz(e)): each thread
// loads the 2^K elements a group combines, runs the stages in registers and
// stores them back, so LDS is read and written once per pass.
template <int K>
__device__ __forceinline__ void dit_pass(double2 *data, int logn, int ncol, int stride, int s, const Tw &T)
{
	constexpr int R = 1 << K;
	const int sub = 1 << s;                     // 2^s: element spacing inside a group
	const int total = ncol << (logn - K);       // groups: n / R per sequence
	for (int gi = threadIdx.x; gi < total; gi += blockDim.x) {
		const int c = gi >> (logn - K), g = gi - (c << (logn - K));
		const int lo = g & (sub - 1), hi = g >> s;
		double2 *col = data + c * stride;
		// swz is linear over XOR and e0, j << s occupy disjoint bits: the
		// addresses are swz(e0) ^ swz(j << s), the second term uniform
		const int b0 = swz((hi << (s + K)) + lo);
		double2 x[R];
#pragma unroll
		for (int j = 0; j < R; ++j) x[j] = col[b0 ^ swz(j << s)];
		if (K == 3 && s == 0) dit8_first(x);
		else dit_regs<K>(x, s, lo, T);
#pragma unroll
		for (int j = 0; j < R; ++j) col[b0 ^ swz(j << s)] = x[j];
	}
	__syncthreads();
}

// In-place forward DIT, bit-reversed input, natural-order output: passes of
// three stages from stage 1 up, the remaining one or two stages last.
__device__ inline void fft_dit(double2 *data, int logn, int ncol, int stride, const Tw &T, int s = 0)
{
	for (; s + 3 <= logn; s += 3) dit_pass<3>(data, logn, ncol, stride, s, T);
	if (logn - s == 2) dit_pass<2>(data, logn, ncol, stride, s, T);
	else if (logn - s == 1) dit_pass<1>(data, logn, ncol, stride, s, T);
}

// Position j of a three-bit reversal.
__device__ __forceinline__ int rev3(int j) { return ((j & 1) << 2) | (j & 2) | (j >> 2); }

// Forward DIT of `nseq` sequences of length n = 2^logn >= 8 whose first pass
// (stages 1-3) reads its inputs through `load(c, i)` (element i of sequence
// c, natural order) instead of LDS: group g combines the elements at
// positions 8g + j of the bit-reversed sequence, inputs
// rev3(j) * n/8 + bitrev(g), so no bit-reversed scatter goes through LDS.
// Sequence c ends at data + (c << logn) in natural order.
template <class F>
__device__ __forceinline__ void fft_dit_gather(double2 *data, int logn, int nseq, const Tw &T, F load)
{
	const int lg = logn - 3;
	for (int gi = threadIdx.x; gi < (nseq << lg); gi += blockDim.x) {
		const int c = gi >> lg, g = gi - (c << lg);
		const int bg = lg ? (int)bitrev(g, lg) : 0;
		double2 x[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) x[j] = load(c, (rev3(j) << lg) + bg);
		dit8_first(x);
		double2 *col = data + (c << logn);
#pragma unroll
		for (int j = 0; j < 8; ++j) col[swz(8 * g + j)] = x[j];
	}
	__syncthreads();
	fft_dit(data, logn, nseq, 1 << logn, T, 3);
}

// One inverse DIF pass of K stages, top down to top-K+1 (indices differing in
// bits top-1 .. top-K, spacing 2^(top-K)).
template <int K>
__device__ __forceinline__ void dif_pass(double2 *data, int logn, int ncol, int stride, int top, const Tw &T)
{
	constexpr int R = 1 << K;
	const int s = top - K;
	const int sub = 1 << s;
	const int total = ncol << (logn - K);
	for (int gi = threadIdx.x; gi < total; gi += blockDim.x) {
		const int c = gi >> (logn - K), g = gi - (c << (logn - K));
		const int lo = g & (sub - 1), hi = g >> s;
		double2 *col = data + c * stride;
		const int b0 = swz((hi << top) + lo);  // as in dit_pass
		double2 x[R];
#pragma unroll
		for (int j = 0; j < R; ++j) x[j] = col[b0 ^ swz(j << s)];
		if (K == 3 && s == 0) dif8_last(x);
		else dif_regs<K>(x, s, lo, T);
#pragma unroll
		for (int j = 0; j < R; ++j) col[b0 ^ swz(j << s)] = x[j];
	}
	__syncthreads();
}

// In-place inverse DIF, natural-order input, bit-reversed output.
__device__ inline void ifft_dif(double2 *data, int logn, int ncol, int stride, const Tw &T)
{
	int top = logn;
	for (; top - 3 >= 0; top -= 3) dif_pass<3>(data, logn, ncol, stride, top, T);
	if (top == 2) dif_pass<2>(data, logn, ncol, stride, top, T);
	else if (top == 1) dif_pass<1>(data, logn, ncol, stride, top, T);
}

// Inverse DIF of `nseq` sequences of length n = 2^logn >= 8 (sequence c at
// data + (c << logn)) whose last pass (stages 3-1) hands its results to
// `emit(c, g, bg, x)` instead of LDS: x[j] is output rev3(j) * n/8 + bg,
// bg = bitrev(g).  The remaining one or two stages run first.
template <class F>
__device__ __forceinline__ void ifft_dif_scatter(double2 *data, int logn, int nseq, const Tw &T, F emit)
{
	const int stride = 1 << logn;
	int top = logn;
	if (logn % 3 == 2) {
		dif_pass<2>(data, logn, nseq, stride, top, T);
		top -= 2;
	}
	else if (logn % 3 == 1) {
		dif_pass<1>(data, logn, nseq, stride, top, T);
		top -= 1;
	}
	for (; top > 3; top -= 3) dif_pass<3>(data, logn, nseq, stride, top, T);
	const int lg = logn - 3;
	for (int gi = threadIdx.x; gi < (nseq << lg); gi += blockDim.x) {
		const int c = gi >> lg, g = gi - (c << lg);
		const double2 *col = data + (c << logn);
		double2 x[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) x[j] = col[swz(8 * g + j)];
		dif8_last(x);
		emit(c, g, lg ? (int)bitrev(g, lg) : 0, x);
	}
}

}  // namespace

// Forward transforms of `np` pairs of real rows, pair p at rows 2p, 2p + 1 of
// `src` (row 2p + 1 only below nrows; pairs from pmax on are not loaded or
// stored), columns [x0, x0 + nx) of the frame holding data, into the
// half-spectrum rows 2p, 2p + 1 of `out` (pitch P):
// A = (Z[k] + conj Z[N-k]) / 2 and B = (Z[k] - conj Z[N-k]) / 2i, k = 0..N/2,
// for z = a + i b.  Sequence p sits at sh + p * N.
template <class Src>
__device__ __forceinline__ void rows_fwd_pairs(double2 *sh, const Tw &T, int N, int L, int np, const Src *src,
                                               long ld, int x0, int nx, int nrows, int pmax, double2 *out, int P)
{
	const int M = N / 2 + 1;
	fft_dit_gather(sh, L, np, T, [&](int p, int k) {
		const int xi = k - x0;
		const bool ok = p < pmax && xi >= 0 && xi < nx;
		const bool okb = ok && 2 * p + 1 < nrows;
		const long o = ok ? (long)(2 * p) * ld + xi : 0;
		const double a = (double)src[o], bv = (double)src[okb ? o + ld : 0];
		return make_double2(ok ? a : 0.0, okb ? bv : 0.0);
	});
	for (int t = threadIdx.x; t < np * M; t += blockDim.x) {
		const int p = t / M, k = t - p * M;
		if (p >= pmax) continue;
		const double2 *z0 = sh + (p << L);
		const double2 z = z0[swz(k)], zc = conjg(z0[swz((N - k) & (N - 1))]);
		double2 *outa = out + (long)(2 * p) * P;
		outa[k] = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y + zc.y));
		// (z - zc) / 2i = (-i / 2) (z - zc)
		if (2 * p + 1 < nrows) outa[P + k] = make_double2(0.5 * (z.y - zc.y), -0.5 * (z.x - zc.x));
	}
}

// Forward rows, fft_pairs_per_block(N1) pairs of input rows per block, each
// pair packed as z = a + i b:
//   mode 0: canvas rows (DW x DH) into frame rows [0, DH)
//   mode 1: w rows (W x H) at column padx into frame rows [pady, pady + H)
__global__ __launch_bounds__(256) void k_fft_rows_fwd(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                      const int *__restrict__ row_off, int ninst, int mode,
                                                      const double2 *__restrict__ tw, int twn)
{
	extern __shared__ double2 sh[];
	const int b = blockIdx.x;
	const int inst = owner(row_off, ninst, b);
	const InstDev &I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int N = g.fft_n1, L = g.fft_l1, P = fft_pitch(N), np = fft_pairs_per_block(N);
	const Tw T = stage_table(sh + np * N, L, tw, twn);
	const int p0 = (b - row_off[inst]) * np;  // first pair of the block
	// straight-line selects of the mode's geometry
	const bool adj = mode == 1;
	const int nrows = adj ? g.H : g.DH;
	const int width = adj ? g.W : g.DW;
	const int x0 = adj ? g.padx : 0;
	const int y0 = adj ? g.pady : 0;
	const int r0 = 2 * p0;
	const float *src = (adj ? I.w : I.canvas) + (long)r0 * width;
	const int pmax = (nrows - r0 + 1) / 2;
	rows_fwd_pairs<float>(sh, T, N, L, np, src, width, x0, width, nrows - r0, pmax, I.fbuf + (long)(y0 + r0) * P, P);
}

// Forward rows of the FP64 canvas (mode 0 of k_fft_rows_fwd for galaxies with
// GalaxyDev::canvas64), a kernel of its own so that no pointer is selected on
// a branch.
__global__ __launch_bounds__(256) void k_fft_rows_fwd64(const GalaxyDev *__restrict__ gals,
                                                        const InstDev *__restrict__ insts,
                                                        const int *__restrict__ row_off, int ninst,
                                                        const double2 *__restrict__ tw, int twn)
{
	extern __shared__ double2 sh[];
	const int b = blockIdx.x;
	const int inst = owner(row_off, ninst, b);
	const InstDev &I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int N = g.fft_n1, L = g.fft_l1, P = fft_pitch(N), np = fft_pairs_per_block(N);
	const Tw T = stage_table(sh + np * N, L, tw, twn);
	const int r0 = 2 * (b - row_off[inst]) * np;
	const double *src = I.canvas64 + (long)r0 * g.DW;
	const int pmax = (g.DH - r0 + 1) / 2;
	rows_fwd_pairs<double>(sh, T, N, L, np, src, g.DW, 0, g.DW, g.DH - r0, pmax, I.fbuf + (long)r0 * P, P);
}

// Forward rows of a galaxy's real FP64 PSF frame (N1 x N2, `src64`) into the
// half-spectrum frame `dst64`, fft_pairs_per_block(N1) pairs of rows per block.
__global__ __launch_bounds__(256) void k_fft_psf_rows(const GalaxyDev *__restrict__ g, const double2 *__restrict__ tw,
                                                      int twn, const double *__restrict__ src64, double2 *dst64)
{
	extern __shared__ double2 sh[];
	const int N = g->fft_n1, L = g->fft_l1, P = fft_pitch(N), np = fft_pairs_per_block(N);
	const Tw T = stage_table(sh + np * N, L, tw, twn);
	const int r0 = 2 * np * blockIdx.x;
	const int nrows = g->fft_n2 - r0;
	rows_fwd_pairs<double>(sh, T, N, L, np, src64 + (long)r0 * N, N, 0, N, nrows, (nrows + 1) / 2,
	                       dst64 + (long)r0 * P, P);
}

// Column pass over CB adjacent half-spectrum columns of one instance.  Frame
// rows [in_lo, in_hi) hold row spectra, the others are zero.  op 0: forward
// only (all N2 rows of `dst64` written: the PSF spectrum); op 1: product with
// the PSF spectrum; op 2: with its conjugate.  The inverse column transform
// then leaves rows [out_lo, out_hi) in the frame.  Global accesses run along
// rows (CB adjacent entries); columns sit in LDS at a stride of N2 + 2.
__global__ __launch_bounds__(kColThreads, 4) void k_fft_cols(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                  const int2 *__restrict__ blocks, int cb, int op,
                                                  int in_mode, int out_mode, const double2 *__restrict__ tw, int twn,
                                                  double2 *dst64)
{
	extern __shared__ double2 sh[];
	const int2 blk = blocks[blockIdx.x];
	const InstDev &I = insts[blk.x];
	const GalaxyDev &g = gals[I.galaxy];
	const int M = g.fft_n1 / 2 + 1, P = fft_pitch(g.fft_n1), N = g.fft_n2, L = g.fft_l2;
	// column stride: 2 mod 16 entries keeps the columns-fastest pass on
	// distinct LDS banks
	const int S = N + 2;
	const Tw T = stage_table(sh + cb * S, L, tw, twn);
	const int c0 = blk.y * cb;
	double2 *frame = (op == 0) ? dst64 : I.fbuf;
	// rows holding data: 0 canvas rows [0, DH); 1 image rows [pady, pady + H); 2 all rows
	const int in_lo = in_mode == 1 ? g.pady : 0;
	const int in_hi = in_mode == 0 ? g.DH : (in_mode == 1 ? g.pady + g.H : N);
	// cb is a power of two: entry t of the tile is row t >> lcb, column t & (cb - 1)
	const int lcb = __ffs(cb) - 1, total = cb << L;
	// from 64 points on (and for the convolutions), stages 1-3 run on the
	// frame's rows as they are loaded and stages 3-1 of the inverse before
	// the rows are stored, as radix-8 groups: lane l takes column l mod cb,
	// so that each row's cb columns are one coalesced run; group g combines
	// rows rev3(j) * N/8 + bitrev(g)
	const bool edge = op != 0 && L >= 6;
	const int lg3 = L - 3;
	if (edge) {
		for (int gi = threadIdx.x; gi < (cb << lg3); gi += blockDim.x) {
			const int c = gi & (cb - 1), gg = gi >> lcb;
			const int bg = (int)bitrev(gg, lg3);
			const bool okc = c0 + c < M;
			double2 x[8];
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				const int y = (rev3(j) << lg3) + bg;
				const bool ok = okc && y >= in_lo && y < in_hi;
				const double2 v = frame[ok ? (long)y * P + c0 + c : 0];
				x[j] = ok ? v : make_double2(0.0, 0.0);
			}
			dit8_first(x);
			double2 *col = sh + c * S;
#pragma unroll
			for (int j = 0; j < 8; ++j) col[swz(8 * gg + j)] = x[j];
		}
		__syncthreads();
		// stages 4 .. L-3 (the last three run around the product below)
		int s = 3;
		for (; s + 3 <= lg3; s += 3) dit_pass<3>(sh, L, cb, S, s, T);
		if (lg3 - s == 2) dit_pass<2>(sh, L, cb, S, s, T);
		else if (lg3 - s == 1) dit_pass<1>(sh, L, cb, S, s, T);
	}
	else {
		for (int base = threadIdx.x; base < total; base += kBatch * blockDim.x) {
			double2 v[kBatch];
#pragma unroll
			for (int u = 0; u < kBatch; ++u) {
				const int t = base + u * blockDim.x;
				const int y = t >> lcb, c = t & (cb - 1);
				const bool ok = t < total && y >= in_lo && y < in_hi && c0 + c < M;
				const double2 x = frame[ok ? (long)y * P + c0 + c : 0];
				v[u] = ok ? x : make_double2(0.0, 0.0);
			}
#pragma unroll
			for (int u = 0; u < kBatch; ++u) {
				const int t = base + u * blockDim.x;
				if (t < total) sh[(t & (cb - 1)) * S + swz(bitrev(t >> lcb, L))] = v[u];
			}
		}
		__syncthreads();
		if (op == 0) {
			fft_dit(sh, L, cb, S, T);
			for (int t = threadIdx.x; t < total; t += blockDim.x) {
				const int y = t >> lcb, c = t & (cb - 1);
				if (c0 + c < M) frame[(long)y * P + c0 + c] = sh[c * S + swz(y)];
			}
			return;
		}
		// forward DIT with the one or two odd stages first, then passes of
		// three up to stage L-3
		int s = L % 3;
		if (s == 2) dit_pass<2>(sh, L, cb, S, 0, T);
		else if (s == 1) dit_pass<1>(sh, L, cb, S, 0, T);
		for (; s + 3 < L; s += 3) dit_pass<3>(sh, L, cb, S, s, T);
	}
	// the forward DIT's last three stages (L-2 .. L) and the inverse's first
	// three (L .. L-2) combine the same eight elements, so they run in one
	// step around the product with the PSF spectrum
	const int lg = L - 3;
	const double2 *spec = g.fft_spec;
	for (int gi = threadIdx.x; gi < (cb << lg); gi += blockDim.x) {
		// columns fastest: a wave reads whole 128-byte runs of spectrum rows
		const int c = gi & (cb - 1), lo = gi >> lcb;
		const bool ok = c0 + c < M;
		double2 *col = sh + c * S;
		double2 x[8], sp[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) x[j] = col[swz(lo) ^ swz(j << lg)];  // lo < 2^lg: disjoint bits
		dit_regs<3>(x, lg, lo, T);
		// spectrum entries four at a time (the registers of four waves per SIMD)
#pragma unroll
		for (int h = 0; h < 8; h += 4) {
#pragma unroll
			for (int j = h; j < h + 4; ++j) sp[j] = spec[ok ? (long)(lo + (j << lg)) * P + c0 + c : 0];
#pragma unroll
			for (int j = h; j < h + 4; ++j) x[j] = (op == 1) ? cmul(x[j], sp[j]) : cmulc(x[j], sp[j]);
		}
		dif_regs<3>(x, lg, lo, T);
#pragma unroll
		for (int j = 0; j < 8; ++j) col[swz(lo) ^ swz(j << lg)] = x[j];
	}
	__syncthreads();
	// rows written back: 0 image rows [pady, pady + H); 1 canvas rows [0, DH)
	const int out_lo = out_mode == 0 ? g.pady : 0;
	const int out_hi = out_mode == 0 ? g.pady + g.H : g.DH;
	if (edge) {
		// inverse stages L-3 .. 4, then 3 .. 1 on the way out
		int top = lg;
		for (; top - 3 >= 3; top -= 3) dif_pass<3>(sh, L, cb, S, top, T);
		if (top - 3 == 2) dif_pass<2>(sh, L, cb, S, top, T);
		else if (top - 3 == 1) dif_pass<1>(sh, L, cb, S, top, T);
		for (int gi = threadIdx.x; gi < (cb << lg3); gi += blockDim.x) {
			const int c = gi & (cb - 1), gg = gi >> lcb;
			const int bg = (int)bitrev(gg, lg3);
			const double2 *col = sh + c * S;
			double2 x[8];
#pragma unroll
			for (int j = 0; j < 8; ++j) x[j] = col[swz(8 * gg + j)];
			dif8_last(x);
			// x[j] is output row rev3(j) * N/8 + bg
			if (c0 + c < M) {
#pragma unroll
				for (int j = 0; j < 8; ++j) {
					const int y = (rev3(j) << lg3) + bg;
					if (y >= out_lo && y < out_hi) frame[(long)y * P + c0 + c] = x[j];
				}
			}
		}
		return;
	}
	int top = lg;
	for (; top - 3 >= 0; top -= 3) dif_pass<3>(sh, L, cb, S, top, T);
	if (top == 2) dif_pass<2>(sh, L, cb, S, top, T);
	else if (top == 1) dif_pass<1>(sh, L, cb, S, top, T);
	for (int t = threadIdx.x; t < total; t += blockDim.x) {
		const int y = t >> lcb, c = t & (cb - 1);
		if (y >= out_lo && y < out_hi && c0 + c < M) frame[(long)y * P + c0 + c] = sh[c * S + swz(bitrev(y, L))];
	}
}

// Inverse rows, fft_pairs_per_block(N1) pairs of frame rows per block, each
// pair (half-spectra A, B) packed as Z = A + i B over the full length
// (Z[k] = conj A[N-k] + i conj B[N-k] above N/2); the real part of the
// inverse is the first row, the imaginary part the second.  Epilogues:
//   mode 0 (likelihood): image rows (frame rows pady + r); model = value /
//     (N1 N2) at column padx + x; likelihood terms into one partial per
//     block, w written
//   mode 3: mode 0 for batches whose galaxies all take a Gaussian
//     likelihood (chi^2 or normal), without the Poisson and t code
//   mode 1 (model image): canvas rows; G = model masked by calc
//   mode 2 (adjoint): canvas rows; G = value / (N1 N2)
template <int MODE>
__global__ __launch_bounds__(256) void k_fft_rows_inv(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                      const int2 *__restrict__ blocks,
                                                      const double2 *__restrict__ tw, int twn,
                                                      const int *__restrict__ lk_base, double *partial)
{
	constexpr int mode = MODE == 3 ? 0 : MODE;
	extern __shared__ double2 sh[];
	const int2 bi = blocks[blockIdx.x];
	const int inst = bi.x;
	const InstDev &I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int N = g.fft_n1, L = g.fft_l1, M = N / 2 + 1, P = fft_pitch(N), np = fft_pairs_per_block(N);
	const Tw T = stage_table(sh + np * N, L, tw, twn);
	const int blk = bi.y;
	const int r0 = 2 * np * blk;  // first output row of the block (image rows in mode 0)
	const int nrows = mode == 0 ? g.H : g.DH;
	const int y0 = mode == 0 ? r0 + g.pady : r0;  // its frame row
	const int total = np * N;
	const double2 *rows = I.fbuf + (long)y0 * P;
	for (int base = threadIdx.x; base < total; base += kBatch * blockDim.x) {
		double2 va[kBatch], vb[kBatch];
#pragma unroll
		for (int u = 0; u < kBatch; ++u) {
			const int t = base + u * blockDim.x;
			const int p = t >> L, k = t - (p << L);
			const int r = r0 + 2 * p;
			const bool ok = t < total && r < nrows, okb = ok && r + 1 < nrows;
			// above N/2 the entries are conjugates of the ones at N - k
			const long o = (long)(2 * p) * P + (k < M ? k : N - k);
			va[u] = rows[ok ? o : 0];
			vb[u] = rows[okb ? o + P : 0];
			if (!ok) va[u] = make_double2(0.0, 0.0);
			if (!okb) vb[u] = make_double2(0.0, 0.0);
		}
#pragma unroll
		for (int u = 0; u < kBatch; ++u) {
			const int t = base + u * blockDim.x;
			if (t < total) {
				const int p = t >> L, k = t - (p << L);
				const double2 a = k < M ? va[u] : conjg(va[u]), bv = k < M ? vb[u] : conjg(vb[u]);
				sh[(p << L) + swz(k)] = make_double2(a.x - bv.y, a.y + bv.x);  // A + i B
			}
		}
	}
	__syncthreads();
	const double scale = 1.0 / ((double)g.fft_n1 * (double)g.fft_n2);
	const int lg = L - 3;
	if constexpr (mode == 0) {
		double part = 0;
		const bool pois = MODE == 0 && g.likefunc == LIKE_POIS, tl = MODE == 0 && g.likefunc == LIKE_T;
		const double s = g.pois_scale;
		// x[j] is frame column rev3(j) * N/8 + bg of rows 2p (real part) and
		// 2p + 1 (imaginary part); image column = frame column - padx
		ifft_dif_scatter(sh, L, np, T, [&](int p, int, int bg, const double2 *x) {
#pragma unroll
			for (int half = 0; half < 4; ++half) {
				const int row = half >> 1, j0 = (half & 1) * 4;
				const int r = r0 + 2 * p + row;
				// the pixels' inputs for four columns first, then the terms
				PixelIn q[8];
				double ax[8];
#pragma unroll
				for (int j = j0; j < j0 + 4; ++j) {
					const int xi = (rev3(j) << lg) + bg - g.padx;
					const bool ok = r < nrows && xi >= 0 && xi < g.W;
					q[j] = g.pix[ok ? (long)r * g.W + xi : 0];
					ax[j] = I.has_aux && ok ? (double)I.aux[(long)(y0 + 2 * p + row) * g.DW + xi + g.padx] : 0.0;
				}
#pragma unroll
				for (int j = j0; j < j0 + 4; ++j) {
					const int xi = (rev3(j) << lg) + bg - g.padx;
					if (r >= nrows || xi < 0 || xi >= g.W) continue;
					float wv = 0.f;
					if (q[j].flags & 1u) {
						const bool in_calc = (q[j].flags & 2u) != 0;
						const double m = in_calc ? (row ? x[j].y : x[j].x) * scale + ax[j] : 0.0;
						const double d = q[j].data;
						if (pois) {
							const double mm = s * m, dd = s * d;
							part += mm - dd * log(mm);
							wv = (float)(-s * (1.0 - d / m));
						}
						else {
							const double diff = d - m;
							const double iv = q[j].ivar;
							part += diff * diff * iv;
							wv = (float)(diff * iv);
							// t: the standardised residual, for k_like_t
							if (tl) wv = (float)(diff * sqrt(iv));
						}
						if (!in_calc && !tl) wv = 0.f;
					}
					I.w[(long)r * g.W + xi] = wv;
				}
			}
		});
		__shared__ double red[4];
#pragma unroll
		for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
		if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
		__syncthreads();
		if (threadIdx.x == 0) {
			double t = 0;
			for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
			partial[lk_base[inst] + blk] = t;
		}
		return;
	}
	ifft_dif_scatter(sh, L, np, T, [&](int p, int, int bg, const double2 *x) {
#pragma unroll
		for (int row = 0; row < 2; ++row) {
			const int r = r0 + 2 * p + row;
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				const int xc = (rev3(j) << lg) + bg;
				if (r >= nrows || xc >= g.DW) continue;
				const long o = (long)r * g.DW + xc;
				double v = (row ? x[j].y : x[j].x) * scale;
				if (mode == 1 && g.has_calc && !g.calc[o]) v = 0;
				if (mode == 1 && I.has_aux) v += (double)I.aux[o];
				I.G[o] = (float)v;
			}
		}
	});
}

// ------------------------------------------------------------------ launchers

// Raises a kernel's dynamic LDS limit once per device (and only when it
// grows), so that launches recorded into a graph make no attribute calls.
static int ensure_lds(const void *fn, size_t lds, size_t *done)
{
	if (lds <= 64 * 1024) return 0;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess) return (int)e;
	if (dev < 0 || dev >= 64) return (int)hipErrorInvalidDevice;
	if (done[dev] >= lds) return 0;
	e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
	if (e != hipSuccess) return (int)e;
	done[dev] = lds;
	return 0;
}
static size_t lds_rows_fwd[64], lds_rows_fwd64[64], lds_psf_rows[64], lds_cols[64], lds_rows_inv[64], lds_rows_inv1[64], lds_rows_inv2[64], lds_rows_inv3[64];

static int cols_per_block(int n2) { return n2 >= 4096 ? 1 : (n2 >= 2048 ? 2 : (4096 / n2 > 8 ? 8 : 4096 / n2)); }

// LDS of a row pass: the block's pairs and the twiddle table (the largest
// frame sets it)
static size_t row_lds(int nmax) { return sizeof(double2) * ((size_t)fft_pairs_per_block(nmax) * nmax + nmax / 2); }

int launch_fft_rows_fwd(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int *row_off,
                        int ninst, int mode, const double2 *tw, int twn, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	if (mode != 0 && mode != 1) return (int)hipErrorInvalidValue;
	const size_t lds = row_lds(nmax);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
	if (int e = ensure_lds((const void *)k_fft_rows_fwd, lds, lds_rows_fwd)) return e;
	k_fft_rows_fwd<<<nblocks, 256, lds, st>>>(gals, insts, row_off, ninst, mode, tw, twn);
	return (int)hipGetLastError();
}

int launch_fft_rows_fwd64(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int *row_off,
                          int ninst, const double2 *tw, int twn, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	const size_t lds = row_lds(nmax);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
	if (int e = ensure_lds((const void *)k_fft_rows_fwd64, lds, lds_rows_fwd64)) return e;
	k_fft_rows_fwd64<<<nblocks, 256, lds, st>>>(gals, insts, row_off, ninst, tw, twn);
	return (int)hipGetLastError();
}

int launch_fft_psf_rows(int n1, int n2, const GalaxyDev *gal, const double2 *tw, int twn, const double *src64,
                        double2 *dst64, hipStream_t st)
{
	const size_t lds = row_lds(n1);
	if (lds > 160 * 1024 || n2 < 2) return (int)hipErrorInvalidValue;
	if (int e = ensure_lds((const void *)k_fft_psf_rows, lds, lds_psf_rows)) return e;
	const int np = fft_pairs_per_block(n1);
	k_fft_psf_rows<<<(n2 / 2 + np - 1) / np, 256, lds, st>>>(gal, tw, twn, src64, dst64);
	return (int)hipGetLastError();
}

int launch_fft_cols(int nblocks, int n2max, int cb, const GalaxyDev *gals, const InstDev *insts, const int2 *blocks,
                    int op, int in_mode, int out_mode, const double2 *tw, int twn, double2 *dst64, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	const size_t lds = sizeof(double2) * ((size_t)cb * (n2max + 2) + n2max / 2);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
	if (int e = ensure_lds((const void *)k_fft_cols, lds, lds_cols)) return e;
	k_fft_cols<<<nblocks, kColThreads, lds, st>>>(gals, insts, blocks, cb, op, in_mode, out_mode, tw, twn, dst64);
	return (int)hipGetLastError();
}

int launch_fft_rows_inv(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int2 *blocks,
                        int mode, const double2 *tw, int twn, const int *lk_base, double *partial, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	if (mode < 0 || mode > 3) return (int)hipErrorInvalidValue;
	const size_t lds = row_lds(nmax);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
	if (mode == 0) {
		if (int e = ensure_lds((const void *)k_fft_rows_inv<0>, lds, lds_rows_inv)) return e;
		k_fft_rows_inv<0><<<nblocks, 256, lds, st>>>(gals, insts, blocks, tw, twn, lk_base, partial);
	}
	else if (mode == 1) {
		if (int e = ensure_lds((const void *)k_fft_rows_inv<1>, lds, lds_rows_inv1)) return e;
		k_fft_rows_inv<1><<<nblocks, 256, lds, st>>>(gals, insts, blocks, tw, twn, lk_base, partial);
	}
	else if (mode == 2) {
		if (int e = ensure_lds((const void *)k_fft_rows_inv<2>, lds, lds_rows_inv2)) return e;
		k_fft_rows_inv<2><<<nblocks, 256, lds, st>>>(gals, insts, blocks, tw, twn, lk_base, partial);
	}
	else {
		if (int e = ensure_lds((const void *)k_fft_rows_inv<3>, lds, lds_rows_inv3)) return e;
		k_fft_rows_inv<3><<<nblocks, 256, lds, st>>>(gals, insts, blocks, tw, twn, lk_base, partial);
	}
	return (int)hipGetLastError();
}

int fft_cols_per_block(int n2) { return cols_per_block(n2); }

}  // namespace profit_amd
