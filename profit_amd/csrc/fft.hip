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
// the inverse ones, three radix-2 stages per pass in registers.  Two layouts
// keep the LDS banks free of conflicts: element e of a sequence sits at
// swz(e), an XOR of its low bits with the next ones, so the strided accesses
// of every pass fall on distinct banks; and the twiddles are staged one table
// per stage, in the order a stage reads them.  The row passes run one pair of
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
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 conjg(double2 a) { return make_double2(a.x, -a.y); }

__device__ inline double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
	return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// LDS position of element e of a sequence: bits 0-2 ^= bits 3-5 and bit 3 ^=
// bit 6, a bijection on every power-of-two range.  A pass reads elements at
// spacing 1, 8 or 64 (for 2^9 points); after the swizzle the 16 lanes that
// share a bank cycle address distinct 16-byte slots in each case.
__device__ __forceinline__ int swz(int e) { return e ^ ((e >> 3) & 7) ^ (((e >> 6) & 1) << 3); }

// Twiddles of an n-point transform in LDS, one table per radix-2 stage S =
// 1..log2(n): wt[2^(S-1) - 1 + k] = exp(-2 pi i k / 2^S), k < 2^(S-1) (n - 1
// entries), taken from the global table of the largest transform (twn points).
__device__ inline void stage_twiddles(double2 *wt, int n, int logn, const double2 *__restrict__ tw, int twn)
{
	const int step = twn / n;
	for (int i = threadIdx.x; i < n - 1; i += blockDim.x) {
		const int S = 32 - __clz(i + 1);  // floor(log2(i + 1)) + 1
		const int k = i + 1 - (1 << (S - 1));
		wt[i] = tw[(k << (logn - S)) * step];
	}
}

__device__ __forceinline__ double2 twiddle(const double2 *wt, int S, int k) { return wt[(1 << (S - 1)) - 1 + k]; }

// In-place forward DIT of `ncol` sequences of length n (sequence c at
// data + c*stride, element e at swz(e)), bit-reversed input, natural-order
// output.  The log2(n) radix-2 stages run in passes of up to three: each
// thread holds the eight (four, two) elements that the pass's stages combine
// in registers, so LDS is read and written once per pass instead of once per
// stage.  Stage S combines i and i + 2^(S-1) with twiddle
// exp(-2 pi i (i mod 2^(S-1)) / 2^S).
template <int K>
__device__ __forceinline__ void dit_pass(double2 *data, int logn, int ncol, int stride, int s, const double2 *wt)
{
	constexpr int R = 1 << K;
	const int sub = 1 << s;                     // 2^s: element spacing inside a group
	const int total = ncol << (logn - K);       // groups: n / R per sequence
	for (int gi = threadIdx.x; gi < total; gi += blockDim.x) {
		const int c = gi >> (logn - K), g = gi - (c << (logn - K));
		const int lo = g & (sub - 1), hi = g >> s;
		double2 *col = data + c * stride;
		const int e0 = (hi << (s + K)) + lo;
		double2 x[R];
#pragma unroll
		for (int j = 0; j < R; ++j) x[j] = col[swz(e0 + (j << s))];
#pragma unroll
		for (int t = 1; t <= K; ++t) {
			const int h = 1 << (t - 1);
			const int S = s + t;
#pragma unroll
			for (int j = 0; j < R; ++j) {
				if (j & h) continue;
				const int kS = ((j & (h - 1)) << s) + lo;  // index within the stage's half block
				const double2 u = x[j], v = cmul(x[j + h], twiddle(wt, S, kS));
				x[j] = cadd(u, v);
				x[j + h] = csub(u, v);
			}
		}
#pragma unroll
		for (int j = 0; j < R; ++j) col[swz(e0 + (j << s))] = x[j];
	}
	__syncthreads();
}

__device__ inline void fft_dit(double2 *data, int logn, int ncol, int stride, const double2 *wt)
{
	int s = 0;
	for (; s + 3 <= logn; s += 3) dit_pass<3>(data, logn, ncol, stride, s, wt);
	if (logn - s == 2) dit_pass<2>(data, logn, ncol, stride, s, wt);
	else if (logn - s == 1) dit_pass<1>(data, logn, ncol, stride, s, wt);
}

// In-place inverse (conjugate twiddles, unscaled) DIF: natural-order input,
// bit-reversed output, element e at swz(e).  Stage S (from log2(n) down to 1)
// maps (u, v) at i and i + 2^(S-1) to (u + v, (u - v) conj w); passes of up
// to three stages as in the forward transform.
template <int K>
__device__ __forceinline__ void dif_pass(double2 *data, int logn, int ncol, int stride, int top, const double2 *wt)
{
	// stages top, top-1, ..., top-K+1 combine indices differing in bits
	// top-1 .. top-K, i.e. spacing 2^(top-K)
	constexpr int R = 1 << K;
	const int s = top - K;
	const int sub = 1 << s;
	const int total = ncol << (logn - K);
	for (int gi = threadIdx.x; gi < total; gi += blockDim.x) {
		const int c = gi >> (logn - K), g = gi - (c << (logn - K));
		const int lo = g & (sub - 1), hi = g >> s;
		double2 *col = data + c * stride;
		const int e0 = (hi << top) + lo;
		double2 x[R];
#pragma unroll
		for (int j = 0; j < R; ++j) x[j] = col[swz(e0 + (j << s))];
#pragma unroll
		for (int t = K; t >= 1; --t) {
			const int h = 1 << (t - 1);
			const int S = s + t;
#pragma unroll
			for (int j = 0; j < R; ++j) {
				if (j & h) continue;
				const int kS = ((j & (h - 1)) << s) + lo;
				const double2 u = x[j], v = x[j + h];
				x[j] = cadd(u, v);
				x[j + h] = cmul(csub(u, v), conjg(twiddle(wt, S, kS)));
			}
		}
#pragma unroll
		for (int j = 0; j < R; ++j) col[swz(e0 + (j << s))] = x[j];
	}
	__syncthreads();
}

__device__ inline void ifft_dif(double2 *data, int logn, int ncol, int stride, const double2 *wt)
{
	int top = logn;
	for (; top - 3 >= 0; top -= 3) dif_pass<3>(data, logn, ncol, stride, top, wt);
	if (top == 2) dif_pass<2>(data, logn, ncol, stride, top, wt);
	else if (top == 1) dif_pass<1>(data, logn, ncol, stride, top, wt);
}

}  // namespace

// Forward transforms of `np` pairs of real rows, pair p at rows 2p, 2p + 1 of
// `src` (row 2p + 1 only below nrows; pairs from pmax on are not loaded or
// stored), columns [x0, x0 + nx) of the frame holding data, into the
// half-spectrum rows 2p, 2p + 1 of `out` (pitch P):
// A = (Z[k] + conj Z[N-k]) / 2 and B = (Z[k] - conj Z[N-k]) / 2i, k = 0..N/2,
// for z = a + i b.  Sequence p sits at sh + p * N.
template <class T>
__device__ __forceinline__ void rows_fwd_pairs(double2 *sh, const double2 *wt, int N, int L, int np, const T *src,
                                               long ld, int x0, int nx, int nrows, int pmax, double2 *out, int P)
{
	const int M = N / 2 + 1;
	for (int t = threadIdx.x; t < np * N; t += blockDim.x) {
		const int p = t >> L, k = t - (p << L);
		double a = 0, bv = 0;
		const int xi = k - x0;
		if (p < pmax && xi >= 0 && xi < nx) {
			const T *row = src + (long)(2 * p) * ld;
			a = (double)row[xi];
			if (2 * p + 1 < nrows) bv = (double)row[ld + xi];
		}
		sh[(p << L) + swz(bitrev(k, L))] = make_double2(a, bv);
	}
	__syncthreads();
	fft_dit(sh, L, np, N, wt);
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
	double2 *wt = sh + np * N;
	stage_twiddles(wt, N, L, tw, twn);
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
	rows_fwd_pairs<float>(sh, wt, N, L, np, src, width, x0, width, nrows - r0, pmax, I.fbuf + (long)(y0 + r0) * P, P);
}

// Forward rows of a galaxy's real FP64 PSF frame (N1 x N2, `src64`) into the
// half-spectrum frame `dst64`, fft_pairs_per_block(N1) pairs of rows per block.
__global__ __launch_bounds__(256) void k_fft_psf_rows(const GalaxyDev *__restrict__ g, const double2 *__restrict__ tw,
                                                      int twn, const double *__restrict__ src64, double2 *dst64)
{
	extern __shared__ double2 sh[];
	const int N = g->fft_n1, L = g->fft_l1, P = fft_pitch(N), np = fft_pairs_per_block(N);
	double2 *wt = sh + np * N;
	stage_twiddles(wt, N, L, tw, twn);
	const int r0 = 2 * np * blockIdx.x;
	const int nrows = g->fft_n2 - r0;
	rows_fwd_pairs<double>(sh, wt, N, L, np, src64 + (long)r0 * N, N, 0, N, nrows, (nrows + 1) / 2,
	                       dst64 + (long)r0 * P, P);
}

// Column pass over CB adjacent half-spectrum columns of one instance.  Frame
// rows [in_lo, in_hi) hold row spectra, the others are zero.  op 0: forward
// only (all N2 rows of `dst64` written: the PSF spectrum); op 1: product with
// the PSF spectrum; op 2: with its conjugate.  The inverse column transform
// then leaves rows [out_lo, out_hi) in the frame.  Global accesses run along
// rows (CB adjacent entries); columns sit in LDS at a stride of N2 + 1.
__global__ __launch_bounds__(256) void k_fft_cols(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                  const int *__restrict__ col_off, int ninst, int cb, int op,
                                                  int in_mode, int out_mode, const double2 *__restrict__ tw, int twn,
                                                  double2 *dst64)
{
	extern __shared__ double2 sh[];
	const int b = blockIdx.x;
	const int inst = owner(col_off, ninst, b);
	const InstDev &I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int M = g.fft_n1 / 2 + 1, P = fft_pitch(g.fft_n1), N = g.fft_n2, L = g.fft_l2;
	const int S = N + 1;
	double2 *wt = sh + cb * S;
	stage_twiddles(wt, N, L, tw, twn);
	const int c0 = (b - col_off[inst]) * cb;
	double2 *frame = (op == 0) ? dst64 : I.fbuf;
	// rows holding data: 0 canvas rows [0, DH); 1 image rows [pady, pady + H); 2 all rows
	const int in_lo = in_mode == 1 ? g.pady : 0;
	const int in_hi = in_mode == 0 ? g.DH : (in_mode == 1 ? g.pady + g.H : N);
	for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
		const int y = t / cb, c = t - y * cb;
		double2 v = make_double2(0.0, 0.0);
		if (y >= in_lo && y < in_hi && c0 + c < M) v = frame[(long)y * P + c0 + c];
		sh[c * S + swz(bitrev(y, L))] = v;
	}
	__syncthreads();
	fft_dit(sh, L, cb, S, wt);
	if (op == 0) {
		for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
			const int y = t / cb, c = t - y * cb;
			if (c0 + c < M) frame[(long)y * P + c0 + c] = sh[c * S + swz(y)];
		}
		return;
	}
	for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
		const int y = t / cb, c = t - y * cb;
		if (c0 + c < M) {
			const double2 s = g.fft_spec[(long)y * P + c0 + c];
			const double2 v = sh[c * S + swz(y)];
			sh[c * S + swz(y)] = (op == 1) ? cmul(v, s) : cmulc(v, s);
		}
	}
	__syncthreads();
	ifft_dif(sh, L, cb, S, wt);
	// rows written back: 0 image rows [pady, pady + H); 1 canvas rows [0, DH)
	const int out_lo = out_mode == 0 ? g.pady : 0;
	const int out_hi = out_mode == 0 ? g.pady + g.H : g.DH;
	for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
		const int y = t / cb, c = t - y * cb;
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
//   mode 1 (model image): canvas rows; G = model masked by calc
//   mode 2 (adjoint): canvas rows; G = value / (N1 N2)
__global__ __launch_bounds__(256) void k_fft_rows_inv(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                      const int *__restrict__ row_off, int ninst, int mode,
                                                      const double2 *__restrict__ tw, int twn,
                                                      const int *__restrict__ lk_base, double *partial)
{
	extern __shared__ double2 sh[];
	const int b = blockIdx.x;
	const int inst = owner(row_off, ninst, b);
	const InstDev &I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int N = g.fft_n1, L = g.fft_l1, M = N / 2 + 1, P = fft_pitch(N), np = fft_pairs_per_block(N);
	double2 *wt = sh + np * N;
	stage_twiddles(wt, N, L, tw, twn);
	const int blk = b - row_off[inst];
	const int r0 = 2 * np * blk;  // first output row of the block (image rows in mode 0)
	const int nrows = mode == 0 ? g.H : g.DH;
	const int y0 = mode == 0 ? r0 + g.pady : r0;  // its frame row
	for (int t = threadIdx.x; t < np * N; t += blockDim.x) {
		const int p = t >> L, k = t - (p << L);
		const int r = r0 + 2 * p;
		double2 z = make_double2(0.0, 0.0);
		if (r < nrows) {
			const bool second = r + 1 < nrows;
			const double2 *ra = I.fbuf + (long)(y0 + 2 * p) * P;
			const double2 *rb = ra + P;
			if (k < M) {
				const double2 a = ra[k], bv = second ? rb[k] : make_double2(0.0, 0.0);
				z = make_double2(a.x - bv.y, a.y + bv.x);  // A + i B
			}
			else {
				const double2 a = conjg(ra[N - k]), bv = second ? conjg(rb[N - k]) : make_double2(0.0, 0.0);
				z = make_double2(a.x - bv.y, a.y + bv.x);
			}
		}
		sh[(p << L) + swz(k)] = z;
	}
	__syncthreads();
	ifft_dif(sh, L, np, N, wt);
	const double scale = 1.0 / ((double)g.fft_n1 * (double)g.fft_n2);
	if (mode == 0) {
		double part = 0;
		const bool pois = g.likefunc == LIKE_POIS, tl = g.likefunc == LIKE_T;
		const double s = g.pois_scale;
		const int W2 = 2 * g.W;
		for (int t = threadIdx.x; t < np * W2; t += blockDim.x) {
			const int p = t / W2, rem = t - p * W2;
			const int row = rem >= g.W ? 1 : 0, x = rem - row * g.W;
			const int r = r0 + 2 * p + row;
			if (r >= nrows) continue;
			const long o = (long)r * g.W + x;
			float wv = 0.f;
			if (g.region[o]) {
				const long qc = (long)(y0 + 2 * p + row) * g.DW + x + g.padx;
				const bool in_calc = !g.has_calc || g.calc[qc];
				const double2 z = sh[(p << L) + swz(bitrev(x + g.padx, L))];
				const double m = in_calc ? (row ? z.y : z.x) * scale + (I.has_aux ? (double)I.aux[qc] : 0.0) : 0.0;
				const double d = g.data[o];
				if (pois) {
					const double mm = s * m, dd = s * d;
					part += mm - dd * log(mm);
					wv = (float)(-s * (1.0 - d / m));
				}
				else {
					const double diff = d - m;
					const double iv = g.ivar[o];
					part += diff * diff * iv;
					wv = (float)(diff * iv);
					// t: the standardised residual, for k_like_t
					if (tl) wv = (float)(diff * sqrt(iv));
				}
				if (!in_calc && !tl) wv = 0.f;
			}
			I.w[o] = wv;
		}
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
	const int D2 = 2 * g.DW;
	for (int t = threadIdx.x; t < np * D2; t += blockDim.x) {
		const int p = t / D2, rem = t - p * D2;
		const int row = rem >= g.DW ? 1 : 0, x = rem - row * g.DW;
		const int r = r0 + 2 * p + row;
		if (r >= nrows) continue;
		const double2 z = sh[(p << L) + swz(bitrev(x, L))];
		double v = (row ? z.y : z.x) * scale;
		const long o = (long)r * g.DW + x;
		if (mode == 1 && g.has_calc && !g.calc[o]) v = 0;
		if (mode == 1 && I.has_aux) v += (double)I.aux[o];
		I.G[o] = (float)v;
	}
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
static size_t lds_rows_fwd[64], lds_psf_rows[64], lds_cols[64], lds_rows_inv[64];

static int cols_per_block(int n2) { return n2 >= 4096 ? 1 : (n2 >= 2048 ? 2 : (4096 / n2 > 8 ? 8 : 4096 / n2)); }

// LDS of a row pass: the block's pairs and the per-stage twiddles (their
// total grows with N, so the largest frame sets it)
static size_t row_lds(int nmax) { return sizeof(double2) * (size_t)(fft_pairs_per_block(nmax) + 1) * nmax; }

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

int launch_fft_cols(int nblocks, int n2max, int cb, const GalaxyDev *gals, const InstDev *insts, const int *col_off,
                    int ninst, int op, int in_mode, int out_mode, const double2 *tw, int twn, double2 *dst64,
                    hipStream_t st)
{
	if (nblocks <= 0) return 0;
	const size_t lds = sizeof(double2) * ((size_t)cb * (n2max + 1) + n2max);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
	if (int e = ensure_lds((const void *)k_fft_cols, lds, lds_cols)) return e;
	k_fft_cols<<<nblocks, 256, lds, st>>>(gals, insts, col_off, ninst, cb, op, in_mode, out_mode, tw, twn, dst64);
	return (int)hipGetLastError();
}

int launch_fft_rows_inv(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int *row_off,
                        int ninst, int mode, const double2 *tw, int twn, const int *lk_base, double *partial,
                        hipStream_t st)
{
	if (nblocks <= 0) return 0;
	const size_t lds = row_lds(nmax);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
	if (int e = ensure_lds((const void *)k_fft_rows_inv, lds, lds_rows_inv)) return e;
	k_fft_rows_inv<<<nblocks, 256, lds, st>>>(gals, insts, row_off, ninst, mode, tw, twn, lk_base, partial);
	return (int)hipGetLastError();
}

int fft_cols_per_block(int n2) { return cols_per_block(n2); }

}  // namespace profit_amd
