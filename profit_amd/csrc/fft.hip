// FFT convolution path (FFTConvolver, src/libprofit/src/profit/convolve.cpp:233-312).
//
// libprofit computes the 2-D linear convolution through a 1-D real FFT of the
// flattened, zero-extended image (FFTW).  Here the same linear convolution is
// computed with 2-D power-of-two complex FFTs in FP64 on the GPU, in three
// passes over an N1 x N2 frame per instance:
//   k_fft_rows_fwd   row FFTs of the zero-padded real input (canvas or w)
//   k_fft_cols_conv  column FFT, product with the PSF spectrum (its conjugate
//                    for the adjoint correlation), inverse column FFT
//   k_fft_rows_inv   inverse row FFTs with the epilogue of the brute-force
//                    path: fused likelihood, masked model image, or the
//                    adjoint image G
// Every transform runs in LDS (in-place radix-2 after a bit-reversed load).
// FP64 keeps the error relative to the image maximum near 1e-15, so faint
// pixels keep their relative accuracy.
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

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

__device__ inline double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
	return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// In-place iterative radix-2 transforms of `ncol` sequences of length n in
// LDS (sequence c at data + c*stride), input in bit-reversed order.
// tw[t] = exp(-2 pi i t / twn) for t < twn/2.  inverse: conjugate twiddles.
__device__ inline void lds_fft(double2 *data, int n, int logn, int ncol, int stride, const double2 *__restrict__ tw,
                               int twn, bool inverse)
{
	const int half_total = ncol * (n >> 1);
	for (int s = 1; s <= logn; ++s) {
		const int len = 1 << s, half = len >> 1;
		const int step = twn >> s;  // twn / len
		for (int b = threadIdx.x; b < half_total; b += blockDim.x) {
			const int c = b >> (logn - 1), bb = b - (c << (logn - 1));
			const int grp = bb >> (s - 1), k = bb - (grp << (s - 1));
			const int i = c * stride + grp * len + k, j = i + half;
			double2 w = tw[k * step];
			if (inverse) w.y = -w.y;
			const double2 u = data[i], v = cmul(data[j], w);
			data[i] = make_double2(u.x + v.x, u.y + v.y);
			data[j] = make_double2(u.x - v.x, u.y - v.y);
		}
		__syncthreads();
	}
}

}  // namespace

// mode 0: input = canvas (DW x DH) at frame (0,0), one block per canvas row.
// mode 1: input = w (W x H) at frame (padx, pady), one block per image row.
// mode 2: input = real FP64 frame (N1 x N2) `src64`, one block per frame row.
__global__ __launch_bounds__(256) void k_fft_rows_fwd(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                      const int *__restrict__ row_off, int ninst, int mode,
                                                      const double2 *__restrict__ tw, int twn,
                                                      const double *__restrict__ src64, double2 *dst64)
{
	extern __shared__ double2 sh[];
	const int b = blockIdx.x;
	const int inst = owner(row_off, ninst, b);
	const InstDev &I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int N = g.fft_n1, L = g.fft_l1;
	const int r = b - row_off[inst];
	int y, x0, nx;
	const float *src32 = nullptr;
	const double *srcd = nullptr;
	if (mode == 0) {
		y = r;
		x0 = 0;
		nx = g.DW;
		src32 = I.canvas + (long)r * g.DW;
	}
	else if (mode == 1) {
		y = r + g.pady;
		x0 = g.padx;
		nx = g.W;
		src32 = I.w + (long)r * g.W;
	}
	else {
		y = r;
		x0 = 0;
		nx = N;
		srcd = src64 + (long)r * N;
	}
	for (int t = threadIdx.x; t < N; t += blockDim.x) {
		double v = 0;
		const int xi = t - x0;
		if (xi >= 0 && xi < nx) v = src32 ? (double)src32[xi] : srcd[xi];
		sh[bitrev(t, L)] = make_double2(v, 0.0);
	}
	__syncthreads();
	lds_fft(sh, N, L, 1, N, tw, twn, false);
	double2 *out = (mode == 2 ? dst64 : I.fbuf) + (long)y * N;
	for (int t = threadIdx.x; t < N; t += blockDim.x) out[t] = sh[t];
}

// Column pass over CB adjacent columns of one instance.
// in rows [in_lo, in_hi) of the frame hold the row spectra, the others are
// zero.  op 0: forward only (writes all N2 rows of `dst64`, used to build the
// PSF spectrum); op 1: product with the spectrum; op 2: with its conjugate.
// After the inverse column FFT, rows [out_lo, out_hi) are written back.
// Global accesses run along rows (CB adjacent columns per row); the columns
// sit in LDS at a stride of N2 + 1 elements to spread them over the banks.
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
	const int N1 = g.fft_n1, N = g.fft_n2, L = g.fft_l2;
	const int S = N + 1;
	const int c0 = (b - col_off[inst]) * cb;
	double2 *frame = (op == 0) ? dst64 : I.fbuf;
	// rows holding data: 0 canvas rows [0, DH); 1 image rows [pady, pady + H); 2 all rows
	const int in_lo = in_mode == 1 ? g.pady : 0;
	const int in_hi = in_mode == 0 ? g.DH : (in_mode == 1 ? g.pady + g.H : N);
	for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
		const int y = t / cb, c = t - y * cb;
		double2 v = make_double2(0.0, 0.0);
		if (y >= in_lo && y < in_hi && c0 + c < N1) v = frame[(long)y * N1 + c0 + c];
		sh[c * S + bitrev(y, L)] = v;
	}
	__syncthreads();
	lds_fft(sh, N, L, cb, S, tw, twn, false);
	if (op == 0) {
		for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
			const int y = t / cb, c = t - y * cb;
			if (c0 + c < N1) frame[(long)y * N1 + c0 + c] = sh[c * S + y];
		}
		return;
	}
	// product with the PSF spectrum, written straight back in bit-reversed
	// order for the inverse transform
	double2 *tmp = sh + cb * S;
	for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
		const int y = t / cb, c = t - y * cb;
		double2 v = sh[c * S + y];
		if (c0 + c < N1) {
			const double2 s = g.fft_spec[(long)y * N1 + c0 + c];
			v = (op == 1) ? cmul(v, s) : cmulc(v, s);
		}
		tmp[c * S + bitrev(y, L)] = v;
	}
	__syncthreads();
	lds_fft(tmp, N, L, cb, S, tw, twn, true);
	// rows written back: 0 image rows [pady, pady + H); 1 canvas rows [0, DH)
	const int out_lo = out_mode == 0 ? g.pady : 0;
	const int out_hi = out_mode == 0 ? g.pady + g.H : g.DH;
	for (int t = threadIdx.x; t < cb * N; t += blockDim.x) {
		const int y = t / cb, c = t - y * cb;
		if (y >= out_lo && y < out_hi && c0 + c < N1) frame[(long)y * N1 + c0 + c] = tmp[c * S + y];
	}
}

// Inverse row pass with the epilogues.
// mode 0 (likelihood): one block per image row; model = Re / (N1 N2) at frame
//   (padx + x, pady + y), likelihood terms into partial[b], w written.
// mode 1 (model image): one block per canvas row; G = model masked by calc.
// mode 2 (adjoint): one block per canvas row; G = Re / (N1 N2).
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
	const int N = g.fft_n1, L = g.fft_l1;
	const int r = b - row_off[inst];
	const int y = mode == 0 ? r + g.pady : r;
	const double2 *row = I.fbuf + (long)y * N;
	for (int t = threadIdx.x; t < N; t += blockDim.x) sh[bitrev(t, L)] = row[t];
	__syncthreads();
	lds_fft(sh, N, L, 1, N, tw, twn, true);
	const double scale = 1.0 / ((double)g.fft_n1 * (double)g.fft_n2);
	if (mode == 0) {
		double part = 0;
		const bool pois = g.likefunc == LIKE_POIS;
		const double s = g.pois_scale;
		for (int x = threadIdx.x; x < g.W; x += blockDim.x) {
			const long o = (long)r * g.W + x;
			float wv = 0.f;
			if (g.region[o]) {
				const bool in_calc = !g.has_calc || g.calc[(long)y * g.DW + x + g.padx];
				const double m = in_calc ? sh[x + g.padx].x * scale : 0.0;
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
				}
				if (!in_calc) wv = 0.f;
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
			partial[lk_base[inst] + r] = t;
		}
		return;
	}
	float *out = I.G + (long)y * g.DW;
	for (int x = threadIdx.x; x < g.DW; x += blockDim.x) {
		double v = sh[x].x * scale;
		if (mode == 1 && g.has_calc && !g.calc[(long)y * g.DW + x]) v = 0;
		out[x] = (float)v;
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
static size_t lds_rows_fwd[64], lds_cols[64], lds_rows_inv[64];

static int cols_per_block(int n2) { return n2 >= 2048 ? 1 : (2048 / n2 > 8 ? 8 : 2048 / n2); }

int launch_fft_rows_fwd(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int *row_off,
                        int ninst, int mode, const double2 *tw, int twn, const double *src64, double2 *dst64,
                        hipStream_t st)
{
	if (nblocks <= 0) return 0;
	const size_t lds = sizeof(double2) * (size_t)nmax;
	if (int e = ensure_lds((const void *)k_fft_rows_fwd, lds, lds_rows_fwd)) return e;
	k_fft_rows_fwd<<<nblocks, 256, lds, st>>>(gals, insts, row_off, ninst, mode, tw, twn, src64, dst64);
	return (int)hipGetLastError();
}

int launch_fft_cols(int nblocks, int n2max, int cb, const GalaxyDev *gals, const InstDev *insts, const int *col_off,
                    int ninst, int op, int in_mode, int out_mode, const double2 *tw, int twn, double2 *dst64,
                    hipStream_t st)
{
	if (nblocks <= 0) return 0;
	const size_t lds = sizeof(double2) * (size_t)cb * (n2max + 1) * (op == 0 ? 1 : 2);
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
	const size_t lds = sizeof(double2) * (size_t)nmax;
	if (int e = ensure_lds((const void *)k_fft_rows_inv, lds, lds_rows_inv)) return e;
	k_fft_rows_inv<<<nblocks, 256, lds, st>>>(gals, insts, row_off, ninst, mode, tw, twn, lk_base, partial);
	return (int)hipGetLastError();
}

int fft_cols_per_block(int n2) { return cols_per_block(n2); }

}  // namespace profit_amd
