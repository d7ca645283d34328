// Runs the FFT convolution kernels on the CPU (hip_cpu stand-in) for a set of
// geometries with every device buffer at its exact size, so that
// AddressSanitizer reports any access outside one.  The host-side numbers
// (frame sizes, blocks per pass, LDS sizes) are computed the way the engine
// computes them; the dynamic LDS past the size the launcher asks for is
// poisoned, so the kernels' LDS indexing is checked too.  Prints
// "ok <geometry>" per case.
#include <sanitizer/asan_interface.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "fft_kernels.inc"  // fft.hip up to its launchers

namespace profit_amd {
constexpr size_t kLdsMax = 1 << 17;
double2 sh[kLdsMax];  // the kernels' dynamic LDS (blocks run one at a time)
}

// the launcher's dynamic LDS size in double2 entries: accesses beyond it fault
static void set_lds(size_t n)
{
	ASAN_UNPOISON_MEMORY_REGION(profit_amd::sh, sizeof(profit_amd::sh));
	if (n > profit_amd::kLdsMax) std::abort();
	ASAN_POISON_MEMORY_REGION(profit_amd::sh + n, (profit_amd::kLdsMax - n) * sizeof(double2));
}

using namespace profit_amd;

static int next_pow2(int v)
{
	int n = 1;
	while (n < v) n <<= 1;
	return n;
}
static int ilog2(int n)
{
	int l = 0;
	while ((1 << l) < n) ++l;
	return l;
}
template <class T> static T *dev(size_t n)
{
	T *p = (T *)std::malloc(n * sizeof(T));
	std::memset(p, 0, n * sizeof(T));
	return p;
}

static void run_case(int W, int H, int K, int pad)
{
	GalaxyDev g{};
	g.W = W;
	g.H = H;
	g.padx = g.pady = pad;
	g.DW = W + 2 * pad;
	g.DH = H + 2 * pad;
	g.Kx = g.Ky = K;
	g.has_calc = 1;
	g.likefunc = LIKE_CHISQ;
	g.pois_scale = 1;
	const int n1 = std::max(8, next_pow2(g.DW + K)), n2 = std::max(8, next_pow2(g.DH + K)), m1 = n1 / 2 + 1;
	g.fft_n1 = n1;
	g.fft_n2 = n2;
	g.fft_l1 = ilog2(n1);
	g.fft_l2 = ilog2(n2);
	const int twn = n1 > n2 ? n1 : n2;
	double2 *tw = dev<double2>(twn / 2);
	for (int t = 0; t < twn / 2; ++t) tw[t] = make_double2(std::cos(2 * M_PI * t / twn), -std::sin(2 * M_PI * t / twn));
	double2 *spec = dev<double2>((size_t)fft_pitch(n1) * n2);
	double *src = dev<double>((size_t)n1 * n2);
	src[0] = 1;
	g.fft_spec = spec;
	uint8_t *calc = dev<uint8_t>((size_t)g.DW * g.DH), *region = dev<uint8_t>((size_t)W * H);
	std::memset(calc, 1, (size_t)g.DW * g.DH);
	std::memset(region, 1, (size_t)W * H);
	double *data = dev<double>((size_t)W * H);
	float *ivar = dev<float>((size_t)W * H);
	for (int i = 0; i < W * H; ++i) ivar[i] = 1;
	g.calc = calc;
	g.region = region;
	g.data = data;
	g.ivar = ivar;
	PixelIn *pix = dev<PixelIn>((size_t)W * H);
	for (int y = 0; y < H; ++y)
		for (int x = 0; x < W; ++x)
			pix[y * W + x] = PixelIn{data[y * W + x], ivar[y * W + x],
			                         (uint32_t)(region[y * W + x] | (calc[(y + pad) * g.DW + x + pad] << 1))};
	g.pix = pix;
	GalaxyDev *gals = dev<GalaxyDev>(1);
	gals[0] = g;
	InstDev *insts = dev<InstDev>(1);
	insts[0].galaxy = 0;
	insts[0].canvas = dev<float>((size_t)g.DW * g.DH);
	insts[0].w = dev<float>((size_t)W * H);
	insts[0].G = dev<float>((size_t)g.DW * g.DH);
	insts[0].fbuf = dev<double2>((size_t)fft_pitch(n1) * n2);
	const int cb = n2 >= 4096 ? 1 : (n2 >= 2048 ? 2 : (4096 / n2 > 8 ? 8 : 4096 / n2));
	int *off = dev<int>(2);
	const int np = fft_pairs_per_block(n1), rpb = 2 * np;
	auto row_blocks = [rpb](int rows) { return (rows + rpb - 1) / rpb; };
	const size_t row_lds = (size_t)np * n1 + n1 / 2, col_lds = (size_t)cb * (n2 + 2) + n2 / 2;
	int *lk = dev<int>(2);
	double *partial = dev<double>(row_blocks(H));
	lk[1] = row_blocks(H);
	auto rows_fwd = [&](int mode, int nblocks) {
		off[1] = nblocks;
		set_lds(row_lds);
		hip_cpu::launch(nblocks, 256, [&] { k_fft_rows_fwd(gals, insts, off, 1, mode, tw, twn); });
	};
	// block tables (instance 0, block b), at their exact sizes
	auto table = [&](int nb) {
		int2 *t = dev<int2>((size_t)nb);
		for (int b = 0; b < nb; ++b) t[b] = make_int2(0, b);
		return t;
	};
	auto cols = [&](int op, int in_mode, int out_mode, double2 *d64) {
		const int nb = (m1 + cb - 1) / cb;
		const int2 *blk = table(nb);
		set_lds(col_lds);
		hip_cpu::launch(nb, kColThreads, [&] { k_fft_cols(gals, insts, blk, cb, op, in_mode, out_mode, tw, twn, d64); });
	};
	auto rows_inv = [&](int mode, int nblocks) {
		const int2 *blk = table(nblocks);
		set_lds(row_lds);
		hip_cpu::launch(nblocks, 256, [&] {
			if (mode == 0) k_fft_rows_inv<0>(gals, insts, blk, tw, twn, lk, partial);
			else if (mode == 1) k_fft_rows_inv<1>(gals, insts, blk, tw, twn, lk, partial);
			else if (mode == 2) k_fft_rows_inv<2>(gals, insts, blk, tw, twn, lk, partial);
			else k_fft_rows_inv<3>(gals, insts, blk, tw, twn, lk, partial);
		});
	};
	set_lds(row_lds);
	hip_cpu::launch((n2 / 2 + np - 1) / np, 256, [&] { k_fft_psf_rows(gals, tw, twn, src, spec); });
	cols(0, 2, 0, spec);
	// the PSF frame is a unit impulse at the origin, so every convolution is
	// the identity: model = canvas, adjoint = w placed at (padx, pady)
	unsigned st = 12345u;
	auto rnd = [&] {
		st = st * 1664525u + 1013904223u;
		return (float)((st >> 8) & 0xffff) / 65536.f;
	};
	for (int i = 0; i < g.DW * g.DH; ++i) insts[0].canvas[i] = rnd();
	for (int i = 0; i < W * H; ++i) insts[0].w[i] = rnd();
	double err = 0;
	rows_fwd(0, row_blocks(g.DH));
	cols(1, 0, 1, nullptr);
	rows_inv(1, row_blocks(g.DH));
	for (int i = 0; i < g.DW * g.DH; ++i) err = std::fmax(err, std::fabs(insts[0].G[i] - insts[0].canvas[i]));
	// the likelihood epilogue (chi^2 with data 0 and unit weights: w = -model),
	// generic (mode 0) and Gaussian-only (mode 3)
	for (int mode : {0, 3}) {
		std::memset(insts[0].w, 0, sizeof(float) * (size_t)W * H);
		rows_fwd(0, row_blocks(g.DH));
		cols(1, 0, 0, nullptr);
		rows_inv(mode, row_blocks(H));
		for (int y = 0; y < H; ++y)
			for (int x = 0; x < W; ++x)
				err = std::fmax(err, std::fabs(insts[0].w[y * W + x] + insts[0].canvas[(y + pad) * g.DW + x + pad]));
	}
	for (int i = 0; i < W * H; ++i) insts[0].w[i] = rnd();
	rows_fwd(1, row_blocks(H));
	cols(2, 1, 1, nullptr);
	rows_inv(2, row_blocks(g.DH));
	for (int y = 0; y < g.DH; ++y)
		for (int x = 0; x < g.DW; ++x) {
			const bool in = y >= pad && y < pad + H && x >= pad && x < pad + W;
			const float ref = in ? insts[0].w[(y - pad) * W + x - pad] : 0.f;
			err = std::fmax(err, std::fabs(insts[0].G[y * g.DW + x] - ref));
		}
	// a three-tap PSF, one tap at a negative (wrapped) offset, so that the
	// forward and inverse transforms are checked separately: model(x, y) =
	// c(x, y) + 0.5 c(x - 1, y - 2) + 0.25 c(x + 3, y + 1), and the adjoint
	// correlation G(x, y) = v(x, y) + 0.5 v(x + 1, y + 2) + 0.25 v(x - 3,
	// y - 1) of w placed at (padx, pady)
	if (n1 >= 8 && n2 >= 8) {
		std::memset(src, 0, sizeof(double) * (size_t)n1 * n2);
		src[0] = 1;
		src[2 * n1 + 1] = 0.5;
		src[(size_t)(n2 - 1) * n1 + (n1 - 3)] = 0.25;
		set_lds(row_lds);
		hip_cpu::launch((n2 / 2 + np - 1) / np, 256, [&] { k_fft_psf_rows(gals, tw, twn, src, spec); });
		cols(0, 2, 0, spec);
		for (int i = 0; i < g.DW * g.DH; ++i) insts[0].canvas[i] = rnd();
		auto cv = [&](int x, int y) {
			return x >= 0 && x < g.DW && y >= 0 && y < g.DH ? (double)insts[0].canvas[y * g.DW + x] : 0.0;
		};
		rows_fwd(0, row_blocks(g.DH));
		cols(1, 0, 1, nullptr);
		rows_inv(1, row_blocks(g.DH));
		for (int y = 0; y < g.DH; ++y)
			for (int x = 0; x < g.DW; ++x) {
				const double ref = cv(x, y) + 0.5 * cv(x - 1, y - 2) + 0.25 * cv(x + 3, y + 1);
				err = std::fmax(err, std::fabs(insts[0].G[y * g.DW + x] - ref));
			}
		for (int i = 0; i < W * H; ++i) insts[0].w[i] = rnd();
		auto wv = [&](int x, int y) {
			x -= pad;
			y -= pad;
			return x >= 0 && x < W && y >= 0 && y < H ? (double)insts[0].w[y * W + x] : 0.0;
		};
		rows_fwd(1, row_blocks(H));
		cols(2, 1, 1, nullptr);
		rows_inv(2, row_blocks(g.DH));
		for (int y = 0; y < g.DH; ++y)
			for (int x = 0; x < g.DW; ++x) {
				const double ref = wv(x, y) + 0.5 * wv(x + 1, y + 2) + 0.25 * wv(x - 3, y - 1);
				err = std::fmax(err, std::fabs(insts[0].G[y * g.DW + x] - ref));
			}
	}
	std::printf("%s W=%d H=%d K=%d pad=%d frame=%dx%d max err %.3e\n", err < 1e-5 ? "ok" : "BAD", W, H, K, pad, n1,
	            n2, err);
}

int main()
{
	run_case(70, 90, 5, 0);
	run_case(33, 130, 9, 0);
	run_case(64, 40, 25, 12);
	run_case(31, 17, 7, 3);
	// the smallest frames: 8 and 16 points per axis
	run_case(3, 2, 1, 0);
	run_case(9, 4, 3, 1);
	// frames of 2048 and 4096 columns: two pairs, one pair of rows per block
	run_case(1100, 11, 5, 0);
	run_case(2100, 5, 3, 1);
	// column transforms of 64, 512 and 1024 points (radix-8 stages on the
	// frame's loads and stores, with 0, 3 and 4 stages between them)
	run_case(10, 40, 3, 0);
	run_case(20, 300, 5, 2);
	run_case(9, 700, 3, 0);
	set_lds(profit_amd::kLdsMax);
	return 0;
}
