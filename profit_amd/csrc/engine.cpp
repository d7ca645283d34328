// Host engine for the GPU model-image path: galaxy upload, tile lists,
// batched launch sequences and their hipGraph replay.
#include "engine.h"

#include <cstddef>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>

namespace profit_amd {

void hip_check(hipError_t e, const char *what)
{
	if (e != hipSuccess) {
		std::ostringstream os;
		os << what << ": " << hipGetErrorString(e) << " (" << (int)e << ")";
		throw hip_error(os.str());
	}
}

void DevBuf::alloc(size_t bytes)
{
	release();
	if (bytes == 0) return;
	hip_check(hipMalloc(&p_, bytes), "hipMalloc");
	hip_check(hipMemset(p_, 0, bytes), "hipMemset");
	n_ = bytes;
}

void DevBuf::release()
{
	if (p_) (void)hipFree(p_);
	p_ = nullptr;
	n_ = 0;
}

void HostBuf::alloc(size_t bytes)
{
	release();
	if (bytes == 0) return;
	hip_check(hipHostMalloc(&p_, bytes, hipHostMallocDefault), "hipHostMalloc");
	std::memset(p_, 0, bytes);
	n_ = bytes;
}

void HostBuf::release()
{
	if (p_) (void)hipHostFree(p_);
	p_ = nullptr;
	n_ = 0;
}

Engine::Engine(int device) : device_(device)
{
	hip_check(hipSetDevice(device_), "hipSetDevice");
	hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
}

Engine::~Engine()
{
	for (auto &g : graph_exec_)
		if (g) (void)hipGraphExecDestroy(g);
	for (auto &e : events_)
		if (e) (void)hipEventDestroy(e);
	for (auto &row : fft_events_)
		for (auto &e : row)
			if (e) (void)hipEventDestroy(e);
	if (stream_) (void)hipStreamDestroy(stream_);
}

static int next_pow2(int v)
{
	int n = 1;
	while (n < v) n <<= 1;
	return n;
}

// FFT frame length for v points: a power of two, at least 8 (the row passes
// fold the bit reversal into their first and last three radix-2 stages)
static int fft_frame(int v) { return std::max(8, next_pow2(v)); }

static int ilog2(int n)
{
	int l = 0;
	while ((1 << l) < n) ++l;
	return l;
}

void Engine::set_timing(bool on)
{
	timing_ = on;
	if (on && !events_[0]) {
		for (auto &e : events_) hip_check(hipEventCreate(&e), "hipEventCreate");
		for (auto &row : fft_events_)
			for (auto &e : row) hip_check(hipEventCreate(&e), "hipEventCreate");
	}
}

void Engine::mark(int k)
{
	if (timing_) hip_check(hipEventRecord(events_[k], stream_), "hipEventRecord");
}

void Engine::mark_fft(int stage, int k, hipStream_t st)
{
	if (!timing_) return;
	hip_check(hipEventRecord(fft_events_[stage][k], st), "hipEventRecord");
	if (k == 3) fft_marked_[stage] = true;
}

double Engine::conv_flops_forward() const
{
	double f = 0;
	for (int i = 0; i < ninst(); ++i) {
		const GalaxyDev &g = gal_host_[inst_galaxy_[i]];
		f += 2.0 * g.Kx * g.Ky * g.n_region;
	}
	return f;
}

double Engine::conv_flops_adjoint() const
{
	double f = 0;
	for (int i = 0; i < ninst(); ++i) {
		const GalaxyDev &g = gal_host_[inst_galaxy_[i]];
		f += 2.0 * g.Kx * g.Ky * g.n_calc;
	}
	return f;
}

static void invalidate_graphs(hipGraphExec_t *g)
{
	for (int k = 0; k < 2; ++k)
		if (g[k]) {
			(void)hipGraphExecDestroy(g[k]);
			g[k] = nullptr;
		}
}

void Engine::set_conv_rows(int r)
{
	if (!(r == 2 || r == 4 || r == 8 || r == 16)) throw std::invalid_argument("conv rows must be 2, 4, 8 or 16");
	if (r == conv_rows_) return;
	conv_rows_ = r;
	for (size_t g = 0; g < gal_host_.size(); ++g) build_tiles((int)g);
	gals_dirty_ = true;
	invalidate_graphs(graph_exec_);
	if (!inst_galaxy_.empty()) set_batch(inst_galaxy_, inst_ncomp_);
}

void Engine::set_conv_method(int m)
{
	if (m < CK_VALU || m > CK_MFMA)
		throw std::invalid_argument("convolution method must be 0 (brute FP32), 1 (fft) or 2 (brute split-BF16x3)");
	bool overridden = false;
	for (int &k : kind_override_) {
		overridden |= k >= 0;
		k = -1;
	}
	if (m == conv_method_ && !overridden) return;
	conv_method_ = m;
	invalidate_graphs(graph_exec_);
	if (!inst_galaxy_.empty()) set_batch(inst_galaxy_, inst_ncomp_);
}

// forward + adjoint convolution time (ms) per evaluation of the current batch
double Engine::time_conv(const ProfileParams *params, int reps)
{
	const bool was = timing_;
	std::vector<double> ll(ninst()), grad((size_t)ncomp_total_ * NGRAD);
	evaluate(params, ll.data(), grad.data(), false);  // warm: attributes, queue size
	set_timing(true);
	reset_stats();
	for (int r = 0; r < reps; ++r) evaluate(params, ll.data(), grad.data(), false);
	const double ms = (stats_.ms[ST_CONV] + stats_.ms[ST_ADJ_CONV]) / std::max(1, reps);
	timing_ = was;
	reset_stats();
	return ms;
}

// Autotuner: picks each galaxy's convolution kernel (and the brute-force
// rows per thread) for the current batch.  Every candidate runs on the whole
// batch; its measured forward + adjoint convolution time divided by the
// batch's work in that kernel's own unit (brute force: tile pixels x PSF
// taps; matrix cores: 128 x 128 tiles x PSF rows x K chunks; FFT: frame
// points x log2 of the frame) calibrates a cost per unit on this device.
// Each galaxy then takes the kernel of least predicted cost; the mixed
// configuration is timed and kept only when it beats the best uniform one.
Engine::TuneReport Engine::autotune(const ProfileParams *params, int reps, bool allow_bf16x3)
{
	if (ninst() == 0) throw std::invalid_argument("autotune needs a batch");
	reps = std::max(1, reps);
	TuneReport rep;
	if (canvas64_) {
		// an FP64 canvas has one convolver, the FP64 FFT
		rep.ms[3] = rep.ms[5] = time_conv(params, reps);
		rep.kinds.assign(gal_host_.size(), CK_VALU);
		for (size_t g = 0; g < gal_host_.size(); ++g)
			if (gal_host_[g].has_psf) rep.kinds[g] = CK_FFT;
		return rep;
	}
	const int rows_cand[3] = {4, 8, 16};
	std::vector<double> w_valu[3], w_fft, w_mfma;
	const int ng = (int)gal_host_.size();
	auto per_galaxy_work = [&](int kind, std::vector<double> &w) {
		w.assign(ng, 0.0);
		for (int i = 0; i < ninst(); ++i) {
			const GalaxyDev &g = gal_host_[inst_galaxy_[i]];
			if (!g.has_psf) continue;
			double v = 0;
			if (kind == CK_VALU) v = (double)(g.n_ctiles + g.n_atiles) * 256.0 * conv_rows_ * g.Kx * g.Ky;
			else if (kind == CK_MFMA) v = (double)(g.n_mtl + g.n_mtc) * g.Ky * (g.mfma_kc + 2);
			else {
				const double n1 = fft_frame(g.DW + g.Kx), n2 = fft_frame(g.DH + g.Ky);
				v = n1 * n2 * std::log2(n1 * n2);
			}
			w[inst_galaxy_[i]] += v;
		}
	};
	double best_uniform = 1e300;
	int best_kind = CK_VALU, best_rows = conv_rows_;
	double a_valu = 0;
	for (int k = 0; k < 3; ++k) {
		set_conv_method(CK_VALU);
		set_conv_rows(rows_cand[k]);
		rep.ms[k] = time_conv(params, reps);
		per_galaxy_work(CK_VALU, w_valu[k]);
		if (rep.ms[k] < best_uniform) {
			best_uniform = rep.ms[k];
			best_kind = CK_VALU;
			best_rows = rows_cand[k];
		}
	}
	// brute force at the best rows per thread
	{
		int kbest = 0;
		for (int k = 1; k < 3; ++k)
			if (rep.ms[k] < rep.ms[kbest]) kbest = k;
		rep.rows = rows_cand[kbest];
		double tot = 0;
		for (double v : w_valu[kbest]) tot += v;
		a_valu = tot > 0 ? rep.ms[kbest] / tot : 0;
		w_valu[0] = w_valu[kbest];
	}
	set_conv_rows(rep.rows);
	set_conv_method(CK_FFT);
	rep.ms[3] = time_conv(params, reps);
	per_galaxy_work(CK_FFT, w_fft);
	if (rep.ms[3] < best_uniform) {
		best_uniform = rep.ms[3];
		best_kind = CK_FFT;
	}
	w_mfma.assign(ng, 0.0);
	if (allow_bf16x3) {
		set_conv_method(CK_MFMA);
		rep.ms[4] = time_conv(params, reps);
		per_galaxy_work(CK_MFMA, w_mfma);
		if (rep.ms[4] < best_uniform) {
			best_uniform = rep.ms[4];
			best_kind = CK_MFMA;
		}
	}
	auto total = [](const std::vector<double> &w) {
		double t = 0;
		for (double v : w) t += v;
		return t;
	};
	const double a_fft = total(w_fft) > 0 ? rep.ms[3] / total(w_fft) : 0;
	const double a_mfma = total(w_mfma) > 0 ? rep.ms[4] / total(w_mfma) : 0;
	rep.kinds.assign(ng, CK_VALU);
	for (int g = 0; g < ng; ++g) {
		if (!gal_host_[g].has_psf) continue;
		const double c[3] = {a_valu * w_valu[0][g], a_fft * w_fft[g], a_mfma * w_mfma[g]};
		int k = 0;
		for (int j = 1; j < (allow_bf16x3 ? 3 : 2); ++j)
			if (c[j] < c[k]) k = j;
		rep.kinds[g] = k;
	}
	// the mixed configuration
	conv_method_ = best_kind;
	kind_override_ = rep.kinds;
	set_batch(inst_galaxy_, inst_ncomp_);
	rep.ms[5] = time_conv(params, reps);
	if (rep.ms[5] > best_uniform) {
		set_conv_rows(best_kind == CK_VALU ? best_rows : rep.rows);
		set_conv_method(best_kind);
		for (int g = 0; g < ng; ++g)
			if (gal_host_[g].has_psf) rep.kinds[g] = best_kind;
		rep.ms[5] = best_uniform;
	}
	return rep;
}


// PSF spectra for the FFT path: g(d) = psf[K-1-K/2+d] about the origin (the
// brute-force kernel's weights as a function of the offset), wrapped into an
// N1 x N2 frame and transformed with the same FFT kernels.
void Engine::ensure_fft_spectra()
{
	int twn = 2;
	for (const auto &g : gal_host_)
		if (g.conv_kind == CK_FFT) {
			const int n1 = fft_frame(g.DW + g.Kx), n2 = fft_frame(g.DH + g.Ky);
			twn = std::max(twn, std::max(n1, n2));
		}
	if (twn > 4096) throw std::invalid_argument("FFT frame above 4096 points per axis is not supported");
	if (twn != twn_) {
		std::vector<double2> tw(twn / 2);
		const long double pi = 3.141592653589793238462643383279502884L;
		for (int t = 0; t < twn / 2; ++t) {
			const long double a = 2 * pi * (long double)t / (long double)twn;
			tw[t] = make_double2((double)cosl(a), (double)(-sinl(a)));
		}
		d_tw_.alloc(tw.size() * sizeof(double2));
		hip_check(hipMemcpy(d_tw_.as<double2>(), tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice), "tw");
		twn_ = twn;
		for (auto &b : gal_bufs_) b.fft_spec.release();
	}
	for (size_t gi = 0; gi < gal_host_.size(); ++gi) {
		GalaxyDev &g = gal_host_[gi];
		GalaxyBuffers &b = gal_bufs_[gi];
		if (g.conv_kind != CK_FFT || b.fft_spec.size()) continue;
		const int n1 = fft_frame(g.DW + g.Kx), n2 = fft_frame(g.DH + g.Ky);
		g.fft_n1 = n1;
		g.fft_n2 = n2;
		g.fft_l1 = ilog2(n1);
		g.fft_l2 = ilog2(n2);
		const int m1 = n1 / 2 + 1;  // half-spectrum width
		std::vector<double> frame((size_t)n1 * n2, 0.0);
		const int ax = g.Kx - 1 - g.Kx / 2, ay = g.Ky - 1 - g.Ky / 2;
		for (int ky = 0; ky < g.Ky; ++ky)
			for (int kx = 0; kx < g.Kx; ++kx) {
				const int dx = kx - ax, dy = ky - ay;
				frame[(size_t)((dy + n2) % n2) * n1 + (dx + n1) % n1] = b.psf64[(size_t)ky * g.Kx + kx];
			}
		b.fft_spec.alloc((size_t)fft_pitch(n1) * n2 * sizeof(double2));
		DevBuf src(frame.size() * sizeof(double));
		hip_check(hipMemcpy(src.as<double>(), frame.data(), frame.size() * sizeof(double), hipMemcpyHostToDevice), "psf");
		g.fft_spec = b.fft_spec.as<double2>();
		// one-instance context pointing at this galaxy
		DevBuf dg(sizeof(GalaxyDev)), di(sizeof(InstDev));
		InstDev I{};
		I.galaxy = 0;
		hip_check(hipMemcpy(dg.as<GalaxyDev>(), &g, sizeof(g), hipMemcpyHostToDevice), "gal");
		hip_check(hipMemcpy(di.as<InstDev>(), &I, sizeof(I), hipMemcpyHostToDevice), "inst");
		const int cb = fft_cols_per_block(n2);
		const int nblk = (m1 + cb - 1) / cb;
		std::vector<int2> blk(nblk);
		for (int t = 0; t < nblk; ++t) blk[t] = make_int2(0, t);
		DevBuf dblk(nblk * sizeof(int2));
		hip_check(hipMemcpy(dblk.as<int2>(), blk.data(), nblk * sizeof(int2), hipMemcpyHostToDevice), "blocks");
		hip_check(launch_fft_psf_rows(n1, n2, dg.as<GalaxyDev>(), d_tw_.as<double2>(), twn_, src.as<double>(),
		                              b.fft_spec.as<double2>(), stream_),
		          "psf rows");
		hip_check(launch_fft_cols(nblk, n2, cb, dg.as<GalaxyDev>(), di.as<InstDev>(), dblk.as<int2>(), 0, 2, 0,
		                          d_tw_.as<double2>(), twn_, b.fft_spec.as<double2>(), stream_),
		          "psf cols");
		hip_check(hipStreamSynchronize(stream_), "sync");
		gals_dirty_ = true;
	}
}

// forward convolution of every instance's canvas: mode 0 likelihood epilogue,
// mode 1 model image into G (masked by the calc region).  Each galaxy's
// instances go through its kernel: tiled FP32 brute force, the three FFT
// passes, or the split-BF16x3 matrix-core brute force; likelihood partials
// land at the instance's base in the shared partials array.
void Engine::conv_forward(int mode, hipStream_t st)
{
	const int n = ninst();
	const GalaxyDev *gals = d_gals_.as<GalaxyDev>();
	const InstDev *insts = d_insts_.as<InstDev>();
	double *partial = d_partial_.as<double>();
	const int *lk = d_lk_off_.as<int>();
	if (mode == 0) {
		hip_check(launch_conv(CONV_LIKE, conv_rows_, n_ct_blocks_, kx_max_, kc_max_, gals, insts, d_ct_off_.as<int>(),
		                      n, lk, partial, st),
		          "conv");
		hip_check(launch_conv_mfma(CONV_LIKE, ml_off_[n], mfma_pitch_, mfma_kch_, gals, insts, d_ml_off_.as<int>(), n,
		                           lk, partial, st),
		          "conv mfma");
	}
	else {
		hip_check(launch_conv(CONV_MODEL, conv_rows_, n_bt_blocks_, kx_max_, kc_max_, gals, insts, d_bt_off_.as<int>(),
		                      n, lk, partial, st),
		          "conv");
		hip_check(launch_conv_mfma(CONV_MODEL, mc_off_[n], mfma_pitch_, mfma_kch_, gals, insts, d_mc_off_.as<int>(), n,
		                           lk, partial, st),
		          "conv mfma");
	}
	if (rf0_off_[n] == 0) return;
	const bool timed = mode == 0;  // the likelihood evaluation's stage
	if (timed) mark_fft(0, 0, st);
	if (canvas64_)
		hip_check(launch_fft_rows_fwd64(rf0_off_[n], fft_nmax1_, gals, insts, d_rf0_off_.as<int>(), n,
		                                d_tw_.as<double2>(), twn_, st),
		          "fft rows");
	else
		hip_check(launch_fft_rows_fwd(rf0_off_[n], fft_nmax1_, gals, insts, d_rf0_off_.as<int>(), n, 0,
		                              d_tw_.as<double2>(), twn_, st),
		          "fft rows");
	if (timed) mark_fft(0, 1, st);
	hip_check(launch_fft_cols(col_off_[n], fft_nmax2_, fft_cb_, gals, insts, d_col_blk_.as<int2>(), 1, 0,
	                          mode == 0 ? 0 : 1, d_tw_.as<double2>(), twn_, nullptr, st),
	          "fft cols");
	if (timed) mark_fft(0, 2, st);
	if (mode == 0)
		hip_check(launch_fft_rows_inv(ri0_off_[n], fft_nmax1_, gals, insts, d_ri0_blk_.as<int2>(), all_gauss_ ? 3 : 0,
		                              d_tw_.as<double2>(), twn_, lk, partial, st),
		          "fft rows inv");
	else
		hip_check(launch_fft_rows_inv(ri1_off_[n], fft_nmax1_, gals, insts, d_ri1_blk_.as<int2>(), 1,
		                              d_tw_.as<double2>(), twn_, lk, partial, st),
		          "fft rows inv");
	if (timed) mark_fft(0, 3, st);
}

void Engine::conv_adjoint(hipStream_t st)
{
	const int n = ninst();
	const GalaxyDev *gals = d_gals_.as<GalaxyDev>();
	const InstDev *insts = d_insts_.as<InstDev>();
	const int *lk = d_lk_off_.as<int>();
	hip_check(launch_conv(CONV_ADJ, conv_rows_, n_bt_blocks_, kx_max_, kc_max_, gals, insts, d_bt_off_.as<int>(), n, lk,
	                      nullptr, st),
	          "adjoint conv");
	hip_check(launch_conv_mfma(CONV_ADJ, mc_off_[n], mfma_pitch_, mfma_kch_, gals, insts, d_mc_off_.as<int>(), n, lk,
	                           nullptr, st),
	          "adjoint conv mfma");
	if (rf1_off_[n] == 0) return;
	mark_fft(1, 0, st);
	hip_check(launch_fft_rows_fwd(rf1_off_[n], fft_nmax1_, gals, insts, d_rf1_off_.as<int>(), n, 1, d_tw_.as<double2>(),
	                              twn_, st),
	          "fft rows");
	mark_fft(1, 1, st);
	hip_check(launch_fft_cols(col_off_[n], fft_nmax2_, fft_cb_, gals, insts, d_col_blk_.as<int2>(), 2, 1, 1,
	                          d_tw_.as<double2>(), twn_, nullptr, st),
	          "fft cols");
	mark_fft(1, 2, st);
	hip_check(launch_fft_rows_inv(ri1_off_[n], fft_nmax1_, gals, insts, d_ri1_blk_.as<int2>(), 2, d_tw_.as<double2>(),
	                              twn_, lk, nullptr, st),
	          "fft rows inv");
	mark_fft(1, 3, st);
}

void Engine::set_render_fp64(int mode)
{
	if (mode < 0 || mode > 2) throw std::invalid_argument("render precision must be 0 (FP32), 1 (FP64 values) or 2 (FP64)");
	const bool on = mode >= 1, c64 = mode == 2;
	if (on == render_fp64_ && c64 == canvas64_) return;
	render_fp64_ = on;
	canvas64_ = c64;
	for (auto &g : gal_host_) {
		g.render_fp64 = on ? 1 : 0;
		g.canvas64 = c64 ? 1 : 0;
	}
	gals_dirty_ = true;
	invalidate_graphs(graph_exec_);
	if (!inst_galaxy_.empty()) set_batch(inst_galaxy_, inst_ncomp_);
}

void Engine::set_grad_slots(unsigned mask)
{
	mask &= (1u << NGRAD) - 1u;
	if (mask == grad_slots_) return;
	const bool box = (mask >> G_BOX) & 1u;
	grad_slots_ = mask;
	for (auto &g : gal_host_) g.grad_box = box ? 1 : 0;
	gals_dirty_ = true;
	invalidate_graphs(graph_exec_);
	if (!inst_galaxy_.empty()) set_batch(inst_galaxy_, inst_ncomp_);
}

int Engine::add_galaxy(const GalaxySpec &s)
{
	if (s.W <= 0 || s.H <= 0) throw std::invalid_argument("Model's requested dimensions are 0");
	if (s.scale_x <= 0) throw std::invalid_argument("Model's scale_x cannot be negative or zero");
	if (s.scale_y <= 0) throw std::invalid_argument("Model's scale_y cannot be negative or zero");
	if (s.finesampling <= 0) throw std::invalid_argument("finesampling must be > 0");
	hip_check(hipSetDevice(device_), "hipSetDevice");
	GalaxyDev g{};
	GalaxyBuffers bufs;
	g.W = s.W;
	g.H = s.H;
	g.likefunc = s.likefunc;
	g.magzero = s.magzero;
	g.img_sx = s.scale_x;
	g.img_sy = s.scale_y;
	g.sx = s.scale_x / s.finesampling;
	g.sy = s.scale_y / s.finesampling;
	const bool have_psf = s.psf && s.Kx > 0 && s.Ky > 0;
	const bool conv = have_psf && s.psf_convolves;
	// without a PSF the likelihood path convolves with a 1 x 1 unit kernel,
	// which copies the canvas exactly
	g.has_psf = conv ? 1 : 0;
	g.conv_kind = CK_VALU;
	g.Kx = conv ? s.Kx : 1;
	g.Ky = conv ? s.Ky : 1;
	// Model::analyze_expansion_requirements with a pre-adjusted mask
	// (model.cpp:152-184, adjust_mask = false) or without a mask
	bool pad = conv;
	if (s.calc) pad = s.cw > s.W && s.ch > s.H;
	g.padx = pad ? g.Kx / 2 : 0;
	g.pady = pad ? g.Ky / 2 : 0;
	g.DW = s.W + 2 * g.padx;
	g.DH = s.H + 2 * g.pady;
	if (s.calc && (s.cw != g.DW || s.ch != g.DH)) {
		std::ostringstream os;
		os << "Mask dimensions != drawing dimensions: [" << s.cw << ", " << s.ch << "] != [" << g.DW << ", " << g.DH
		   << "]";
		throw std::invalid_argument(os.str());
	}
	g.has_calc = s.calc != nullptr;
	g.render_fp64 = render_fp64_ ? 1 : 0;
	g.canvas64 = canvas64_ ? 1 : 0;
	g.grad_box = (grad_slots_ >> G_BOX) & 1u;
	const size_t npix = (size_t)s.W * s.H, ndraw = (size_t)g.DW * g.DH;
	if (have_psf) {
		// the model's PSF for point sources (psf.cpp:58-145 reads model.psf, normalised by set_psf)
		std::vector<double> p(s.psf, s.psf + (size_t)s.Kx * s.Ky);
		double sum = 0;
		for (double v : p) sum += v;
		std::vector<float> pt(p.size());
		for (size_t k = 0; k < p.size(); ++k) pt[k] = (float)((s.normalise_psf && sum > 0) ? p[k] / sum : p[k]);
		bufs.psf_pt.alloc(pt.size() * sizeof(float));
		hip_check(hipMemcpy(bufs.psf_pt.as<float>(), pt.data(), pt.size() * 4, hipMemcpyHostToDevice), "psf");
		g.psf_pt = bufs.psf_pt.as<float>();
		g.pkx = s.Kx;
		g.pky = s.Ky;
	}
	g.psf_sx = s.psf_scale_x;
	g.psf_sy = s.psf_scale_y;
	if (!(g.psf_sx > 0 && g.psf_sy > 0)) throw std::invalid_argument("PSF pixel scale must be positive");
	{
		std::vector<double> p(1, 1.0);
		if (conv) {
			p.assign(s.psf, s.psf + (size_t)s.Kx * s.Ky);
			double sum = 0;
			for (double v : p) sum += v;
			if (s.normalise_psf && sum > 0)
				for (double &v : p) v /= sum;  // Model::set_psf normalises (model.h:160-170)
		}
		bufs.psf64 = p;
		std::vector<float> fwd(p.size()), adj(p.size());
		for (int ky = 0; ky < g.Ky; ++ky)
			for (int kx = 0; kx < g.Kx; ++kx) {
				fwd[(size_t)ky * g.Kx + kx] = (float)p[(size_t)(g.Ky - 1 - ky) * g.Kx + (g.Kx - 1 - kx)];
				adj[(size_t)ky * g.Kx + kx] = (float)p[(size_t)ky * g.Kx + kx];
			}
		bufs.psf_fwd.alloc(fwd.size() * sizeof(float));
		bufs.psf_adj.alloc(adj.size() * sizeof(float));
		hip_check(hipMemcpy(bufs.psf_fwd.as<float>(), fwd.data(), fwd.size() * 4, hipMemcpyHostToDevice), "psf");
		hip_check(hipMemcpy(bufs.psf_adj.as<float>(), adj.data(), adj.size() * 4, hipMemcpyHostToDevice), "psf");
		// Toeplitz B fragments of both weight arrays for the matrix-core path
		g.mfma_kc = mfma_kc(g.Kx);
		std::vector<uint16_t> fr;
		mfma_fragments(fwd.data(), g.Kx, g.Ky, fr);
		bufs.mfrag_fwd.alloc(fr.size() * 2);
		hip_check(hipMemcpy(bufs.mfrag_fwd.as<void>(), fr.data(), fr.size() * 2, hipMemcpyHostToDevice), "frag");
		mfma_fragments(adj.data(), g.Kx, g.Ky, fr);
		bufs.mfrag_adj.alloc(fr.size() * 2);
		hip_check(hipMemcpy(bufs.mfrag_adj.as<void>(), fr.data(), fr.size() * 2, hipMemcpyHostToDevice), "frag");
	}
	std::vector<uint8_t> calc(ndraw, 1);
	if (s.calc)
		for (size_t k = 0; k < ndraw; ++k) calc[k] = s.calc[k] ? 1 : 0;
	g.n_calc = 0;
	for (size_t k = 0; k < ndraw; ++k) g.n_calc += calc[k];
	bufs.calc.alloc(ndraw);
	hip_check(hipMemcpy(bufs.calc.as<uint8_t>(), calc.data(), ndraw, hipMemcpyHostToDevice), "calc");
	std::vector<uint8_t> region(npix, 1);
	double nreg = 0;
	for (int j = 0; j < s.H; ++j)
		for (int i = 0; i < s.W; ++i) {
			const size_t o = (size_t)j * s.W + i;
			const uint8_t r = s.region ? (s.region[o] ? 1 : 0) : 1;
			region[o] = r;
			nreg += r;
		}
	g.n_region = nreg;
	bufs.region.alloc(npix);
	hip_check(hipMemcpy(bufs.region.as<uint8_t>(), region.data(), npix, hipMemcpyHostToDevice), "region");
	g.pois_scale = 1;
	if (s.data && s.sigma) {
		std::vector<double> d(npix);
		std::vector<float> iv(npix);
		double smax = 0;
		for (size_t k = 0; k < npix; ++k) {
			d[k] = s.data[k];
			const double sg = s.sigma[k];
			iv[k] = (float)(1.0 / (sg * sg));
			if (region[k]) smax = std::max(smax, std::fabs(s.data[k]) / (std::fabs(sg) * std::fabs(sg)));
		}
		// R/profitLikeModel.R:324-328
		if (smax < 0.1 || smax > 10) g.pois_scale = smax;
		bufs.data.alloc(npix * sizeof(double));
		bufs.ivar.alloc(npix * 4);
		hip_check(hipMemcpy(bufs.data.as<double>(), d.data(), npix * sizeof(double), hipMemcpyHostToDevice), "data");
		hip_check(hipMemcpy(bufs.ivar.as<float>(), iv.data(), npix * 4, hipMemcpyHostToDevice), "ivar");
		std::vector<PixelIn> px(npix);
		for (int j = 0; j < s.H; ++j)
			for (int i = 0; i < s.W; ++i) {
				const size_t o = (size_t)j * s.W + i;
				const size_t qc = (size_t)(j + g.pady) * g.DW + i + g.padx;
				px[o] = PixelIn{d[o], iv[o], (uint32_t)(region[o] | (calc[qc] << 1))};
			}
		bufs.pix.alloc(npix * sizeof(PixelIn));
		hip_check(hipMemcpy(bufs.pix.as<PixelIn>(), px.data(), npix * sizeof(PixelIn), hipMemcpyHostToDevice), "pix");
	}
	g.psf_fwd = bufs.psf_fwd.as<float>();
	g.psf_adj = bufs.psf_adj.as<float>();
	g.mfrag_fwd = bufs.mfrag_fwd.as<uint4>();
	g.mfrag_adj = bufs.mfrag_adj.as<uint4>();
	g.calc = bufs.calc.as<uint8_t>();
	g.region = bufs.region.as<uint8_t>();
	g.data = bufs.data.as<double>();
	g.ivar = bufs.ivar.as<float>();
	g.pix = bufs.pix.as<PixelIn>();
	gal_host_.push_back(g);
	gal_bufs_.push_back(std::move(bufs));
	build_tiles((int)gal_host_.size() - 1);
	gals_dirty_ = true;
	invalidate_graphs(graph_exec_);
	return (int)gal_host_.size() - 1;
}

// Tiles of 64 x (4*conv_rows) pixels: on the image where the likelihood region
// has a pixel (forward convolution), and on the canvas where the calc region
// has one (rendering, adjoint convolution and adjoint rendering).
void Engine::build_tiles(int gi)
{
	GalaxyDev &g = gal_host_[gi];
	GalaxyBuffers &b = gal_bufs_[gi];
	const int th = 4 * conv_rows_;
	g.conv_rows = conv_rows_;
	std::vector<uint8_t> calc((size_t)g.DW * g.DH), region((size_t)g.W * g.H);
	hip_check(hipMemcpy(calc.data(), g.calc, calc.size(), hipMemcpyDeviceToHost), "calc");
	hip_check(hipMemcpy(region.data(), g.region, region.size(), hipMemcpyDeviceToHost), "region");
	auto tiles = [&](const std::vector<uint8_t> &m, int w, int h) {
		std::vector<int2> out;
		for (int ty = 0; ty * th < h; ++ty)
			for (int tx = 0; tx * kTileW < w; ++tx) {
				bool any = false;
				for (int j = ty * th; j < std::min(h, (ty + 1) * th) && !any; ++j)
					for (int i = tx * kTileW; i < std::min(w, (tx + 1) * kTileW); ++i)
						if (m[(size_t)j * w + i]) {
							any = true;
							break;
						}
				if (any) out.push_back(make_int2(tx, ty));
			}
		return out;
	};
	b.ct_host = tiles(region, g.W, g.H);
	b.at_host = tiles(calc, g.DW, g.DH);
	// 128 x 128 tiles of the matrix-core path
	auto big_tiles = [&](const std::vector<uint8_t> &m, int w, int h) {
		std::vector<int2> out;
		for (int ty = 0; ty * kMfmaTile < h; ++ty)
			for (int tx = 0; tx * kMfmaTile < w; ++tx) {
				bool any = false;
				for (int j = ty * kMfmaTile; j < std::min(h, (ty + 1) * kMfmaTile) && !any; ++j)
					for (int i = tx * kMfmaTile; i < std::min(w, (tx + 1) * kMfmaTile); ++i)
						if (m[(size_t)j * w + i]) {
							any = true;
							break;
						}
				if (any) out.push_back(make_int2(tx, ty));
			}
		return out;
	};
	const std::vector<int2> ml = big_tiles(region, g.W, g.H), mc = big_tiles(calc, g.DW, g.DH);
	b.mtiles_like.alloc(std::max<size_t>(1, ml.size()) * sizeof(int2));
	b.mtiles_canvas.alloc(std::max<size_t>(1, mc.size()) * sizeof(int2));
	if (!ml.empty())
		hip_check(hipMemcpy(b.mtiles_like.as<int2>(), ml.data(), ml.size() * sizeof(int2), hipMemcpyHostToDevice),
		          "tiles");
	if (!mc.empty())
		hip_check(hipMemcpy(b.mtiles_canvas.as<int2>(), mc.data(), mc.size() * sizeof(int2), hipMemcpyHostToDevice),
		          "tiles");
	g.mtiles_like = b.mtiles_like.as<int2>();
	g.mtiles_canvas = b.mtiles_canvas.as<int2>();
	g.n_mtl = (int)ml.size();
	g.n_mtc = (int)mc.size();
	b.ctiles.alloc(std::max<size_t>(1, b.ct_host.size()) * sizeof(int2));
	b.atiles.alloc(std::max<size_t>(1, b.at_host.size()) * sizeof(int2));
	if (!b.ct_host.empty())
		hip_check(hipMemcpy(b.ctiles.as<int2>(), b.ct_host.data(), b.ct_host.size() * sizeof(int2), hipMemcpyHostToDevice),
		          "tiles");
	if (!b.at_host.empty())
		hip_check(hipMemcpy(b.atiles.as<int2>(), b.at_host.data(), b.at_host.size() * sizeof(int2), hipMemcpyHostToDevice),
		          "tiles");
	g.ctiles = b.ctiles.as<int2>();
	g.atiles = b.atiles.as<int2>();
	g.n_ctiles = (int)b.ct_host.size();
	g.n_atiles = (int)b.at_host.size();
}

void Engine::upload_galaxies()
{
	if (!gals_dirty_) return;
	d_gals_.alloc(std::max<size_t>(1, gal_host_.size()) * sizeof(GalaxyDev));
	hip_check(hipMemcpy(d_gals_.as<GalaxyDev>(), gal_host_.data(), gal_host_.size() * sizeof(GalaxyDev),
	                    hipMemcpyHostToDevice),
	          "galaxies");
	gals_dirty_ = false;
}

// need: the light-list capacity of each profile family
void Engine::ensure_queue(const long *need, long hneed)
{
	bool grow = false;
	for (int f = 0; f < kFamilies; ++f) grow |= need[f] > lcap_[f];
	if (!grow && hneed <= hcap_) return;
	if (grow) {
		long total = 0;
		for (int f = 0; f < kFamilies; ++f) {
			lcap_[f] = std::max(lcap_[f], need[f]);
			total += lcap_[f];
		}
		if (total > INT32_MAX) throw std::runtime_error("sub-pixel work lists beyond 2^31 entries");
		d_queue_.alloc((size_t)total * sizeof(QEntry));
		d_qjac_.alloc((size_t)total * kJacStride * sizeof(float));
		d_qdeep_.alloc((size_t)total * sizeof(int));
	}
	if (hneed > hcap_) {
		hcap_ = hneed;
		d_hqueue_.alloc((size_t)hcap_ * sizeof(QEntry));
		d_hjac_.alloc((size_t)hcap_ * kJacStride * sizeof(float));
	}
	invalidate_graphs(graph_exec_);
}

WorkQueue Engine::work_queue()
{
	WorkQueue q{d_queue_.as<QEntry>(), d_hqueue_.as<QEntry>(), d_qcount_.as<int>(), (int)hcap_, {},
	            d_qjac_.as<float>(), d_hjac_.as<float>(), d_qdeep_.as<int>()};
	q.lbase[0] = 0;
	for (int f = 0; f < kFamilies; ++f) q.lbase[f + 1] = q.lbase[f] + (int)lcap_[f];
	return q;
}

// whether the last evaluation's work lists fitted
bool Engine::queue_ok() const
{
	const int *q = h_qcount_.as<int>();
	for (int f = 0; f < kFamilies; ++f)
		if (q[light_count(f)] > lcap_[f]) return false;
	return q[1] <= hcap_;
}

// grows the work lists to what the last evaluation needed (drops the graphs)
void Engine::grow_queue()
{
	const int *q = h_qcount_.as<int>();
	long need[kFamilies];
	for (int f = 0; f < kFamilies; ++f) {
		const long n = q[light_count(f)];
		need[f] = n > lcap_[f] ? n + n / 4 + 1024 : lcap_[f];
	}
	ensure_queue(need, std::max<long>(hcap_, (long)q[1] + 64));
}

void Engine::set_batch(const std::vector<int> &inst_galaxy, const std::vector<int> &inst_ncomp)
{
	if (inst_galaxy.size() != inst_ncomp.size()) throw std::invalid_argument("instance arrays differ in length");
	hip_check(hipSetDevice(device_), "hipSetDevice");
	invalidate_graphs(graph_exec_);
	bool any_fft = false;
	kind_override_.resize(gal_host_.size(), -1);
	for (size_t gi = 0; gi < gal_host_.size(); ++gi) {
		GalaxyDev &g = gal_host_[gi];
		// an FP64 canvas is convolved by the FP64 FFT
		const int k = !g.has_psf ? CK_VALU : canvas64_ ? CK_FFT : kind_override_[gi] >= 0 ? kind_override_[gi] : conv_method_;
		if (k != g.conv_kind) {
			g.conv_kind = k;
			gals_dirty_ = true;
		}
		any_fft |= k == CK_FFT;
	}
	if (any_fft) ensure_fft_spectra();
	upload_galaxies();
	const int n = (int)inst_galaxy.size();
	inst_galaxy_ = inst_galaxy;
	inst_ncomp_ = inst_ncomp;
	inst_comp0_.assign(n, 0);
	ct_off_.assign(n + 1, 0);
	at_off_.assign(n + 1, 0);
	ncomp_total_ = 0;
	any_t_ = false;
	all_gauss_ = true;
	kx_max_ = 0;
	long qneed = 1024, hneed = 64;
	inst_bufs_.clear();
	std::vector<InstDev> insts(n);
	std::vector<int> comp_inst;
	for (int i = 0; i < n; ++i) {
		const int gi = inst_galaxy[i];
		if (gi < 0 || gi >= (int)gal_host_.size()) throw std::invalid_argument("unknown galaxy id");
		if (inst_ncomp[i] < 0) throw std::invalid_argument("negative component count");
		const GalaxyDev &g = gal_host_[gi];
		any_t_ |= g.likefunc == LIKE_T;
		all_gauss_ &= g.likefunc == LIKE_CHISQ || g.likefunc == LIKE_NORM;
		inst_comp0_[i] = ncomp_total_;
		ncomp_total_ += inst_ncomp[i];
		for (int k = 0; k < inst_ncomp[i]; ++k) comp_inst.push_back(i);
		ct_off_[i + 1] = ct_off_[i] + (g.conv_kind == CK_VALU ? g.n_ctiles : 0);
		at_off_[i + 1] = at_off_[i] + g.n_atiles;
		if (g.conv_kind == CK_VALU) kx_max_ = std::max(kx_max_, g.Kx);
		const size_t ndraw = (size_t)g.DW * g.DH, npix = (size_t)g.W * g.H;
		DevBuf canvas(ndraw * 4), w(npix * 4), G(ndraw * 4), aux(ndraw * 4), fbuf;
		if (g.conv_kind == CK_FFT) fbuf.alloc((size_t)fft_pitch(g.fft_n1) * g.fft_n2 * sizeof(double2));
		DevBuf dmask(std::max<size_t>(8, (size_t)inst_ncomp[i] * g.DH * ((g.DW + 63) / 64) * sizeof(uint64_t)));
		DevBuf canvas64;
		if (canvas64_) canvas64.alloc(ndraw * sizeof(double));
		insts[i] = InstDev{gi,          inst_comp0_[i], inst_ncomp[i],   0,
		                   canvas.as<float>(), w.as<float>(), G.as<float>(), aux.as<float>(), fbuf.as<double2>(),
		                   dmask.as<uint64_t>(), canvas64.as<double>()};
		inst_bufs_.push_back(std::move(canvas));
		inst_bufs_.push_back(std::move(w));
		inst_bufs_.push_back(std::move(G));
		inst_bufs_.push_back(std::move(aux));
		inst_bufs_.push_back(std::move(fbuf));
		inst_bufs_.push_back(std::move(dmask));
		inst_bufs_.push_back(std::move(canvas64));
		qneed += (long)inst_ncomp[i] * (long)std::min<size_t>(ndraw, 4096);
		// centre pixels: within one image pixel of the centre, (2 fs)^2 at most
		const long fs = std::lround(g.img_sx / g.sx);
		hneed += (long)inst_ncomp[i] * 4 * std::max(1L, fs * fs);
	}
	n_ct_blocks_ = ct_off_[n];
	n_at_blocks_ = at_off_[n];
	inst_aux_.assign(n, 0);
	any_nonradial_ = false;
	// brute-force model / adjoint tiles and the FFT passes, per instance
	bt_off_.assign(n + 1, 0);
	ml_off_.assign(n + 1, 0);
	mc_off_.assign(n + 1, 0);
	lk_off_.assign(n + 1, 0);
	rf0_off_.assign(n + 1, 0);
	rf1_off_.assign(n + 1, 0);
	col_off_.assign(n + 1, 0);
	ri0_off_.assign(n + 1, 0);
	ri1_off_.assign(n + 1, 0);
	fft_nmax1_ = fft_nmax2_ = 0;
	int kc_m = 0, ky_m = 0;
	for (int i = 0; i < n; ++i) {
		const GalaxyDev &g = gal_host_[inst_galaxy[i]];
		if (g.conv_kind == CK_MFMA) {
			kc_m = std::max(kc_m, g.mfma_kc);
			ky_m = std::max(ky_m, g.Ky);
		}
		if (g.conv_kind != CK_FFT) continue;
		fft_nmax1_ = std::max(fft_nmax1_, g.fft_n1);
		fft_nmax2_ = std::max(fft_nmax2_, g.fft_n2);
	}
	mfma_pitch_ = mfma_kch_ = 0;
	if (kc_m) {
		mfma_pitch_ = mfma_pitch(kc_m);
		mfma_kch_ = mfma_rows_per_chunk(mfma_pitch_, ky_m, 150 * 1024);
		if (mfma_kch_ < 1) throw std::invalid_argument("PSF too wide for the split-BF16x3 convolution");
	}
	fft_cb_ = fft_cols_per_block(std::max(fft_nmax2_, 2));
	for (int i = 0; i < n; ++i) {
		const GalaxyDev &g = gal_host_[inst_galaxy[i]];
		const bool f = g.conv_kind == CK_FFT;
		bt_off_[i + 1] = bt_off_[i] + (g.conv_kind == CK_VALU ? g.n_atiles : 0);
		ml_off_[i + 1] = ml_off_[i] + (g.conv_kind == CK_MFMA ? g.n_mtl : 0);
		mc_off_[i + 1] = mc_off_[i] + (g.conv_kind == CK_MFMA ? g.n_mtc : 0);
		// row passes take fft_pairs_per_block pairs of rows per block;
		// columns span the half-spectrum
		const int rpb = f ? 2 * fft_pairs_per_block(g.fft_n1) : 2;
		auto row_blocks = [rpb](int rows) { return (rows + rpb - 1) / rpb; };
		lk_off_[i + 1] = lk_off_[i] + (g.conv_kind == CK_VALU   ? g.n_ctiles
		                               : g.conv_kind == CK_MFMA ? g.n_mtl
		                                                        : row_blocks(g.H));
		rf0_off_[i + 1] = rf0_off_[i] + (f ? row_blocks(g.DH) : 0);
		rf1_off_[i + 1] = rf1_off_[i] + (f ? row_blocks(g.H) : 0);
		col_off_[i + 1] = col_off_[i] + (f ? (g.fft_n1 / 2 + 1 + fft_cb_ - 1) / fft_cb_ : 0);
		ri0_off_[i + 1] = ri0_off_[i] + (f ? row_blocks(g.H) : 0);
		ri1_off_[i + 1] = ri1_off_[i] + (f ? row_blocks(g.DH) : 0);
	}
	n_bt_blocks_ = bt_off_[n];
	// the adjoint render gives each thread about 32 canvas pixels of a component
	adj_grp_ = std::max(1, 32 / conv_rows_);
	ag_off_.assign(n + 1, 0);
	for (int i = 0; i < n; ++i)
		ag_off_[i + 1] = ag_off_[i] + (gal_host_[inst_galaxy[i]].n_atiles + adj_grp_ - 1) / adj_grp_;
	n_ag_blocks_ = ag_off_[n];
	auto upload_off = [&](DevBuf &d, const std::vector<int> &h) {
		d.alloc((n + 1) * sizeof(int));
		hip_check(hipMemcpy(d.as<int>(), h.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "off");
	};
	upload_off(d_bt_off_, bt_off_);
	upload_off(d_ml_off_, ml_off_);
	upload_off(d_mc_off_, mc_off_);
	upload_off(d_lk_off_, lk_off_);
	upload_off(d_rf0_off_, rf0_off_);
	upload_off(d_rf1_off_, rf1_off_);
	upload_off(d_ag_off_, ag_off_);
	// block tables: galaxy by galaxy (in order of first appearance), block
	// index, then the galaxy's instances
	{
		std::vector<std::vector<int>> by_gal(gal_host_.size());
		std::vector<int> gorder;
		for (int i = 0; i < n; ++i) {
			const int gi = inst_galaxy[i];
			if (by_gal[gi].empty()) gorder.push_back(gi);
			by_gal[gi].push_back(i);
		}
		auto table = [&](DevBuf &d, const std::vector<int> &off) {
			std::vector<int2> t;
			t.reserve(off[n]);
			for (int gi : gorder) {
				const int nb = off[by_gal[gi][0] + 1] - off[by_gal[gi][0]];
				for (int b = 0; b < nb; ++b)
					for (int i : by_gal[gi]) t.push_back(make_int2(i, b));
			}
			if ((int)t.size() != off[n]) throw std::logic_error("block table size");
			d.alloc(std::max<size_t>(1, t.size()) * sizeof(int2));
			if (!t.empty())
				hip_check(hipMemcpy(d.as<int2>(), t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice), "blocks");
		};
		table(d_col_blk_, col_off_);
		table(d_ri0_blk_, ri0_off_);
		table(d_ri1_blk_, ri1_off_);
	}
	d_insts_.alloc(std::max(1, n) * sizeof(InstDev));
	if (n) hip_check(hipMemcpy(d_insts_.as<InstDev>(), insts.data(), n * sizeof(InstDev), hipMemcpyHostToDevice), "insts");
	d_ct_off_.alloc((n + 1) * sizeof(int));
	d_at_off_.alloc((n + 1) * sizeof(int));
	hip_check(hipMemcpy(d_ct_off_.as<int>(), ct_off_.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "off");
	hip_check(hipMemcpy(d_at_off_.as<int>(), at_off_.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice), "off");
	const int nc = std::max(1, ncomp_total_);
	d_comp_inst_.alloc(nc * sizeof(int));
	if (ncomp_total_)
		hip_check(hipMemcpy(d_comp_inst_.as<int>(), comp_inst.data(), ncomp_total_ * sizeof(int), hipMemcpyHostToDevice),
		          "comp_inst");
	{
		// the identity order until the first parameters give the families
		std::vector<int> order(nc);
		for (int c = 0; c < nc; ++c) order[c] = c;
		d_setup_order_.alloc(nc * sizeof(int));
		hip_check(hipMemcpy(d_setup_order_.as<int>(), order.data(), nc * sizeof(int), hipMemcpyHostToDevice), "order");
		setup_types_.assign(nc, -1);
	}
	d_params_.alloc(nc * sizeof(ProfileParams));
	d_comps_.alloc(nc * sizeof(Comp));
	d_err_.alloc(nc * sizeof(int));
	d_partial_.alloc(std::max(1, lk_off_[n]) * sizeof(double));
	d_ll_.alloc(std::max(1, n) * sizeof(double));
	d_gscale_.alloc(std::max(1, n) * sizeof(double));
	d_grad_acc_.alloc((size_t)nc * NGRAD * sizeof(double));
	d_grad_out_.alloc((size_t)nc * NGRAD * sizeof(double));
	d_qcount_.alloc(kQueueCounters * sizeof(int));
	h_params_.alloc(nc * sizeof(ProfileParams));
	h_ll_.alloc(std::max(1, n) * sizeof(double));
	h_grad_.alloc((size_t)nc * NGRAD * sizeof(double));
	h_err_.alloc(nc * sizeof(int));
	h_qcount_.alloc(kQueueCounters * sizeof(int));
	for (int f = 0; f < kFamilies; ++f) lcap_[f] = 0;
	hcap_ = 0;
	// each family may take a third of the estimate; a list that overflows
	// grows and the evaluation runs again
	long need[kFamilies];
	for (int f = 0; f < kFamilies; ++f) need[f] = qneed / 3 + 1024;
	ensure_queue(need, hneed);
	int ncu = 256;
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, device_) == hipSuccess) ncu = prop.multiProcessorCount;
	subsample_blocks_ = ncu * 8;
}

void Engine::record(bool grad)
{
	hipStream_t st = stream_;
	const int n = ninst();
	const int nc = ncomp_total_;
	const GalaxyDev *gals = d_gals_.as<GalaxyDev>();
	const InstDev *insts = d_insts_.as<InstDev>();
	hip_check(hipMemcpyAsync(d_params_.as<ProfileParams>(), h_params_.as<ProfileParams>(), nc * sizeof(ProfileParams),
	                         hipMemcpyHostToDevice, st),
	          "params");
	mark(0);
	hip_check(launch_setup(d_params_.as<ProfileParams>(), d_setup_order_.as<int>(), d_comp_inst_.as<int>(), insts, gals, d_comps_.as<Comp>(),
	                       d_err_.as<int>(), nc, st),
	          "setup");
	if (grad && (radial_types_ & ((1u << KING) | (1u << CORESERSIC) | (1u << BROKENEXP))))
		hip_check(launch_setup_grad(d_setup_order_.as<int>(), d_comps_.as<Comp>(), d_err_.as<int>(), nc, st),
		          "setup grad");
	hip_check(launch_zero_i(d_qcount_.as<int>(), kQueueCounters, st), "zero");
	if (grad) hip_check(launch_zero_d(d_grad_acc_.as<double>(), (long)nc * NGRAD, st), "zero");
	mark(1);
	// a batch of one radial family alone takes the render compiled for it
	int rtype = -1;
	for (int t = 0; t < 6; ++t)
		if (!any_nonradial_ && radial_types_ == (1u << t)) rtype = t;
	// the render's FP64 value path only where a pixel can take it
	const bool fp64 = render_fp64_ || canvas64_ || (radial_types_ & ((1u << KING) | (1u << FERRER)));
	hip_check(launch_render(gals, insts, d_at_off_.as<int>(), n, n_at_blocks_, d_comps_.as<Comp>(),
	                        work_queue(), -1, 0, nullptr, st, rtype, fp64),
	          "render");
	mark(2);
	launch_walkers_fwd(0, grad, st);
	mark(3);
	conv_forward(0, st);
	mark(4);
	hip_check(launch_like_finalize(gals, insts, d_lk_off_.as<int>(), n, d_partial_.as<double>(), d_ll_.as<double>(),
	                               d_gscale_.as<double>(), st),
	          "like");
	if (any_t_)
		hip_check(launch_like_t(gals, insts, n, grad ? 1 : 0, d_ll_.as<double>(), d_gscale_.as<double>(), st), "t like");
	mark(5);
	if (grad) {
		conv_adjoint(st);
		mark(6);
		for (int key = 0; key < 12; ++key)
			if ((grad_keys_ >> key) & 1u)
				hip_check(launch_adj_render(key >> 1, key & 1, gals, insts, d_ag_off_.as<int>(), n, n_ag_blocks_, adj_grp_,
				                            d_comps_.as<Comp>(), d_grad_acc_.as<double>(), st),
				          "adjoint render");
		if (any_nonradial_)
			hip_check(launch_adj_nonradial(gals, insts, n, d_comps_.as<Comp>(), d_grad_acc_.as<double>(), st),
			          "adjoint point sources and sky");
		mark(7);
		// the sub-sampled pixels: their Jacobians from the forward walk,
		// weighted by G(q)
		for (int t = 0; t < 6; ++t)
			if ((radial_types_ >> t) & 1u)
				hip_check(launch_adj_queue(t, subsample_blocks_, insts, d_comps_.as<Comp>(), work_queue(),
				                           d_grad_acc_.as<double>(), st),
				          "adjoint subsample");
		mark(8);
		hip_check(launch_grad_finalize(d_comps_.as<Comp>(), d_gscale_.as<double>(), d_grad_acc_.as<double>(),
		                               d_grad_out_.as<double>(), nc, grad_slots_, st),
		          "grad");
		mark(9);
		hip_check(hipMemcpyAsync(h_grad_.as<double>(), d_grad_out_.as<double>(), (size_t)nc * NGRAD * sizeof(double),
		                         hipMemcpyDeviceToHost, st),
		          "grad copy");
	}
	hip_check(hipMemcpyAsync(h_ll_.as<double>(), d_ll_.as<double>(), n * sizeof(double), hipMemcpyDeviceToHost, st),
	          "ll copy");
	hip_check(hipMemcpyAsync(h_err_.as<int>(), d_err_.as<int>(), nc * sizeof(int), hipMemcpyDeviceToHost, st), "err");
	hip_check(hipMemcpyAsync(h_qcount_.as<int>(), d_qcount_.as<int>(), kQueueCounters * sizeof(int), hipMemcpyDeviceToHost, st), "qcount");
}

void Engine::check_setup_errors()
{
	const int *err = h_err_.as<int>();
	for (int c = 0; c < ncomp_total_; ++c)
		if (err[c] != SE_OK) {
			std::ostringstream os;
			os << "component " << c << ": " << setup_error_message(err[c]);
			throw std::invalid_argument(os.str());
		}
}

// Point sources and sky among the current parameters (h_params_): whether
// any is present (the adjoint launches k_adj_nonradial; a change re-records
// the graphs) and, per instance, whether one is drawn without convolution,
// which puts the aux canvas in use (InstDev::has_aux, read by the kernels).
// With enable = false every instance's flag is cleared (the model-canvas path
// draws its aux canvas itself).  Returns whether anything changed.
bool Engine::sync_aux_flags(bool enable)
{
	const ProfileParams *p = h_params_.as<ProfileParams>();
	// the set-up runs one thread per component with the components grouped
	// by profile family, so that a wavefront runs one family's code
	{
		bool same = (int)setup_types_.size() >= ncomp_total_;
		for (int c = 0; same && c < ncomp_total_; ++c) same = setup_types_[c] == p[c].type;
		if (!same && ncomp_total_ > 0) {
			std::vector<int> order;
			order.reserve(ncomp_total_);
			for (int f = 0; f <= kFamilies; ++f)
				for (int c = 0; c < ncomp_total_; ++c)
					if ((is_radial(p[c].type) ? p[c].type : kFamilies) == f) order.push_back(c);
			hip_check(hipMemcpy(d_setup_order_.as<int>(), order.data(), ncomp_total_ * sizeof(int),
			                    hipMemcpyHostToDevice),
			          "order");
			setup_types_.assign(ncomp_total_, 0);
			for (int c = 0; c < ncomp_total_; ++c) setup_types_[c] = p[c].type;
		}
	}
	bool any = false, changed = false;
	unsigned types = 0, keys = 0;
	for (int i = 0; i < ninst(); ++i) {
		int aux = 0;
		for (int k = 0; k < inst_ncomp_[i]; ++k) {
			const ProfileParams &q = p[inst_comp0_[i] + k];
			if (is_radial(q.type)) {
				types |= 1u << q.type;
				keys |= 1u << (2 * q.type + (wants_grad64(q) ? 1 : 0));
				continue;
			}
			any = true;
			if (!q.convolve) aux = 1;
		}
		if (!enable) aux = 0;
		if (aux != inst_aux_[i]) {
			inst_aux_[i] = aux;
			const size_t off = (size_t)i * sizeof(InstDev) + offsetof(InstDev, has_aux);
			hip_check(hipMemcpy(d_insts_.as<char>() + off, &aux, sizeof(int), hipMemcpyHostToDevice), "insts");
			changed = true;
		}
	}
	if (any != any_nonradial_ || types != radial_types_ || keys != grad_keys_) {
		any_nonradial_ = any;
		radial_types_ = types;
		grad_keys_ = keys;
		invalidate_graphs(graph_exec_);
		changed = true;
	}
	return changed;
}

// the forward walk of every radial family present
void Engine::launch_walkers_fwd(int split, bool jac, hipStream_t st)
{
	for (int t = 0; t < 6; ++t)
		if ((radial_types_ >> t) & 1u)
			hip_check(launch_subsample_fwd(t, subsample_blocks_, d_gals_.as<GalaxyDev>(), d_insts_.as<InstDev>(),
			                               d_comps_.as<Comp>(), work_queue(), split, jac, st),
			          "subsample");
}

void Engine::run(bool grad, bool use_graph)
{
	hip_check(hipSetDevice(device_), "hipSetDevice");
	if (timing_) use_graph = false;  // events between kernels need the eager sequence
	for (int attempt = 0; attempt < 4; ++attempt) {
		const int gi = grad ? 1 : 0;
		if (use_graph) {
			if (!graph_exec_[gi]) {
				// first call runs eagerly (sets kernel attributes), later calls replay a graph
				record(grad);
				hip_check(hipStreamSynchronize(stream_), "sync");
				if (queue_ok()) {
					hipGraph_t graph;
					hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture");
					record(grad);
					hip_check(hipStreamEndCapture(stream_, &graph), "capture end");
					hip_check(hipGraphInstantiate(&graph_exec_[gi], graph, nullptr, nullptr, 0), "instantiate");
					(void)hipGraphDestroy(graph);
				}
			}
			else {
				hip_check(hipGraphLaunch(graph_exec_[gi], stream_), "graph launch");
				hip_check(hipStreamSynchronize(stream_), "sync");
			}
		}
		else {
			record(grad);
			hip_check(hipStreamSynchronize(stream_), "sync");
		}
		{
			const int *q = h_qcount_.as<int>();
			int total = q[1];
			for (int f = 0; f < kFamilies; ++f) total += q[light_count(f)];
			stats_.queue_count = total;
		}
		if (queue_ok()) {
			if (timing_) {
				const int last = grad ? ST_COUNT : ST_ADJ_CONV;
				for (int k = 0; k < last; ++k) {
					float ms = 0;
					hip_check(hipEventElapsedTime(&ms, events_[k], events_[k + 1]), "elapsed");
					stats_.ms[k] += ms;
					stats_.calls[k] += 1;
				}
				for (int f = 0; f < 2; ++f) {
					if (!fft_marked_[f]) continue;
					for (int k = 0; k < 3; ++k) {
						float ms = 0;
						hip_check(hipEventElapsedTime(&ms, fft_events_[f][k], fft_events_[f][k + 1]), "elapsed");
						stats_.fft_ms[f][k] += ms;
					}
					stats_.fft_calls[f] += 1;
				}
			}
			fft_marked_[0] = fft_marked_[1] = false;
			stats_.evaluations += 1;
			return;
		}
		grow_queue();  // drops the graphs; run again
	}
	throw std::runtime_error("sub-pixel work queue kept overflowing");
}

void Engine::evaluate(const ProfileParams *params, double *ll, double *grad, bool use_graph)
{
	if (ninst() == 0) return;
	std::memcpy(h_params_.as<ProfileParams>(), params, ncomp_total_ * sizeof(ProfileParams));
	sync_aux_flags(true);
	run(grad != nullptr, use_graph);
	check_setup_errors();
	if (ll) std::memcpy(ll, h_ll_.as<double>(), ninst() * sizeof(double));
	if (grad) std::memcpy(grad, h_grad_.as<double>(), (size_t)ncomp_total_ * NGRAD * sizeof(double));
}

void Engine::diagnostics(const ProfileParams *params, int *cls, int *nodes)
{
	if (ninst() != 1) throw std::invalid_argument("diagnostics need a batch of one instance");
	hip_check(hipSetDevice(device_), "hipSetDevice");
	const GalaxyDev &g = gal_host_[inst_galaxy_[0]];
	const size_t plane = (size_t)g.DW * g.DH;
	const int nc = ncomp_total_;
	DevBuf dcls(plane * nc * sizeof(int)), dnodes(plane * nc * sizeof(int));
	std::memcpy(h_params_.as<ProfileParams>(), params, nc * sizeof(ProfileParams));
	sync_aux_flags(false);
	hipStream_t st = stream_;
	const GalaxyDev *gals = d_gals_.as<GalaxyDev>();
	const InstDev *insts = d_insts_.as<InstDev>();
	for (int attempt = 0; attempt < 4; ++attempt) {
		hip_check(hipMemcpyAsync(d_params_.as<ProfileParams>(), h_params_.as<ProfileParams>(), nc * sizeof(ProfileParams),
		                         hipMemcpyHostToDevice, st),
		          "params");
		hip_check(launch_setup(d_params_.as<ProfileParams>(), d_setup_order_.as<int>(), d_comp_inst_.as<int>(), insts, gals, d_comps_.as<Comp>(),
		                       d_err_.as<int>(), nc, st),
		          "setup");
		hip_check(launch_zero_i(d_qcount_.as<int>(), kQueueCounters, st), "zero");
		hip_check(launch_render(gals, insts, d_at_off_.as<int>(), 1, n_at_blocks_, d_comps_.as<Comp>(),
		                        work_queue(), -1, 0, dcls.as<int>(), st),
		          "render");
		hip_check(launch_subsample_nodes(subsample_blocks_, gals, insts, d_comps_.as<Comp>(), work_queue(), dnodes.as<int>(), st),
		          "nodes");
		hip_check(hipMemcpyAsync(h_err_.as<int>(), d_err_.as<int>(), nc * sizeof(int), hipMemcpyDeviceToHost, st), "err");
		hip_check(hipMemcpyAsync(h_qcount_.as<int>(), d_qcount_.as<int>(), kQueueCounters * sizeof(int), hipMemcpyDeviceToHost, st),
		          "qcount");
		hip_check(hipStreamSynchronize(st), "sync");
		check_setup_errors();
		if (queue_ok()) break;
		grow_queue();
	}
	hip_check(hipMemcpy(cls, dcls.as<int>(), plane * nc * sizeof(int), hipMemcpyDeviceToHost), "cls");
	hip_check(hipMemcpy(nodes, dnodes.as<int>(), plane * nc * sizeof(int), hipMemcpyDeviceToHost), "nodes");
}

void Engine::adjusted_settings(const ProfileParams *params, int *res, int *maxrec, double *acc, double *rsw,
                               double *rmax)
{
	hip_check(hipSetDevice(device_), "hipSetDevice");
	const int nc = ncomp_total_;
	if (nc == 0) return;
	std::memcpy(h_params_.as<ProfileParams>(), params, nc * sizeof(ProfileParams));
	hipStream_t st = stream_;
	hip_check(hipMemcpyAsync(d_params_.as<ProfileParams>(), h_params_.as<ProfileParams>(), nc * sizeof(ProfileParams),
	                         hipMemcpyHostToDevice, st),
	          "params");
	hip_check(launch_setup(d_params_.as<ProfileParams>(), d_setup_order_.as<int>(), d_comp_inst_.as<int>(), d_insts_.as<InstDev>(),
	                       d_gals_.as<GalaxyDev>(), d_comps_.as<Comp>(), d_err_.as<int>(), nc, st),
	          "setup");
	std::vector<Comp> comps(nc);
	hip_check(hipMemcpyAsync(comps.data(), d_comps_.as<Comp>(), nc * sizeof(Comp), hipMemcpyDeviceToHost, st), "comps");
	hip_check(hipMemcpyAsync(h_err_.as<int>(), d_err_.as<int>(), nc * sizeof(int), hipMemcpyDeviceToHost, st), "err");
	hip_check(hipStreamSynchronize(st), "sync");
	check_setup_errors();
	for (int c = 0; c < nc; ++c) {
		const bool radial = is_radial(comps[c].type);
		if (res) res[c] = radial ? comps[c].res : 0;
		if (maxrec) maxrec[c] = radial ? comps[c].maxrec : 0;
		if (acc) acc[c] = radial ? comps[c].acc : 0;
		if (rsw) rsw[c] = radial ? comps[c].rscale_switch : 0;
		if (rmax) rmax[c] = radial ? comps[c].rscale_max : 0;
	}
}

void Engine::model_images(const ProfileParams *params, float *out)
{
	hip_check(hipSetDevice(device_), "hipSetDevice");
	const int n = ninst(), nc = ncomp_total_;
	std::memcpy(h_params_.as<ProfileParams>(), params, nc * sizeof(ProfileParams));
	sync_aux_flags(true);
	hipStream_t st = stream_;
	const GalaxyDev *gals = d_gals_.as<GalaxyDev>();
	const InstDev *insts = d_insts_.as<InstDev>();
	for (int attempt = 0; attempt < 4; ++attempt) {
		hip_check(hipMemcpyAsync(d_params_.as<ProfileParams>(), h_params_.as<ProfileParams>(), nc * sizeof(ProfileParams),
		                         hipMemcpyHostToDevice, st),
		          "params");
		hip_check(launch_setup(d_params_.as<ProfileParams>(), d_setup_order_.as<int>(), d_comp_inst_.as<int>(), insts, gals, d_comps_.as<Comp>(),
		                       d_err_.as<int>(), nc, st),
		          "setup");
		hip_check(launch_zero_i(d_qcount_.as<int>(), kQueueCounters, st), "zero");
		hip_check(launch_render(gals, insts, d_at_off_.as<int>(), n, n_at_blocks_, d_comps_.as<Comp>(),
		                        work_queue(), -1, 0, nullptr, st),
		          "render");
		launch_walkers_fwd(0, false, st);
		conv_forward(1, st);
		hip_check(hipMemcpyAsync(h_err_.as<int>(), d_err_.as<int>(), nc * sizeof(int), hipMemcpyDeviceToHost, st), "err");
		hip_check(hipMemcpyAsync(h_qcount_.as<int>(), d_qcount_.as<int>(), kQueueCounters * sizeof(int), hipMemcpyDeviceToHost, st),
		          "qcount");
		hip_check(hipStreamSynchronize(st), "sync");
		check_setup_errors();
		if (queue_ok()) break;
		grow_queue();
	}
	size_t off = 0;
	for (int i = 0; i < n; ++i) {
		const GalaxyDev &g = gal_host_[inst_galaxy_[i]];
		std::vector<float> canvas((size_t)g.DW * g.DH);
		const float *src = inst_bufs_[kInstBufs * i + 2].as<float>();
		hip_check(hipMemcpy(canvas.data(), src, canvas.size() * 4, hipMemcpyDeviceToHost), "model");
		for (int j = 0; j < g.H; ++j)
			for (int k = 0; k < g.W; ++k) out[off + (size_t)j * g.W + k] = canvas[(size_t)(j + g.pady) * g.DW + k + g.padx];
		off += (size_t)g.W * g.H;
	}
}

void Engine::model_canvas(const ProfileParams *params, float *out, float *out_direct)
{
	if (ninst() != 1) throw std::invalid_argument("model_canvas needs a batch of one instance");
	hip_check(hipSetDevice(device_), "hipSetDevice");
	const int nc = ncomp_total_;
	std::memcpy(h_params_.as<ProfileParams>(), params, nc * sizeof(ProfileParams));
	sync_aux_flags(false);
	hipStream_t st = stream_;
	const GalaxyDev *gals = d_gals_.as<GalaxyDev>();
	const InstDev *insts = d_insts_.as<InstDev>();
	const GalaxyDev &g = gal_host_[inst_galaxy_[0]];
	const bool conv = g.has_psf != 0;
	for (int attempt = 0; attempt < 4; ++attempt) {
		hip_check(hipMemcpyAsync(d_params_.as<ProfileParams>(), h_params_.as<ProfileParams>(), nc * sizeof(ProfileParams),
		                         hipMemcpyHostToDevice, st),
		          "params");
		hip_check(launch_setup(d_params_.as<ProfileParams>(), d_setup_order_.as<int>(), d_comp_inst_.as<int>(), insts, gals, d_comps_.as<Comp>(),
		                       d_err_.as<int>(), nc, st),
		          "setup");
		hip_check(launch_zero_i(d_qcount_.as<int>(), kQueueCounters, st), "zero");
		// convolved profiles into the canvas, the others into the aux canvas
		hip_check(launch_render(gals, insts, d_at_off_.as<int>(), 1, n_at_blocks_, d_comps_.as<Comp>(),
		                        work_queue(), conv ? 1 : -1, 0, nullptr, st),
		          "render");
		if (conv)
			hip_check(launch_render(gals, insts, d_at_off_.as<int>(), 1, n_at_blocks_, d_comps_.as<Comp>(),
			                        work_queue(), 0, 1, nullptr, st),
			          "render");
		launch_walkers_fwd(conv ? 1 : 0, false, st);
		if (conv && !out_direct) conv_forward(1, st);
		hip_check(hipMemcpyAsync(h_err_.as<int>(), d_err_.as<int>(), nc * sizeof(int), hipMemcpyDeviceToHost, st), "err");
		hip_check(hipMemcpyAsync(h_qcount_.as<int>(), d_qcount_.as<int>(), kQueueCounters * sizeof(int), hipMemcpyDeviceToHost, st),
		          "qcount");
		hip_check(hipStreamSynchronize(st), "sync");
		check_setup_errors();
		if (queue_ok()) break;
		grow_queue();
	}
	const size_t nd = (size_t)g.DW * g.DH;
	if (out_direct) {
		// the two canvases before any convolution
		if (conv) {
			hip_check(hipMemcpy(out, inst_bufs_[0].as<float>(), nd * 4, hipMemcpyDeviceToHost), "model");
			hip_check(hipMemcpy(out_direct, inst_bufs_[3].as<float>(), nd * 4, hipMemcpyDeviceToHost), "model");
		}
		else {
			std::memset(out, 0, nd * 4);
			hip_check(hipMemcpy(out_direct, inst_bufs_[0].as<float>(), nd * 4, hipMemcpyDeviceToHost), "model");
		}
	}
	else if (conv) {
		std::vector<float> aux(nd);
		hip_check(hipMemcpy(out, inst_bufs_[2].as<float>(), nd * 4, hipMemcpyDeviceToHost), "model");
		hip_check(hipMemcpy(aux.data(), inst_bufs_[3].as<float>(), nd * 4, hipMemcpyDeviceToHost), "model");
		for (size_t k = 0; k < nd; ++k) out[k] += aux[k];
	}
	else {
		hip_check(hipMemcpy(out, inst_bufs_[0].as<float>(), nd * 4, hipMemcpyDeviceToHost), "model");
	}
}

void Engine::convolve_canvas(const float *canvas, float *out)
{
	if (ninst() != 1) throw std::invalid_argument("convolve_canvas needs a batch of one instance");
	hip_check(hipSetDevice(device_), "hipSetDevice");
	const GalaxyDev &g = gal_host_[inst_galaxy_[0]];
	const size_t nd = (size_t)g.DW * g.DH;
	hip_check(hipMemcpy(inst_bufs_[0].as<float>(), canvas, nd * 4, hipMemcpyHostToDevice), "canvas");
	if (g.has_psf) {
		conv_forward(1, stream_);
		hip_check(hipStreamSynchronize(stream_), "sync");
		hip_check(hipMemcpy(out, inst_bufs_[2].as<float>(), nd * 4, hipMemcpyDeviceToHost), "out");
	}
	else {
		std::memcpy(out, canvas, nd * 4);
	}
}

void Engine::dilate(const uint8_t *in, uint8_t *out, int w, int h, int padx, int pady, int device)
{
	hip_check(hipSetDevice(device), "hipSetDevice");
	const size_t n = (size_t)w * h;
	DevBuf a(n), b(n), c(n);
	hip_check(hipMemcpy(a.as<uint8_t>(), in, n, hipMemcpyHostToDevice), "mask");
	hip_check(launch_dilate(a.as<uint8_t>(), b.as<uint8_t>(), c.as<uint8_t>(), w, h, padx, pady, nullptr), "dilate");
	hip_check(hipDeviceSynchronize(), "sync");
	hip_check(hipMemcpy(out, c.as<uint8_t>(), n, hipMemcpyDeviceToHost), "mask");
}

void convolve_image(const double *src, int w, int h, const double *krn, int kw, int kh, const uint8_t *mask,
                    double *out, int device, int method)
{
	hip_check(hipSetDevice(device), "hipSetDevice");
	if (method == CK_FFT || method == CK_MFMA) {
		// FFT or matrix-core path: a one-instance engine whose canvas is the source image
		// (no padding, the kernel as given, the mask as its calc region)
		const size_t n = (size_t)w * h;
		std::vector<double> zeros(n, 0.0), ones(n, 1.0);
		std::vector<float> s(n), r(n);
		// an all-true mask when none is given keeps the canvas unpadded
		std::vector<uint8_t> all;
		if (!mask) {
			all.assign(n, 1);
			mask = all.data();
		}
		for (size_t k = 0; k < n; ++k) s[k] = (float)src[k];
		GalaxySpec gs;
		gs.W = w;
		gs.H = h;
		gs.psf = krn;
		gs.Kx = kw;
		gs.Ky = kh;
		gs.normalise_psf = false;
		gs.calc = mask;
		gs.cw = w;
		gs.ch = h;
		gs.data = zeros.data();
		gs.sigma = ones.data();
		Engine eng(device);
		eng.set_conv_method(method);
		eng.add_galaxy(gs);
		eng.set_batch({0}, {0});
		if (eng.galaxy(0).DW != w || eng.galaxy(0).DH != h) throw std::logic_error("convolution canvas was padded");
		eng.convolve_canvas(s.data(), r.data());
		for (size_t k = 0; k < n; ++k) out[k] = r[k];
		return;
	}
	const int rows = 4;
	const size_t n = (size_t)w * h;
	std::vector<float> s(n), f((size_t)kw * kh);
	for (size_t k = 0; k < n; ++k) s[k] = (float)src[k];
	for (int ky = 0; ky < kh; ++ky)
		for (int kx = 0; kx < kw; ++kx) f[(size_t)ky * kw + kx] = (float)krn[(size_t)(kh - 1 - ky) * kw + (kw - 1 - kx)];
	std::vector<int2> tiles;
	const int th = 4 * rows;
	for (int ty = 0; ty * th < h; ++ty)
		for (int tx = 0; tx * kTileW < w; ++tx) tiles.push_back(make_int2(tx, ty));
	DevBuf d_src(n * 4), d_out(n * 4), d_krn(f.size() * 4), d_mask(mask ? n : 1), d_tiles(tiles.size() * sizeof(int2));
	DevBuf d_gal(sizeof(GalaxyDev)), d_inst(sizeof(InstDev)), d_off(2 * sizeof(int));
	hip_check(hipMemcpy(d_src.as<float>(), s.data(), n * 4, hipMemcpyHostToDevice), "src");
	hip_check(hipMemcpy(d_krn.as<float>(), f.data(), f.size() * 4, hipMemcpyHostToDevice), "krn");
	if (mask) hip_check(hipMemcpy(d_mask.as<uint8_t>(), mask, n, hipMemcpyHostToDevice), "mask");
	hip_check(hipMemcpy(d_tiles.as<int2>(), tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice), "tiles");
	GalaxyDev g{};
	g.W = g.DW = w;
	g.H = g.DH = h;
	g.Kx = kw;
	g.Ky = kh;
	g.has_calc = mask != nullptr;
	g.calc = mask ? d_mask.as<uint8_t>() : nullptr;
	g.psf_fwd = d_krn.as<float>();
	g.atiles = d_tiles.as<int2>();
	g.n_atiles = (int)tiles.size();
	g.conv_rows = rows;
	InstDev I{0, 0, 0, 0, d_src.as<float>(), nullptr, d_out.as<float>(), nullptr, nullptr, nullptr};
	const int off[2] = {0, (int)tiles.size()};
	hip_check(hipMemcpy(d_gal.as<GalaxyDev>(), &g, sizeof(g), hipMemcpyHostToDevice), "gal");
	hip_check(hipMemcpy(d_inst.as<InstDev>(), &I, sizeof(I), hipMemcpyHostToDevice), "inst");
	hip_check(hipMemcpy(d_off.as<int>(), off, sizeof(off), hipMemcpyHostToDevice), "off");
	const int kc = std::min(kh, 64);
	hip_check(launch_conv(1, rows, (int)tiles.size(), kw, kc, d_gal.as<GalaxyDev>(), d_inst.as<InstDev>(), d_off.as<int>(),
	                      1, nullptr, nullptr, nullptr),
	          "conv");
	hip_check(hipDeviceSynchronize(), "sync");
	std::vector<float> o(n);
	hip_check(hipMemcpy(o.data(), d_out.as<float>(), n * 4, hipMemcpyDeviceToHost), "out");
	for (size_t k = 0; k < n; ++k) out[k] = o[k];
}

}  // namespace profit_amd
