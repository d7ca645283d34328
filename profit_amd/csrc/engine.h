// Host engine: device-resident galaxies (PSF, masks, data) and batches of
// instances (parameter sets) evaluated by one launch sequence per batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.h"

namespace profit_amd {

// kernel launchers (kernels.hip)
// order: the components grouped by profile family (one thread each)
int launch_setup(const ProfileParams *params, const int *order, const int *comp_inst, const InstDev *insts,
                 const GalaxyDev *gals, Comp *comps, int *errors, int ncomp, hipStream_t st);
// the QUADPACK luminosity gradients (King, core-Sersic, broken exponential)
int launch_setup_grad(const int *order, Comp *comps, const int *errors, int ncomp, hipStream_t st);
int launch_zero_d(double *p, long n, hipStream_t st);
int launch_zero_i(int *p, long n, hipStream_t st);
// fp64: some pixel may take an FP64 value (the FP64 model modes, King or
// Ferrer components in the batch); false only when none can
int launch_render(const GalaxyDev *gals, const InstDev *insts, const int *tile_off, int ninst, int nblocks,
                  const Comp *comps, WorkQueue wq, int filter, int target, int *diag_cls, hipStream_t st, int type = -1,
                  bool fp64 = true);
// the forward walker and the adjoint render, one launch per radial family
// (and per sum precision in the adjoint render: 1 = FP64 per-thread sums,
// steep cores only).  jac: the forward walk also writes each entry's Jacobian
// (wq.ljac, wq.hjac), which launch_adj_queue weighs by G(q).
int launch_subsample_fwd(int type, int nblocks, const GalaxyDev *gals, const InstDev *insts, const Comp *comps,
                         WorkQueue wq, int split, bool jac, hipStream_t st);
int launch_adj_queue(int type, int nblocks, const InstDev *insts, const Comp *comps, WorkQueue wq, double *grad_acc,
                     hipStream_t st);
int launch_subsample_nodes(int nblocks, const GalaxyDev *gals, const InstDev *insts, const Comp *comps,
                           WorkQueue wq, int *nodes, hipStream_t st);
int launch_conv(int mode, int rows, int nblocks, int kx_max, int kc, const GalaxyDev *gals, const InstDev *insts,
                const int *tile_off, int ninst, const int *lk_base, double *partial, hipStream_t st);
int launch_like_finalize(const GalaxyDev *gals, const InstDev *insts, const int *lk_off, int ninst,
                         const double *partial, double *ll, double *gscale, hipStream_t st);
int launch_adj_nonradial(const GalaxyDev *gals, const InstDev *insts, int ninst, const Comp *comps, double *grad_acc,
                         hipStream_t st);
int launch_like_t(const GalaxyDev *gals, const InstDev *insts, int ninst, int grad, double *ll, double *gscale,
                  hipStream_t st);
int launch_adj_render(int type, int prec, const GalaxyDev *gals, const InstDev *insts, const int *grp_off, int ninst,
                      int nblocks, int grp, const Comp *comps, double *grad_acc, hipStream_t st);
int launch_grad_finalize(const Comp *comps, const double *gscale, const double *grad_acc, double *grad_out,
                         int ncomp, unsigned slot_mask, hipStream_t st);
int launch_dilate(const uint8_t *in, uint8_t *tmp, uint8_t *out, int w, int h, int padx, int pady, hipStream_t st);
// FFT path (fft.hip)
int launch_fft_rows_fwd(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int *row_off,
                        int ninst, int mode, const double2 *tw, int twn, hipStream_t st);
// mode 0 of launch_fft_rows_fwd from the FP64 canvas (GalaxyDev::canvas64)
int launch_fft_rows_fwd64(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int *row_off,
                          int ninst, const double2 *tw, int twn, hipStream_t st);
int launch_fft_psf_rows(int n1, int n2, const GalaxyDev *gal, const double2 *tw, int twn, const double *src64,
                        double2 *dst64, hipStream_t st);
// The column pass and the inverse rows take a block table: (instance, block
// of that instance) per launched block.
int launch_fft_cols(int nblocks, int n2max, int cb, const GalaxyDev *gals, const InstDev *insts, const int2 *blocks,
                    int op, int in_mode, int out_mode, const double2 *tw, int twn, double2 *dst64, hipStream_t st);
// mode 0 likelihood, 1 model image, 2 adjoint image, 3 likelihood of a batch
// whose galaxies all take a Gaussian likelihood (chi^2 or normal)
int launch_fft_rows_inv(int nblocks, int nmax, const GalaxyDev *gals, const InstDev *insts, const int2 *blocks,
                        int mode, const double2 *tw, int twn, const int *lk_base, double *partial, hipStream_t st);
int fft_cols_per_block(int n2);
// split-BF16x3 matrix-core path (conv_mfma.hip)
int mfma_kc(int kx);
int mfma_pitch(int kc);
int mfma_rows_per_chunk(int pitch, int ky, size_t budget);
void mfma_fragments(const float *wt, int kx, int ky, std::vector<uint16_t> &out);
int launch_conv_mfma(int mode, int nblocks, int pitch, int kch, const GalaxyDev *gals, const InstDev *insts,
                     const int *tile_off, int ninst, const int *lk_base, double *partial, hipStream_t st);

// Convolver::convolve_impl on the GPU, masked output pixels are 0: method 0
// tiled brute force, 1 FFT.  The kernel is used as given (libprofit
// normalises only the Model's PSF).
void convolve_image(const double *src, int w, int h, const double *krn, int kw, int kh, const uint8_t *mask,
                    double *out, int device, int method = 0);

class hip_error : public std::runtime_error {
  public:
	using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what);
inline void hip_check(int e, const char *what) { hip_check((hipError_t)e, what); }

// owning device allocation
class DevBuf {
  public:
	DevBuf() = default;
	explicit DevBuf(size_t bytes) { alloc(bytes); }
	DevBuf(const DevBuf &) = delete;
	DevBuf &operator=(const DevBuf &) = delete;
	DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
	DevBuf &operator=(DevBuf &&o) noexcept
	{
		if (this != &o) {
			release();
			p_ = o.p_;
			n_ = o.n_;
			o.p_ = nullptr;
			o.n_ = 0;
		}
		return *this;
	}
	~DevBuf() { release(); }
	void alloc(size_t bytes);
	void release();
	template <class T> T *as() const { return static_cast<T *>(p_); }
	size_t size() const { return n_; }

  private:
	void *p_ = nullptr;
	size_t n_ = 0;
};

// pinned host staging
class HostBuf {
  public:
	HostBuf() = default;
	HostBuf(const HostBuf &) = delete;
	HostBuf &operator=(const HostBuf &) = delete;
	~HostBuf() { release(); }
	void alloc(size_t bytes);
	void release();
	template <class T> T *as() const { return static_cast<T *>(p_); }
	size_t size() const { return n_; }

  private:
	void *p_ = nullptr;
	size_t n_ = 0;
};

// A galaxy as the likelihood engine sees it: ProFit's profitSetupData output.
struct GalaxySpec {
	int W = 0, H = 0;                    // image dimensions
	double scale_x = 1, scale_y = 1;     // image pixel scale
	double magzero = 0;
	const double *psf = nullptr;         // PSF (normalised here), Ky rows of Kx
	int Kx = 0, Ky = 0;
	const uint8_t *calc = nullptr;       // pre-adjusted calc region (cw x ch) or null
	int cw = 0, ch = 0;
	const uint8_t *region = nullptr;     // W*H likelihood region, null = every pixel
	const double *data = nullptr;        // W*H
	const double *sigma = nullptr;       // W*H
	int likefunc = LIKE_CHISQ;
	int finesampling = 1;                // W, H are then the fine grid; profiles use scale/finesampling
	bool normalise_psf = true;           // Model::set_psf normalises; a bare Convolver does not
	bool psf_convolves = true;           // false: the PSF only draws point sources (no padding)
	double psf_scale_x = 1, psf_scale_y = 1;  // Model::set_psf_pixel_scale
};

// Accumulated per-stage device time (HIP events on the engine stream) over the
// evaluations run while timing is on.
enum Stage { ST_SETUP = 0, ST_RENDER, ST_SUBSAMPLE, ST_CONV, ST_LIKE, ST_ADJ_CONV, ST_ADJ_RENDER, ST_ADJ_SUB, ST_GRAD,
             ST_COUNT };
struct EngineStats {
	double ms[ST_COUNT] = {0};
	long calls[ST_COUNT] = {0};
	int queue_count = 0;
	long evaluations = 0;
	// the FFT stages kernel by kernel: [forward, adjoint][rows forward,
	// columns, rows inverse], over the evaluations that ran FFT galaxies
	double fft_ms[2][3] = {{0}};
	long fft_calls[2] = {0};
};

class Engine {
  public:
	explicit Engine(int device = 0);
	~Engine();
	Engine(const Engine &) = delete;
	Engine &operator=(const Engine &) = delete;

	int add_galaxy(const GalaxySpec &spec);
	// convolution method for every galaxy: 0 brute force (LDS-tiled FP32 with
	// FP64 column sums), 1 FFT (FP64, power-of-two frames up to 4096)
	void set_conv_method(int m);
	int conv_method() const { return conv_method_; }
	// Per-galaxy convolution kernels picked on the current batch (see
	// autotune in engine.cpp).  ms: the forward + adjoint convolution time per
	// evaluation of FP32 brute force at 4, 8 and 16 rows per thread, FFT,
	// split-BF16x3 and the chosen configuration.
	struct TuneReport {
		double ms[6] = {0};
		int rows = 8;
		std::vector<int> kinds;  // per galaxy (CK_*)
	};
	TuneReport autotune(const ProfileParams *params, int reps, bool allow_bf16x3 = false);
	// a single instance's canvas supplied directly (no profiles): the model-mode
	// convolution of `canvas` (DW*DH floats) into out (DW*DH floats)
	void convolve_canvas(const float *canvas, float *out);
	void set_batch(const std::vector<int> &inst_galaxy, const std::vector<int> &inst_ncomp);
	// params: one ProfileParams per component, instance-major.  ll: one per
	// instance; grad (nullable): NGRAD per component.
	void evaluate(const ProfileParams *params, double *ll, double *grad, bool use_graph = true);
	// rendering diagnostics for parity tests: per-component class and node maps
	void diagnostics(const ProfileParams *params, int *cls, int *nodes);
	// the adjusted sub-pixel settings of every component of the batch (set-up
	// only): resolution, max_recursions, acc, rscale_switch, rscale_max
	void adjusted_settings(const ProfileParams *params, int *res, int *maxrec, double *acc, double *rsw,
	                       double *rmax);
	// the convolved model of every instance (W*H floats each, region-independent)
	void model_images(const ProfileParams *params, float *out);
	// Model::evaluate's drawing canvas for a single instance: convolved profiles
	// (masked by the calc region) plus the profiles drawn without convolution.
	// With out_direct given, nothing is convolved: out receives the canvas of
	// the profiles to be convolved and out_direct that of the others, as
	// Model::produce_image holds them before its convolver runs.
	void model_canvas(const ProfileParams *params, float *out, float *out_direct = nullptr);
	// Mask::expand_by on the device
	static void dilate(const uint8_t *in, uint8_t *out, int w, int h, int padx, int pady, int device = 0);

	int ninst() const { return (int)inst_galaxy_.size(); }
	int ncomp() const { return ncomp_total_; }
	int conv_rows() const { return conv_rows_; }
	void set_conv_rows(int r);
	// profile values (direct pixels, recursion leaves) in FP32 (default) or FP64
	// (1); 2: FP64 values, an FP64 canvas and the FFT for every PSF
	// convolution, the model image FP64 end to end
	void set_render_fp64(int mode);
	// gradient slots computed (bit s = slot s); the others come back as 0.
	// Leaving out the box slot spares the per-sample logarithms of d / d box.
	void set_grad_slots(unsigned mask);
	unsigned grad_slots() const { return grad_slots_; }
	bool render_fp64() const { return render_fp64_; }
	bool canvas_fp64() const { return canvas64_; }
	void set_timing(bool on);
	void reset_stats() { stats_ = EngineStats{}; }
	// algorithmic work of the batch: pixels of the likelihood region and of the
	// calc region summed over instances, and the PSF taps of each
	double conv_flops_forward() const;
	double conv_flops_adjoint() const;
	const EngineStats &stats() const { return stats_; }
	const GalaxyDev &galaxy(int g) const { return gal_host_.at(g); }
	hipStream_t stream() const { return stream_; }

  private:
	struct GalaxyBuffers {
		DevBuf psf_fwd, psf_adj, calc, region, data, ivar, pix, rtiles, ctiles, atiles, fft_spec;
		DevBuf mfrag_fwd, mfrag_adj, mtiles_like, mtiles_canvas, psf_pt;
		std::vector<int2> ct_host, at_host;
		std::vector<double> psf64;  // normalised (or raw) PSF, Ky rows of Kx
	};
	void ensure_fft_spectra();
	void conv_forward(int mode, hipStream_t st);  // 0 likelihood, 1 model image
	void conv_adjoint(hipStream_t st);
	void build_tiles(int g);
	void upload_galaxies();
	void run(bool grad, bool use_graph);
	void record(bool grad);
	void ensure_queue(const long *need, long hneed);
	WorkQueue work_queue();
	bool queue_ok() const;
	void grow_queue();
	void check_setup_errors();

	int device_ = 0;
	hipStream_t stream_ = nullptr;
	int conv_rows_ = 8;
	int conv_method_ = 0;
	std::vector<int> kind_override_;  // per galaxy: -1 follows conv_method_
	double time_conv(const ProfileParams *params, int reps);
	DevBuf d_tw_;
	int twn_ = 0;
	int fft_nmax1_ = 0, fft_nmax2_ = 0;
	// per-instance launch tables (prefix sums of blocks): brute-force model and
	// adjoint tiles, and the FFT passes (row forward on the canvas / image,
	// columns, row inverse on the image / canvas)
	// split-BF16x3 tiles (likelihood / canvas), and the likelihood partials
	// base of each instance
	std::vector<int> bt_off_, rf0_off_, rf1_off_, col_off_, ri0_off_, ri1_off_, ml_off_, mc_off_, lk_off_;
	DevBuf d_bt_off_, d_rf0_off_, d_rf1_off_, d_ml_off_, d_mc_off_, d_lk_off_;
	// block tables of the column pass and the inverse rows, ordered by galaxy,
	// then block, then instance: the blocks in flight together share their
	// galaxy's PSF spectrum columns (and pixel data) in the L2
	DevBuf d_col_blk_, d_ri0_blk_, d_ri1_blk_;
	int n_bt_blocks_ = 0;
	int mfma_pitch_ = 0, mfma_kch_ = 0;
	int fft_cb_ = 8;
	bool render_fp64_ = false;
	bool canvas64_ = false;
	unsigned grad_slots_ = (1u << NGRAD) - 1u;
	int kc_max_ = 64;
	bool timing_ = false;
	EngineStats stats_;

	std::vector<GalaxyDev> gal_host_;
	std::vector<GalaxyBuffers> gal_bufs_;
	DevBuf d_gals_;
	bool gals_dirty_ = true;

	std::vector<int> inst_galaxy_, inst_ncomp_, inst_comp0_;
	int ncomp_total_ = 0;
	bool any_t_ = false;  // an instance of the batch uses the t likelihood
	bool all_gauss_ = true;  // every instance of the batch takes chi^2 or the normal likelihood
	// point sources and sky of the current parameters: any at all (adjoint
	// pass), and per instance whether one is drawn without convolution
	bool any_nonradial_ = false;
	std::vector<int> inst_aux_;
	// radial families present (bit t) and the adjoint's (family, precision)
	// keys present (bit 2t + prec), which decide the walker launches
	unsigned radial_types_ = 0, grad_keys_ = 0;
	void launch_walkers_fwd(int split, bool jac, hipStream_t st);
	bool sync_aux_flags(bool enable);
	std::vector<int> ct_off_, at_off_, ag_off_;
	int n_ct_blocks_ = 0, n_at_blocks_ = 0, n_ag_blocks_ = 0;
	int adj_grp_ = 1;  // canvas tiles per block of the adjoint render
	int kx_max_ = 0;
	DevBuf d_insts_, d_ct_off_, d_at_off_, d_ag_off_, d_comp_inst_;
	// the set-up's order of the components (grouped by profile family) and
	// the types it was formed from
	DevBuf d_setup_order_;
	std::vector<int> setup_types_;
	// per instance: canvas, w, G, aux, FFT frame, direct-pixel mask, FP64 canvas
	static constexpr int kInstBufs = 7;
	std::vector<DevBuf> inst_bufs_;
	DevBuf d_params_, d_comps_, d_err_, d_queue_, d_qcount_, d_partial_, d_ll_, d_gscale_, d_grad_acc_, d_grad_out_;
	long lcap_[kFamilies] = {}, hcap_ = 0;
	DevBuf d_hqueue_, d_qjac_, d_hjac_, d_qdeep_;
	HostBuf h_params_, h_ll_, h_grad_, h_err_, h_qcount_;

	hipGraphExec_t graph_exec_[2] = {nullptr, nullptr};
	hipEvent_t events_[ST_COUNT + 1] = {};
	void mark(int k);
	// around the three FFT kernels of the forward (0) and adjoint (1) stage
	hipEvent_t fft_events_[2][4] = {};
	bool fft_marked_[2] = {false, false};
	void mark_fft(int stage, int k, hipStream_t st);
	int subsample_blocks_ = 2048;
};

}  // namespace profit_amd
