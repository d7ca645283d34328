// C-ABI of profit_amd (declared in include/profit_amd.h).  Catches every
// exception and reports it through profit_amd_last_error().
#include "../../include/profit_amd.h"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "engine.h"
#include "noalloc.h"
#include "profit.h"

static_assert(sizeof(profit_amd_profile) == sizeof(profit_amd::ProfileParams), "profile layout");

namespace {

thread_local std::string g_error;

template <class F>
int guarded(F &&f)
{
	try {
		f();
		g_error.clear();
		return 0;
	} catch (const std::exception &e) {
		g_error = e.what();
	} catch (...) {
		g_error = "unknown error";
	}
	return -1;
}

const char *kProfileNames[6] = {"sersic", "moffat", "ferrer", "king", "coresersic", "brokenexp"};

void build_model(profit::Model &m, const profit_amd_model_desc *d, const profit_amd_profile *profiles, int n)
{
	m.set_image_pixel_scale({d->scale_x, d->scale_y});
	m.set_magzero(d->magzero);
	m.set_finesampling(d->finesampling);
	m.set_return_finesampled(d->return_finesampled != 0);
	m.set_crop(d->crop != 0);
	m.set_adjust_mask(d->adjust_mask != 0);
	m.set_dry_run(d->dry_run != 0);
	m.set_device(d->device);
	if (d->psf && d->psf_w > 0 && d->psf_h > 0) {
		std::vector<double> p(d->psf, d->psf + (size_t)d->psf_w * d->psf_h);
		m.set_psf(profit::Image(p, d->psf_w, d->psf_h));
	}
	if (d->mask && d->mask_w > 0 && d->mask_h > 0) m.set_mask(profit::Mask(d->mask, d->mask_w, d->mask_h));
	m.set_psf_pixel_scale({d->psf_scale_x > 0 ? d->psf_scale_x : 1.0, d->psf_scale_y > 0 ? d->psf_scale_y : 1.0});
	switch (d->convolver) {
	case PROFIT_AMD_CONV_BRUTE_OLD: m.set_convolver(profit::create_convolver(profit::BRUTE_OLD)); break;
	case PROFIT_AMD_CONV_BRUTE: m.set_convolver(profit::create_convolver(profit::BRUTE)); break;
	case PROFIT_AMD_CONV_FFT: m.set_convolver(profit::create_convolver(profit::FFT)); break;
	case PROFIT_AMD_CONV_NULL: m.set_convolver(profit::create_convolver(profit::NO_OP)); break;
	case PROFIT_AMD_CONV_BRUTE_BF16X3: m.set_convolver(profit::create_convolver(profit::BRUTE_BF16X3)); break;
	default: throw profit::invalid_parameter("Unsupported convolver type: " + std::to_string(d->convolver));
	}
	for (int k = 0; k < n; ++k) {
		const profit_amd_profile &pp = profiles[k];
		profit::ProfilePtr p;
		if (pp.type >= 0 && pp.type < 6) {
			p = m.add_profile(kProfileNames[pp.type]);
			std::memcpy(&p->params(), &pp, sizeof(pp));
			p->parameter("convolve", pp.convolve != 0);
		}
		else if (pp.type == PROFIT_AMD_SKY) {
			p = m.add_profile("sky");
			p->parameter("bg", pp.shape[0]);
			p->parameter("convolve", pp.convolve != 0);
		}
		else if (pp.type == PROFIT_AMD_PSF) {
			p = m.add_profile("psf");
			p->parameter("xcen", pp.xcen);
			p->parameter("ycen", pp.ycen);
			p->parameter("mag", pp.mag);
			p->parameter("convolve", pp.convolve != 0);
		}
		else if (pp.type == PROFIT_AMD_NULL) {
			p = m.add_profile("null");
		}
		else {
			throw profit::invalid_parameter("Unknown profile type " + std::to_string(pp.type));
		}
	}
}

}  // namespace

struct profit_amd_engine {
	explicit profit_amd_engine(int device) : eng(device) {}
	profit_amd::Engine eng;
};

extern "C" {

const char *profit_amd_last_error(void) { return g_error.c_str(); }

const char *profit_amd_version(void) { return "profit_amd 0.1 (libprofit 1.9.3 model-image path)"; }

int profit_amd_profile_defaults(int32_t type, profit_amd_profile *p)
{
	return guarded([&] {
		if (!p) throw std::invalid_argument("null profile pointer");
		profit_amd::ProfileParams q;
		profit_amd::default_params(type, &q);
		std::memcpy(p, &q, sizeof(q));
	});
}

int profit_amd_make_model(const profit_amd_model_desc *model, const profit_amd_profile *profiles, int32_t nprofiles,
                          double *out, size_t out_cap, uint32_t *out_w, uint32_t *out_h, int32_t *offset_x,
                          int32_t *offset_y)
{
	return guarded([&] {
		if (!model) throw std::invalid_argument("null model description");
		profit::Model m(model->width, model->height);
		build_model(m, model, profiles, nprofiles);
		if (!m.has_profiles()) throw profit::invalid_parameter("No valid profiles found in profiles list");
		profit::Image img;
		profit::Point off;
		m.evaluate(img, off);
		if (img.size() > out_cap) throw std::invalid_argument("output buffer too small");
		if (img.size()) std::memcpy(out, img.data(), img.size() * sizeof(double));
		if (out_w) *out_w = img.getWidth();
		if (out_h) *out_h = img.getHeight();
		if (offset_x) *offset_x = (int32_t)off.x;
		if (offset_y) *offset_y = (int32_t)off.y;
	});
}

int profit_amd_drawing_dims(const profit_amd_model_desc *model, const profit_amd_profile *profiles,
                            int32_t nprofiles, uint32_t *width, uint32_t *height)
{
	return guarded([&] {
		profit::Model m(model->width, model->height);
		build_model(m, model, profiles, nprofiles);
		const profit::Dimensions d = m.get_drawing_dimensions();
		*width = d.x;
		*height = d.y;
	});
}

int profit_amd_output_dims(const profit_amd_model_desc *model, const profit_amd_profile *profiles, int32_t nprofiles,
                           uint32_t *width, uint32_t *height)
{
	return guarded([&] {
		if (!model) throw std::invalid_argument("null model description");
		profit::Model m(model->width, model->height);
		build_model(m, model, profiles, nprofiles);
		const profit::Dimensions d = m.get_output_dimensions();
		*width = d.x;
		*height = d.y;
	});
}

int profit_amd_clear_cache(void)
{
	return guarded([&] { profit::clear_cache(); });
}

int profit_amd_adjust_mask(const uint8_t *mask, uint32_t mask_w, uint32_t mask_h, uint32_t img_w, uint32_t img_h,
                           const double *psf, uint32_t psf_w, uint32_t psf_h, uint32_t finesampling, uint8_t *out,
                           size_t out_cap, uint32_t *out_w, uint32_t *out_h)
{
	return guarded([&] {
		profit::Mask m(mask, mask_w, mask_h);
		profit::Image p;
		if (psf && psf_w && psf_h) p = profit::Image(std::vector<double>(psf, psf + (size_t)psf_w * psf_h), psf_w, psf_h);
		profit::Model::adjust(m, {img_w, img_h}, p, finesampling);
		if (m.size() > out_cap) throw std::invalid_argument("output buffer too small");
		for (size_t k = 0; k < m.size(); ++k) out[k] = m[k] ? 1 : 0;
		*out_w = m.getWidth();
		*out_h = m.getHeight();
	});
}

int profit_amd_convolve(int32_t convolver, const double *src, uint32_t src_w, uint32_t src_h, const double *krn,
                        uint32_t krn_w, uint32_t krn_h, const uint8_t *mask, double *out, int32_t device)
{
	return guarded([&] {
		profit::ConvolverCreationPreferences prefs;
		prefs.device = device;
		profit::ConvolverPtr c;
		switch (convolver) {
		case PROFIT_AMD_CONV_BRUTE_OLD: c = profit::create_convolver(profit::BRUTE_OLD, prefs); break;
		case PROFIT_AMD_CONV_BRUTE: c = profit::create_convolver(profit::BRUTE, prefs); break;
		case PROFIT_AMD_CONV_FFT: c = profit::create_convolver(profit::FFT, prefs); break;
		case PROFIT_AMD_CONV_NULL: c = profit::create_convolver(profit::NO_OP, prefs); break;
		case PROFIT_AMD_CONV_BRUTE_BF16X3: c = profit::create_convolver(profit::BRUTE_BF16X3, prefs); break;
		default: throw profit::invalid_parameter("Unsupported convolver type: " + std::to_string(convolver));
		}
		profit::Image s(std::vector<double>(src, src + (size_t)src_w * src_h), src_w, src_h);
		profit::Image k(std::vector<double>(krn, krn + (size_t)krn_w * krn_h), krn_w, krn_h);
		profit::Mask m;
		if (mask) m = profit::Mask(mask, src_w, src_h);
		profit::Image r = c->convolve(s, k, m);
		std::memcpy(out, r.data(), r.size() * sizeof(double));
	});
}

int profit_amd_convolve_noalloc(int32_t convolver, const double *src, uint32_t src_w, uint32_t src_h, const double *krn,
                                uint32_t krn_w, uint32_t krn_h, const uint8_t *mask, double *out, int32_t device,
                                void *stream)
{
	return guarded([&] {
		int method;
		switch (convolver) {
		case PROFIT_AMD_CONV_BRUTE_OLD:
		case PROFIT_AMD_CONV_BRUTE: method = profit_amd::NOALLOC_BRUTE; break;
		case PROFIT_AMD_CONV_NULL: method = profit_amd::NOALLOC_NULL; break;
		default: throw profit::invalid_parameter("the allocation-free convolution supports brute, brute-old and null");
		}
		if (src_w > (uint32_t)INT32_MAX || src_h > (uint32_t)INT32_MAX)
			throw std::invalid_argument("source dimensions out of range");
		if (method != profit_amd::NOALLOC_NULL &&
		    (src_w > (uint32_t)profit_amd::kNoAllocMaxSide || src_h > (uint32_t)profit_amd::kNoAllocMaxSide))
			throw std::invalid_argument("source too large for the allocation-free convolution: at most " +
			                            std::to_string(profit_amd::kNoAllocMaxSide) + " pixels per side");
		if (krn_w > (uint32_t)profit_amd::kNoAllocMaxKernel || krn_h > (uint32_t)profit_amd::kNoAllocMaxKernel)
			throw std::invalid_argument("kernel dimensions must be 1 to " + std::to_string(profit_amd::kNoAllocMaxKernel));
		profit_amd::convolve_noalloc(method, src, (int)src_w, (int)src_h, krn, (int)krn_w, (int)krn_h, mask, out, device,
		                             static_cast<hipStream_t>(stream));
	});
}

const char *profit_amd_convolvers(void) { return "brute,brute-old,fft,null,brute-bf16x3"; }

// image.cpp:246-367
int profit_amd_downsample(const double *img, uint32_t w, uint32_t h, uint32_t f, int32_t mode, double *out,
                          size_t out_cap, uint32_t *out_w, uint32_t *out_h)
{
	return guarded([&] {
		if (f == 0) throw std::invalid_argument("downsampling factor is 0");
		const uint32_t dw = (w + f - 1) / f, dh = (h + f - 1) / f;
		if ((size_t)dw * dh > out_cap) throw std::invalid_argument("output buffer too small");
		std::memset(out, 0, sizeof(double) * (size_t)dw * dh);
		if (mode == 0) {
			for (uint32_t rd = 0; rd < dh; ++rd)
				for (uint32_t cd = 0; cd < dw; ++cd) out[cd + (size_t)rd * dw] = img[cd * f + (size_t)(rd * f) * w];
		}
		else if (mode == 1) {
			for (uint32_t r = 0; r < h; ++r)
				for (uint32_t c = 0; c < w; ++c) out[c / f + (size_t)(r / f) * dw] += img[c + (size_t)r * w];
		}
		else {
			for (uint32_t rd = 0; rd < dh; ++rd)
				for (uint32_t cd = 0; cd < dw; ++cd) {
					double t = 0;
					unsigned cnt = 0;
					for (uint32_t r = rd * f; r < std::min(rd * f + f, h); ++r)
						for (uint32_t c = cd * f; c < std::min(cd * f + f, w); ++c) {
							t += img[c + (size_t)r * w];
							++cnt;
						}
					out[cd + (size_t)rd * dw] = t / cnt;
				}
		}
		*out_w = dw;
		*out_h = dh;
	});
}

int profit_amd_upsample(const double *img, uint32_t w, uint32_t h, uint32_t f, int32_t mode, double *out,
                        size_t out_cap, uint32_t *out_w, uint32_t *out_h)
{
	return guarded([&] {
		if (f == 0) throw std::invalid_argument("upsampling factor is 0");
		const uint32_t uw = w * f, uh = h * f;
		if ((size_t)uw * uh > out_cap) throw std::invalid_argument("output buffer too small");
		const double div = (mode == 1) ? (double)f * f : 1.0;
		for (uint32_t r = 0; r < uh; ++r)
			for (uint32_t c = 0; c < uw; ++c) out[c + (size_t)r * uw] = img[c / f + (size_t)(r / f) * w] / div;
		*out_w = uw;
		*out_h = uh;
	});
}

int profit_amd_engine_create(int32_t device, profit_amd_engine **engine)
{
	return guarded([&] {
		if (!engine) throw std::invalid_argument("null engine pointer");
		*engine = new profit_amd_engine(device);
	});
}

void profit_amd_engine_destroy(profit_amd_engine *engine) { delete engine; }

int profit_amd_engine_add_galaxy(profit_amd_engine *engine, const profit_amd_galaxy_desc *g, int32_t *galaxy_id)
{
	return guarded([&] {
		if (!engine || !g) throw std::invalid_argument("null argument");
		profit_amd::GalaxySpec s;
		s.W = (int)g->width;
		s.H = (int)g->height;
		s.scale_x = g->scale_x;
		s.scale_y = g->scale_y;
		s.magzero = g->magzero;
		s.psf = g->psf;
		s.Kx = (int)g->psf_w;
		s.Ky = (int)g->psf_h;
		s.calc = g->calcregion;
		s.cw = (int)g->calc_w;
		s.ch = (int)g->calc_h;
		s.region = g->region;
		s.psf_scale_x = g->psf_scale_x > 0 ? g->psf_scale_x : 1.0;
		s.psf_scale_y = g->psf_scale_y > 0 ? g->psf_scale_y : 1.0;
		s.data = g->data;
		s.sigma = g->sigma;
		s.likefunc = g->likefunc;
		if (s.likefunc < 0 || s.likefunc > 3) throw profit::invalid_parameter("unknown likelihood function");
		const int id = engine->eng.add_galaxy(s);
		if (galaxy_id) *galaxy_id = id;
	});
}

int profit_amd_engine_set_batch(profit_amd_engine *engine, const int32_t *inst_galaxy, const int32_t *inst_ncomp,
                                int32_t ninst)
{
	return guarded([&] {
		if (!engine) throw std::invalid_argument("null engine");
		std::vector<int> g(inst_galaxy, inst_galaxy + ninst), n(inst_ncomp, inst_ncomp + ninst);
		engine->eng.set_batch(g, n);
	});
}

int profit_amd_engine_set_tile_rows(profit_amd_engine *engine, int32_t rows)
{
	return guarded([&] { engine->eng.set_conv_rows(rows); });
}

int profit_amd_engine_set_render_fp64(profit_amd_engine *engine, int32_t on)
{
	return guarded([&] { engine->eng.set_render_fp64(on); });
}

int profit_amd_engine_set_grad_slots(profit_amd_engine *engine, uint32_t mask)
{
	return guarded([&] {
		if (!engine) throw std::invalid_argument("null engine");
		engine->eng.set_grad_slots(mask);
	});
}

int profit_amd_engine_set_conv_method(profit_amd_engine *engine, int32_t convolver)
{
	return guarded([&] {
		if (convolver == PROFIT_AMD_CONV_BRUTE) engine->eng.set_conv_method(profit_amd::CK_VALU);
		else if (convolver == PROFIT_AMD_CONV_FFT) engine->eng.set_conv_method(profit_amd::CK_FFT);
		else if (convolver == PROFIT_AMD_CONV_BRUTE_BF16X3) engine->eng.set_conv_method(profit_amd::CK_MFMA);
		else
			throw std::invalid_argument(
			    "the engine convolves with PROFIT_AMD_CONV_BRUTE, PROFIT_AMD_CONV_FFT or PROFIT_AMD_CONV_BRUTE_BF16X3");
	});
}

int profit_amd_engine_autotune(profit_amd_engine *engine, const profit_amd_profile *params, int32_t reps,
                               int32_t flags, int32_t *kinds, double *ms)
{
	return guarded([&] {
		const profit_amd::Engine::TuneReport r = engine->eng.autotune(
		    reinterpret_cast<const profit_amd::ProfileParams *>(params), reps, (flags & PROFIT_AMD_TUNE_ALLOW_BF16X3) != 0);
		if (kinds)
			for (size_t g = 0; g < r.kinds.size(); ++g)
				kinds[g] = r.kinds[g] == profit_amd::CK_FFT    ? PROFIT_AMD_CONV_FFT
				           : r.kinds[g] == profit_amd::CK_MFMA ? PROFIT_AMD_CONV_BRUTE_BF16X3
				                                               : PROFIT_AMD_CONV_BRUTE;
		if (ms)
			for (int k = 0; k < 6; ++k) ms[k] = r.ms[k];
	});
}

int profit_amd_engine_like(profit_amd_engine *engine, const profit_amd_profile *params, double *loglike)
{
	return guarded([&] {
		engine->eng.evaluate(reinterpret_cast<const profit_amd::ProfileParams *>(params), loglike, nullptr);
	});
}

int profit_amd_engine_like_grad(profit_amd_engine *engine, const profit_amd_profile *params, double *loglike,
                                double *grad)
{
	return guarded([&] {
		if (!grad) throw std::invalid_argument("null gradient buffer");
		engine->eng.evaluate(reinterpret_cast<const profit_amd::ProfileParams *>(params), loglike, grad);
	});
}

int profit_amd_engine_model_images(profit_amd_engine *engine, const profit_amd_profile *params, float *out)
{
	return guarded([&] {
		engine->eng.model_images(reinterpret_cast<const profit_amd::ProfileParams *>(params), out);
	});
}

int profit_amd_engine_diagnostics(profit_amd_engine *engine, const profit_amd_profile *params, int32_t *cls,
                                  int32_t *nodes)
{
	return guarded([&] {
		engine->eng.diagnostics(reinterpret_cast<const profit_amd::ProfileParams *>(params), cls, nodes);
	});
}

int profit_amd_engine_galaxy_geometry(profit_amd_engine *engine, int32_t galaxy_id, uint32_t *draw_w,
                                      uint32_t *draw_h, uint32_t *pad_x, uint32_t *pad_y)
{
	return guarded([&] {
		const profit_amd::GalaxyDev &g = engine->eng.galaxy(galaxy_id);
		*draw_w = (uint32_t)g.DW;
		*draw_h = (uint32_t)g.DH;
		*pad_x = (uint32_t)g.padx;
		*pad_y = (uint32_t)g.pady;
	});
}

int profit_amd_engine_set_timing(profit_amd_engine *engine, int32_t on)
{
	return guarded([&] { engine->eng.set_timing(on != 0); });
}

int profit_amd_engine_stats(profit_amd_engine *engine, double *stage_ms, int64_t *stage_calls, int64_t *evaluations,
                            int32_t *queue_length, double *conv_flops, double *adjoint_flops)
{
	return guarded([&] {
		const profit_amd::EngineStats &s = engine->eng.stats();
		for (int k = 0; k < profit_amd::ST_COUNT; ++k) {
			if (stage_ms) stage_ms[k] = s.ms[k];
			if (stage_calls) stage_calls[k] = s.calls[k];
		}
		if (evaluations) *evaluations = s.evaluations;
		if (queue_length) *queue_length = s.queue_count;
		if (conv_flops) *conv_flops = engine->eng.conv_flops_forward();
		if (adjoint_flops) *adjoint_flops = engine->eng.conv_flops_adjoint();
	});
}

int profit_amd_engine_adjusted_settings(profit_amd_engine *engine, const profit_amd_profile *params, int32_t *resolution,
                                        int32_t *max_recursions, double *acc, double *rscale_switch,
                                        double *rscale_max)
{
	return guarded([&] {
		engine->eng.adjusted_settings(reinterpret_cast<const profit_amd::ProfileParams *>(params), resolution,
		                              max_recursions, acc, rscale_switch, rscale_max);
	});
}

int profit_amd_engine_fft_stats(profit_amd_engine *engine, double *kernel_ms, int64_t *calls)
{
	return guarded([&] {
		const profit_amd::EngineStats &s = engine->eng.stats();
		for (int f = 0; f < 2; ++f) {
			for (int k = 0; k < 3; ++k)
				if (kernel_ms) kernel_ms[3 * f + k] = s.fft_ms[f][k];
			if (calls) calls[f] = s.fft_calls[f];
		}
	});
}

int profit_amd_engine_reset_stats(profit_amd_engine *engine)
{
	return guarded([&] { engine->eng.reset_stats(); });
}

}  // extern "C"
