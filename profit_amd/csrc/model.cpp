// libprofit-compatible Model / Profile / Convolver on top of the GPU engine.
//
// Host orchestration follows src/libprofit/src/profit/model.cpp:103-380
// (add_profile registry, analyze_inputs, analyze_expansion_requirements, mask
// adjustment, evaluate, produce_image) and convolve.cpp:55-99, 526-594; the
// per-pixel work (rendering, sub-pixel recursion, convolution, dilation) runs
// in the HIP kernels.
#include "profit.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <list>
#include <mutex>
#include <sstream>

#include "engine.h"

namespace profit {

Point Convolver::NO_OFFSET;
Point Model::NO_OFFSET;

// ------------------------------------------------------------------ Image / Mask

Image::Image(const std::vector<double> &data, unsigned int w, unsigned int h) : dims_(w, h), data_(data)
{
	if (data_.size() != (size_t)w * h) throw std::invalid_argument("data size != width * height");
}

void Image::zero() { std::fill(data_.begin(), data_.end(), 0.0); }

double Image::total() const
{
	double s = 0;
	for (double v : data_) s += v;
	return s;
}

Image Image::normalize() const
{
	Image out(*this);
	const double s = total();
	if (s > 0)
		for (double &v : out.data_) v /= s;
	return out;
}

Image Image::crop(Dimensions d, Point start) const
{
	if (d.x + start.x > dims_.x || d.y + start.y > dims_.y) throw std::invalid_argument("crop exceeds image");
	Image out(d);
	for (unsigned j = 0; j < d.y; ++j)
		for (unsigned i = 0; i < d.x; ++i) out[i + (size_t)j * d.x] = data_[(i + start.x) + (size_t)(j + start.y) * dims_.x];
	return out;
}

Image Image::extend(Dimensions d, Point start) const
{
	if (dims_.x + start.x > d.x || dims_.y + start.y > d.y) throw std::invalid_argument("extension not possible");
	Image out(d);
	for (unsigned j = 0; j < dims_.y; ++j)
		for (unsigned i = 0; i < dims_.x; ++i) out[(i + start.x) + (size_t)(j + start.y) * d.x] = data_[i + (size_t)j * dims_.x];
	return out;
}

// image.cpp:283-296 (SUM mode)
Image Image::downsample_sum(unsigned int f) const
{
	if (f == 0) throw std::invalid_argument("downsampling factor is 0");
	if (f == 1) return *this;
	Dimensions dd{(dims_.x + f - 1) / f, (dims_.y + f - 1) / f};
	Image out(dd);
	for (unsigned row = 0; row < dims_.y; ++row)
		for (unsigned col = 0; col < dims_.x; ++col) out[col / f + (size_t)(row / f) * dd.x] += data_[col + (size_t)row * dims_.x];
	return out;
}

Mask::Mask(const std::vector<bool> &data, unsigned int w, unsigned int h) : dims_(w, h), data_(data.size())
{
	if (data.size() != (size_t)w * h) throw std::invalid_argument("data size != width * height");
	for (size_t k = 0; k < data.size(); ++k) data_[k] = data[k] ? 1 : 0;
}

Mask::Mask(const uint8_t *data, unsigned int w, unsigned int h) : dims_(w, h), data_((size_t)w * h)
{
	for (size_t k = 0; k < data_.size(); ++k) data_[k] = data[k] ? 1 : 0;
}

// image.h:362-392
std::pair<Point, Point> Mask::bounding_box() const
{
	Point lb = dims_, ub{0, 0};
	bool only_zeros = true;
	for (unsigned j = 0; j < dims_.y; ++j)
		for (unsigned i = 0; i < dims_.x; ++i) {
			if (!data_[i + (size_t)j * dims_.x]) continue;
			lb.x = std::min(lb.x, i);
			lb.y = std::min(lb.y, j);
			ub.x = std::max(ub.x, i + 1);
			ub.y = std::max(ub.y, j + 1);
			only_zeros = false;
		}
	if (only_zeros) return {Point{0, 0}, Point{0, 0}};
	return {lb, ub};
}

Mask Mask::upsample(unsigned int f) const
{
	Mask out(dims_.x * f, dims_.y * f);
	for (unsigned r = 0; r < out.dims_.y; ++r)
		for (unsigned c = 0; c < out.dims_.x; ++c) out.data_[c + (size_t)r * out.dims_.x] = data_[c / f + (size_t)(r / f) * dims_.x];
	return out;
}

Mask Mask::extend(Dimensions d, Point start) const
{
	if (dims_.x + start.x > d.x || dims_.y + start.y > d.y) throw std::invalid_argument("extension not possible");
	Mask out(d.x, d.y);
	for (unsigned j = 0; j < dims_.y; ++j)
		for (unsigned i = 0; i < dims_.x; ++i)
			out.data_[(i + start.x) + (size_t)(j + start.y) * d.x] = data_[i + (size_t)j * dims_.x];
	return out;
}

Mask Mask::expand_by(Dimensions pad) const
{
	Mask out(dims_.x, dims_.y);
	if (data_.empty()) return out;
	profit_amd::Engine::dilate(data_.data(), out.data_.data(), (int)dims_.x, (int)dims_.y, (int)pad.x, (int)pad.y);
	return out;
}

// ------------------------------------------------------------------ Profile

static const char *kShapeNames[6][5] = {{"re", "nser", nullptr, nullptr, nullptr},
                                        {"fwhm", "con", nullptr, nullptr, nullptr},
                                        {"rout", "a", "b", nullptr, nullptr},
                                        {"rc", "rt", "a", nullptr, nullptr},
                                        {"re", "rb", "nser", "a", "b"},
                                        {"h1", "h2", "rb", "a", nullptr}};

Profile::Profile(const std::string &name, int kind, int type) : name_(name), kind_(kind)
{
	profit_amd::default_params(type, &params_);
}

bool Profile::set_double(const std::string &n, double v)
{
	if (kind_ == SKY) {
		// sky.cpp:65-71
		if (n == "bg") {
			params_.shape[0] = v;
			return true;
		}
		return false;
	}
	if (kind_ == POINT) {
		// psf.cpp:147-156
		if (n == "xcen") params_.xcen = v;
		else if (n == "ycen") params_.ycen = v;
		else if (n == "mag") params_.mag = v;
		else return false;
		return true;
	}
	if (kind_ != RADIAL) return false;
	auto &p = params_;
	if (n == "xcen") p.xcen = v;
	else if (n == "ycen") p.ycen = v;
	else if (n == "mag") p.mag = v;
	else if (n == "ang") p.ang = v;
	else if (n == "axrat") p.axrat = v;
	else if (n == "box") p.box = v;
	else if (n == "acc") p.acc = v;
	else if (n == "rscale_switch") p.rscale_switch = v;
	else if (n == "rscale_max") p.rscale_max = v;
	else {
		for (int k = 0; k < 5; ++k) {
			const char *sn = kShapeNames[p.type][k];
			if (sn && n == sn) {
				p.shape[k] = v;
				return true;
			}
		}
		return false;
	}
	return true;
}

bool Profile::set_uint(const std::string &n, unsigned int v)
{
	if (kind_ != RADIAL) return false;
	if (n == "resolution") params_.resolution = v;
	else if (n == "max_recursions") params_.max_recursions = v;
	else return false;
	return true;
}

bool Profile::set_bool(const std::string &n, bool v)
{
	if (n == "convolve") {
		convolve_ = v;
		params_.convolve = v;
		return true;
	}
	if (kind_ != RADIAL) return false;
	if (n == "rough") params_.rough = v;
	else if (n == "adjust") params_.adjust = v;
	else if (n == "force_cpu") { /* accepted for compatibility: there is no CPU path */ }
	else if (n == "rescale_flux" && params_.type == profit_amd::SERSIC) params_.rescale_flux = v;
	else return false;
	return true;
}

// profile.cpp:78-158
void Profile::parameter(const std::string &n, double v)
{
	if (!set_double(n, v)) throw invalid_parameter("Unknown double parameter in profile " + name_ + ": " + n);
}
void Profile::parameter(const std::string &n, unsigned int v)
{
	if (!set_uint(n, v)) throw invalid_parameter("Unknown unsigned int parameter in profile " + name_ + ": " + n);
}
void Profile::parameter(const std::string &n, bool v)
{
	if (!set_bool(n, v)) throw invalid_parameter("Unknown bool parameter in profile " + name_ + ": " + n);
}

static std::string trim(const std::string &s)
{
	const auto b = s.find_first_not_of(" \t\n\r");
	if (b == std::string::npos) return "";
	const auto e = s.find_last_not_of(" \t\n\r");
	return s.substr(b, e - b + 1);
}

void Profile::parameter(const std::string &spec)
{
	const auto eq = spec.find('=');
	if (eq == std::string::npos || spec.find('=', eq + 1) != std::string::npos)
		throw invalid_parameter("missing = in parameter: " + spec);
	const std::string n = trim(spec.substr(0, eq)), val = trim(spec.substr(eq + 1));
	auto parse_error = [&](const char *tn) {
		std::ostringstream os;
		os << "Parameter " << n << " in profile " << name_ << " is of type " << tn
		   << ", but given value cannot be parsed as " << tn << ": " << val;
		return invalid_parameter(os.str());
	};
	// bool, then unsigned int, then double (profile.cpp:147-149)
	static const char *bools[] = {"convolve", "rough", "adjust", "force_cpu", "rescale_flux"};
	for (const char *b : bools)
		if (n == b) {
			unsigned long u;
			try {
				u = std::stoul(val, nullptr, 10);
			} catch (const std::exception &) {
				throw parse_error("bool");
			}
			if (set_bool(n, u != 0)) return;
		}
	if (n == "resolution" || n == "max_recursions") {
		unsigned long u;
		try {
			u = std::stoul(val, nullptr, 10);
		} catch (const std::exception &) {
			throw parse_error("unsigned int");
		}
		if (set_uint(n, (unsigned)u)) return;
	}
	double d = 0;
	bool parsed = true;
	try {
		d = std::stod(val);
	} catch (const std::exception &) {
		parsed = false;
	}
	if (parsed && set_double(n, d)) return;
	if (!parsed) {
		profit_amd::ProfileParams probe = params_;
		if (set_double(n, 0.0)) {
			params_ = probe;
			throw parse_error("double");
		}
	}
	throw unknown_parameter("Profile " + name_ + " doesn't support parameter " + n + ", or parameter has invalid value: " + val);
}

bool Profile::do_convolve() const { return convolve_; }

void Profile::validate() const
{
	if (kind_ != RADIAL) return;
	const int e = profit_amd::validate_params(params_);
	if (e != profit_amd::SE_OK) throw invalid_parameter(profit_amd::setup_error_message(e));
}

// ------------------------------------------------------------------ Convolver

// FFTConvolver::offset_after_convolution (convolve.cpp:269-281): where the
// source starts within the 2 * max(src, krn) frame of the result
static Point fft_offset_after_convolution(const Dimensions &sd, const Dimensions &kd)
{
	Point off = sd / 2;
	if (sd.x % 2 == 0 || kd.x % 2 == 0) off.x -= 1;
	if (sd.y % 2 == 0 || kd.y % 2 == 0) off.y -= 1;
	return off;
}

PointPair Convolver::padding(const Dimensions &sd, const Dimensions &kd) const
{
	if (type_ != FFT) return {};
	// FFTConvolver::padding (convolve.cpp:262-267)
	const Dimensions ext = max(sd, kd) * 2;
	const Point off = fft_offset_after_convolution(sd, kd);
	return {off, ext - ext / 2 - off};
}

static int kernel_method(ConvolverType t)
{
	return t == FFT ? profit_amd::CK_FFT : t == BRUTE_BF16X3 ? profit_amd::CK_MFMA : profit_amd::CK_VALU;
}

Image Convolver::convolve_impl(const Image &src, const Image &krn, const Mask &mask, bool crop, Point &offset_out)
{
	if (type_ == NO_OP) return src;  // NullConvolver (convolve.cpp:521-524)
	const Dimensions sd = src.getDimensions(), kd = krn.getDimensions();
	if (type_ != FFT || crop) {
		Image out(sd);
		profit_amd::convolve_image(src.data(), (int)sd.x, (int)sd.y, krn.data(), (int)kd.x, (int)kd.y,
		                           mask ? mask.bytes() : nullptr, out.data(), device_, kernel_method(type_));
		return out;
	}
	// FFTConvolver with crop = false (convolve.cpp:283-311): the whole frame of
	// 2 * max(src, krn), holding the linear convolution with the kernel placed at
	// (src - krn) / 2, masked by the mask extended to the frame at the offset.
	// The frame is the GPU convolution of the source placed at s = k0 + c in a
	// zero image (c = kw - 1 - kw/2, the centre the same-size sum uses), whose
	// same-size result at X is sum_i src(i) krn(X - k0 - i).
	const Dimensions ext = max(sd, kd) * 2;
	if (!(ext >= sd)) throw std::invalid_argument("extension not possible");
	const Point k0 = (sd - kd) / 2;
	const Point s{k0.x + (kd.x - 1 - kd.x / 2), k0.y + (kd.y - 1 - kd.y / 2)};
	const Dimensions work{std::max(ext.x, s.x + sd.x), std::max(ext.y, s.y + sd.y)};
	Image placed = src.extend(work, s);
	Image conv(work);
	profit_amd::convolve_image(placed.data(), (int)work.x, (int)work.y, krn.data(), (int)kd.x, (int)kd.y, nullptr,
	                           conv.data(), device_, profit_amd::CK_FFT);
	Image out = conv.crop(ext, {0, 0});
	const Point off = fft_offset_after_convolution(sd, kd);
	if (mask) {
		const Mask em = mask.extend(ext, off);
		for (size_t k = 0; k < out.size(); ++k)
			if (!em[k]) out[k] = 0;
	}
	if (&offset_out != &NO_OFFSET) offset_out = off;
	return out;
}

Image Convolver::convolve(const Image &src, const Image &krn, const Mask &mask, bool crop, Point &offset_out)
{
	const Dimensions sd = src.getDimensions(), kd = krn.getDimensions();
	if (!(sd >= kd)) {
		// convolve.cpp:58-74: extend to max(src, krn), whose x extent serves both axes
		const Dimensions nd = max(sd, kd);
		if (nd.x < sd.x || nd.y < sd.y) throw std::invalid_argument("extension not possible");
		const Dimensions diff = nd - sd;
		Image es = src.extend(nd, {0, 0});
		Mask em;
		if (mask) em = mask.extend(nd, {0, 0});
		Image r = convolve_impl(es, krn, em, crop, offset_out);
		return r.crop(r.getDimensions() - diff, {0, 0});
	}
	return convolve_impl(src, krn, mask, crop, offset_out);
}

ConvolverPtr create_convolver(ConvolverType type, const ConvolverCreationPreferences &prefs)
{
	switch (type) {
	case BRUTE_OLD:
	case BRUTE:
	case FFT:
	case BRUTE_BF16X3:
	case NO_OP: return std::make_shared<Convolver>(type, prefs.device);
	default: throw invalid_parameter("Unsupported convolver type: " + std::to_string((int)type));
	}
}

ConvolverPtr create_convolver(const std::string &type, const ConvolverCreationPreferences &prefs)
{
	if (type == "brute-old") return create_convolver(BRUTE_OLD, prefs);
	if (type == "brute") return create_convolver(BRUTE, prefs);
	if (type == "fft") return create_convolver(FFT, prefs);
	if (type == "null") return create_convolver(NO_OP, prefs);
	if (type == "brute-bf16x3") return create_convolver(BRUTE_BF16X3, prefs);
	throw invalid_parameter("Convolver of type " + type + " is not supported");
}

// ------------------------------------------------------------------ Model

ProfilePtr Model::add_profile(const std::string &n)
{
	ProfilePtr p;
	if (n == "null") p = std::make_shared<Profile>(n, Profile::NULLPROF, profit_amd::NULLPROF);
	else if (n == "sky") p = std::make_shared<Profile>(n, Profile::SKY, profit_amd::SKY);
	else if (n == "psf") p = std::make_shared<Profile>(n, Profile::POINT, profit_amd::POINTSRC);
	else if (n == "sersic") p = std::make_shared<Profile>(n, Profile::RADIAL, profit_amd::SERSIC);
	else if (n == "moffat") p = std::make_shared<Profile>(n, Profile::RADIAL, profit_amd::MOFFAT);
	else if (n == "ferrer" || n == "ferrers") p = std::make_shared<Profile>(n, Profile::RADIAL, profit_amd::FERRER);
	else if (n == "coresersic") p = std::make_shared<Profile>(n, Profile::RADIAL, profit_amd::CORESERSIC);
	else if (n == "king") p = std::make_shared<Profile>(n, Profile::RADIAL, profit_amd::KING);
	else if (n == "brokenexp") p = std::make_shared<Profile>(n, Profile::RADIAL, profit_amd::BROKENEXP);
	else throw invalid_parameter("Unknown profile name: " + n);
	profiles_.push_back(p);
	return p;
}

// model.cpp:152-184
void Model::analyze_expansion(const Dimensions &dims, const Mask &mask, const Image &psf, unsigned int fs, Analysis &a,
                              bool adjust_mask)
{
	a.mask_needs_convolution = false;
	a.mask_needs_psf_padding = false;
	bool model_needs_psf_padding = a.convolution_required;
	if (mask && !adjust_mask) {
		model_needs_psf_padding = mask.getDimensions() > dims * fs;
	}
	else if (mask && a.convolution_required) {
		auto bounds = mask.bounding_box();
		const Point low = bounds.first * fs;
		const Point up = mask.getDimensions() * fs - bounds.second * fs;
		const Dimensions needed = psf.getDimensions() / 2;
		model_needs_psf_padding = !(low >= needed) || !(up >= needed);
		a.mask_needs_convolution = true;
		a.mask_needs_psf_padding = model_needs_psf_padding;
	}
	a.psf_padding = model_needs_psf_padding ? psf.getDimensions() / 2 : Dimensions{0, 0};
	a.mask_needs_adjustment = mask && (a.mask_needs_convolution || a.mask_needs_psf_padding || fs > 1);
	a.drawing_dims = dims * fs + a.psf_padding * 2;
}

Model::Analysis Model::analyze_inputs() const
{
	if (!requested_dimensions_) throw invalid_parameter("Model's requested dimensions are 0");
	if (scale_.x <= 0) throw invalid_parameter("Model's scale_x cannot be negative or zero");
	if (scale_.y <= 0) throw invalid_parameter("Model's scale_y cannot be negative or zero");
	if (mask_ && adjust_mask_ && mask_.getDimensions() != requested_dimensions_)
		throw invalid_parameter("Mask dimensions != model dimensions");
	Analysis a;
	a.convolution_required = std::any_of(profiles_.begin(), profiles_.end(), [](const ProfilePtr &p) { return p->do_convolve(); });
	if (a.convolution_required && !psf_) throw invalid_parameter("No psf provided but profile(s) requested convolution");
	for (auto &p : profiles_) {
		p->validate();
		// PsfProfile::validate (psf.cpp:40-45)
		if (p->kind() == Profile::POINT && !psf_)
			throw invalid_parameter("No psf present in the model, cannot produce a psf profile");
	}
	analyze_expansion(requested_dimensions_, mask_, psf_, finesampling_, a, adjust_mask_);
	return a;
}

Dimensions Model::get_drawing_dimensions() const { return analyze_inputs().drawing_dims; }

// The dimensions of the image evaluate() returns: the drawing dimensions,
// then as evaluate changes them (model.cpp:282-329): the FFT convolver's
// uncropped frame, the crop of the PSF padding and the downsampling.
Dimensions Model::get_output_dimensions() const
{
	const Analysis a = analyze_inputs();
	Dimensions d = a.drawing_dims;
	if (dry_run_) return d;
	const ConvolverPtr convolver = convolver_ ? convolver_ : create_convolver(BRUTE);
	const Dimensions kd = psf_.getDimensions();
	if (a.convolution_required && convolver->type() == FFT && !crop_) {
		Dimensions src = d, diff{0, 0};
		if (!(d >= kd)) {
			src = max(d, kd);
			diff = src - d;
		}
		d = max(src, kd) * 2 - diff;
	}
	if (a.psf_padding) {
		Dimensions crop_dims = a.drawing_dims - a.psf_padding * 2;
		if (!crop_) {
			const PointPair actual = convolver->padding(a.drawing_dims, kd);
			const PointPair intended = convolver->padding(a.drawing_dims - a.psf_padding * 2, kd);
			crop_dims = d - a.psf_padding * 2;
			crop_dims -= (actual.first - intended.first) + (actual.second - intended.second);
		}
		d = crop_dims;
	}
	if (finesampling_ > 1 && !return_finesampled_)
		d = Dimensions{(d.x + finesampling_ - 1) / finesampling_, (d.y + finesampling_ - 1) / finesampling_};
	return d;
}

// model.cpp:230-246
void Model::adjust_impl(Mask &mask, const Image &psf, unsigned int fs, const Analysis &a)
{
	if (!mask) return;
	if (fs > 1) mask = mask.upsample(fs);
	if (a.mask_needs_psf_padding) mask = mask.extend(a.drawing_dims, a.psf_padding);
	if (a.mask_needs_convolution) mask = mask.expand_by(psf.getDimensions() / 2);
}

void Model::adjust(Mask &mask, const Dimensions &dims, const Image &psf, unsigned int fs)
{
	Analysis a;
	a.convolution_required = psf.size() > 0;
	analyze_expansion(dims, mask, psf, fs, a, true);
	if (a.mask_needs_adjustment) adjust_impl(mask, psf, fs, a);
}

Image Model::evaluate(Point &offset_out)
{
	Image image;
	evaluate(image, offset_out);
	return image;
}

// ------------------------------------------------------------------ engine cache
//
// A Model's drawing set-up (dimensions, PSF, calc region, convolution kernel)
// becomes one engine galaxy: its device buffers, tile lists, FFT spectrum and
// work lists.  Repeated evaluations with the same set-up (the parameters
// change, the geometry does not) reuse that engine instead of building it
// again; a few set-ups are kept, the least recently used dropped first.
// clear_cache() releases them (R_profit_clear_cache, src/r_profit.cpp:822-825).

namespace {

struct CachedEngine {
	uint64_t key = 0;
	std::unique_ptr<profit_amd::Engine> engine;
	int ncomp = -1;
};

std::mutex g_cache_mu;
std::list<CachedEngine> g_cache;  // most recently used first
constexpr size_t kCacheEntries = 4;

struct Fnv {
	uint64_t h = 1469598103934665603ull;
	void add(const void *p, size_t n)
	{
		const unsigned char *b = static_cast<const unsigned char *>(p);
		for (size_t k = 0; k < n; ++k) h = (h ^ b[k]) * 1099511628211ull;
	}
	template <class T> void add(const T &v) { add(&v, sizeof(T)); }
};

uint64_t spec_key(const profit_amd::GalaxySpec &s, int method, int device)
{
	Fnv f;
	f.add(s.W);
	f.add(s.H);
	f.add(s.scale_x);
	f.add(s.scale_y);
	f.add(s.magzero);
	f.add(s.finesampling);
	f.add(s.psf_convolves);
	f.add(s.psf_scale_x);
	f.add(s.psf_scale_y);
	f.add(s.Kx);
	f.add(s.Ky);
	if (s.psf) f.add(s.psf, sizeof(double) * (size_t)s.Kx * s.Ky);
	f.add(s.cw);
	f.add(s.ch);
	f.add(s.calc != nullptr);
	if (s.calc) f.add(s.calc, (size_t)s.cw * s.ch);
	f.add(method);
	f.add(device);
	return f.h;
}

// the cached engine of a set-up (built on first use), moved to the front;
// the caller holds g_cache_mu
profit_amd::Engine &cached_engine(const profit_amd::GalaxySpec &gs, int method, int device, int ncomp)
{
	const uint64_t key = spec_key(gs, method, device);
	for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
		if (it->key == key) {
			g_cache.splice(g_cache.begin(), g_cache, it);
			break;
		}
	if (g_cache.empty() || g_cache.front().key != key) {
		CachedEngine ce;
		ce.key = key;
		ce.engine.reset(new profit_amd::Engine(device));
		ce.engine->set_conv_method(method);
		ce.engine->add_galaxy(gs);
		g_cache.push_front(std::move(ce));
		while (g_cache.size() > kCacheEntries) g_cache.pop_back();
	}
	CachedEngine &ce = g_cache.front();
	if (ce.ncomp != ncomp) {
		ce.engine->set_batch({0}, {ncomp});
		ce.ncomp = ncomp;
	}
	return *ce.engine;
}

}  // namespace

void clear_cache()
{
	std::lock_guard<std::mutex> lock(g_cache_mu);
	g_cache.clear();
}

std::map<std::string, std::shared_ptr<ProfileStats>> Model::get_stats() const
{
	std::map<std::string, std::shared_ptr<ProfileStats>> out;
	for (auto &p : profiles_) out[p->get_name()] = p->stats();
	return out;
}

// model.cpp:271-380
void Model::evaluate(Image &image, Point &offset_out)
{
	const Analysis a = analyze_inputs();
	if (image.getDimensions() == a.drawing_dims) image.zero();
	else image = Image(a.drawing_dims);
	if (dry_run_) {
		if (&offset_out != &NO_OFFSET) offset_out = Point{0, 0};
		return;
	}
	Mask adjusted;
	const Mask *draw_mask = &mask_;
	if (adjust_mask_ && a.mask_needs_adjustment) {
		adjusted = mask_;
		adjust_impl(adjusted, psf_, finesampling_, a);
		draw_mask = &adjusted;
	}
	else if (!adjust_mask_ && mask_.getDimensions() != a.drawing_dims) {
		std::ostringstream os;
		os << "Mask dimensions != drawing dimensions: [" << mask_.getWidth() << ", " << mask_.getHeight() << "] != ["
		   << a.drawing_dims.x << ", " << a.drawing_dims.y << "]";
		throw invalid_parameter(os.str());
	}
	const bool has_mask = (bool)*draw_mask;

	// every profile but "null" is drawn by the engine: radial profiles, the sky
	// (sky.cpp:47-63) and point sources (psf.cpp:58-145), each into the
	// convolved or the unconvolved canvas by its convolve flag
	std::vector<profit_amd::ProfileParams> drawn;
	bool any_point = false;
	for (auto &p : profiles_) {
		if (p->kind() == Profile::NULLPROF) continue;
		drawn.push_back(p->params());
		any_point |= p->kind() == Profile::POINT;
	}
	// ensure_convolver (model.cpp:144-150): brute force unless one was set
	const ConvolverPtr convolver = convolver_ ? convolver_ : create_convolver(BRUTE);
	const ConvolverType ctype = convolver->type();
	// produce_image step by step (the two canvases come back unconvolved, then
	// the convolver runs) where its result is not the engine's fused
	// convolution: the null convolver, and the FFT convolver's uncropped frame
	const bool literal = a.convolution_required && (ctype == NO_OP || (ctype == FFT && !crop_));
	Point offset{0, 0};
	RadialProfileStats times;
	if (!drawn.empty() || literal) {
		profit_amd::GalaxySpec gs;
		gs.W = (int)(requested_dimensions_.x * finesampling_);
		gs.H = (int)(requested_dimensions_.y * finesampling_);
		gs.scale_x = scale_.x;
		gs.scale_y = scale_.y;
		gs.magzero = magzero_;
		gs.finesampling = (int)finesampling_;
		if (a.convolution_required || any_point) {
			gs.psf = psf_.data();
			gs.Kx = (int)psf_.getWidth();
			gs.Ky = (int)psf_.getHeight();
			gs.psf_convolves = a.convolution_required;
		}
		gs.psf_scale_x = psf_scale_.x;
		gs.psf_scale_y = psf_scale_.y;
		if (has_mask) {
			gs.calc = draw_mask->bytes();
			gs.cw = (int)draw_mask->getWidth();
			gs.ch = (int)draw_mask->getHeight();
		}
		const size_t nd = (size_t)a.drawing_dims.x * a.drawing_dims.y;
		std::vector<float> canvas(nd, 0.f), direct(literal ? nd : 0, 0.f);
		if (!drawn.empty()) {
			const int method = literal ? profit_amd::CK_VALU : kernel_method(ctype);
			std::lock_guard<std::mutex> lock(g_cache_mu);
			profit_amd::Engine &eng = cached_engine(gs, method, device_, (int)drawn.size());
			const profit_amd::GalaxyDev &g = eng.galaxy(0);
			if ((unsigned)g.DW != a.drawing_dims.x || (unsigned)g.DH != a.drawing_dims.y)
				throw std::logic_error("engine drawing dimensions differ from the model analysis");
			const auto t0 = std::chrono::steady_clock::now();
			try {
				eng.model_canvas(drawn.data(), canvas.data(), literal ? direct.data() : nullptr);
			} catch (const std::invalid_argument &e) {
				throw invalid_parameter(e.what());
			}
			times.total = (nsecs_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
			                  std::chrono::steady_clock::now() - t0)
			                  .count();
		}
		if (!literal) {
			for (size_t k = 0; k < nd; ++k) image[k] += canvas[k];
		}
		else {
			// model.cpp:355-377: the unconvolved profiles on the canvas, then
			// convolve, extend the canvas to the result's frame and add
			for (size_t k = 0; k < nd; ++k) image[k] += direct[k];
			Image to_convolve(a.drawing_dims);
			for (size_t k = 0; k < nd; ++k) to_convolve[k] = canvas[k];
			const Image conv = convolver->convolve(to_convolve, psf_, *draw_mask, crop_, offset);
			if (conv.getDimensions() != a.drawing_dims) image = image.extend(conv.getDimensions(), offset);
			for (size_t k = 0; k < image.size(); ++k) image[k] += conv[k];
		}
	}
	for (auto &p : profiles_) p->set_stats(times);

	if (a.psf_padding) {
		// model.cpp:306-324
		Point crop_offset = a.psf_padding;
		Dimensions crop_dims = a.drawing_dims - a.psf_padding * 2;
		if (!crop_) {
			const PointPair actual = convolver->padding(a.drawing_dims, psf_.getDimensions());
			const PointPair intended = convolver->padding(a.drawing_dims - a.psf_padding * 2, psf_.getDimensions());
			const Point offset_diff = actual.first - intended.first;
			const Dimensions dim_diff = actual.second - intended.second;
			crop_dims = image.getDimensions() - a.psf_padding * 2;
			crop_dims -= offset_diff + dim_diff;
			crop_offset += offset_diff;
			offset -= offset_diff;
		}
		image = image.crop(crop_dims, crop_offset);
	}
	if (finesampling_ > 1 && !return_finesampled_) {
		image = image.downsample_sum(finesampling_);
		offset = offset / finesampling_;
	}
	if (adjusted) {
		const size_t n = std::min(image.size(), mask_.size());
		for (size_t k = 0; k < n; ++k)
			if (!mask_[k]) image[k] = 0;
	}
	if (&offset_out != &NO_OFFSET) offset_out = offset;
}

// ------------------------------------------------------------------ library

void init() {}
void finish() {}
bool has_openmp() { return false; }
// libprofit answers whether it was built with FFTW, i.e. whether its FFT
// convolver exists (library.h:86-88); this one always has an FFT convolver
// (its own GPU FFT, not FFTW)
bool has_fftw() { return true; }
bool has_opencl() { return false; }
std::string version() { return "1.9.3"; }

}  // namespace profit
