// libprofit-compatible C++ API backed by the MI355X engine.
//
// Mirrors the names and semantics of libprofit 1.9.3 (paths relative to
// src/libprofit/src/profit/):
//   model.h:65-363         Model (setters, add_profile, evaluate, adjust)
//   profile.h:79-261       Profile parameter registry (parameter(name, value),
//                          parameter("name = value"))
//   convolve.h:68-209      Convolver, ConvolverType, create_convolver
//   image.h:541-788        Image / Mask (row-major, x fastest: i + j*width)
//   exceptions.h:41-105    exception, invalid_parameter, unknown_parameter
//   library.h:67-124       init / finish / has_* queries
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "profile_math.h"

namespace profit {

class exception : public std::exception {
  public:
	explicit exception(const std::string &what) : what_(what) {}
	const char *what() const noexcept override { return what_.c_str(); }

  private:
	std::string what_;
};
class invalid_parameter : public exception {
  public:
	using exception::exception;
};
class unknown_parameter : public exception {
  public:
	using exception::exception;
};
class fft_error : public exception {
  public:
	using exception::exception;
};

struct Dimensions {
	unsigned int x = 0, y = 0;
	Dimensions() = default;
	Dimensions(unsigned int x_, unsigned int y_) : x(x_), y(y_) {}
	explicit operator bool() const { return x > 0 && y > 0; }
	bool operator==(const Dimensions &o) const { return x == o.x && y == o.y; }
	bool operator!=(const Dimensions &o) const { return !(*this == o); }
	bool operator>=(const Dimensions &o) const { return x >= o.x && y >= o.y; }
	bool operator>(const Dimensions &o) const { return x > o.x && y > o.y; }
	Dimensions operator*(unsigned int f) const { return {x * f, y * f}; }
	Dimensions operator/(unsigned int f) const { return {x / f, y / f}; }
	Dimensions operator+(const Dimensions &o) const { return {x + o.x, y + o.y}; }
	Dimensions operator-(const Dimensions &o) const { return {x - o.x, y - o.y}; }
	Dimensions &operator+=(const Dimensions &o) { return *this = *this + o; }
	Dimensions &operator-=(const Dimensions &o) { return *this = *this - o; }
};
// coordinates.h:182-186: libprofit's max() takes the x extent for both axes
inline Dimensions max(const Dimensions &a, const Dimensions &b)
{
	const unsigned int m = a.x > b.x ? a.x : b.x;
	return {m, m};
}
using Point = Dimensions;
using PointPair = std::pair<Point, Point>;

struct PixelScale {
	double x = 1, y = 1;
	PixelScale() = default;
	PixelScale(double x_, double y_) : x(x_), y(y_) {}
};

class Image {
  public:
	Image() = default;
	Image(unsigned int w, unsigned int h) : dims_(w, h), data_((size_t)w * h, 0.0) {}
	explicit Image(Dimensions d) : Image(d.x, d.y) {}
	Image(const std::vector<double> &data, unsigned int w, unsigned int h);
	unsigned int getWidth() const { return dims_.x; }
	unsigned int getHeight() const { return dims_.y; }
	Dimensions getDimensions() const { return dims_; }
	size_t size() const { return data_.size(); }
	bool empty() const { return data_.empty(); }
	explicit operator bool() const { return (bool)dims_; }
	double *data() { return data_.data(); }
	const double *data() const { return data_.data(); }
	double &operator[](size_t i) { return data_[i]; }
	double operator[](size_t i) const { return data_[i]; }
	void zero();
	double total() const;
	Image normalize() const;
	Image crop(Dimensions d, Point start) const;
	Image extend(Dimensions d, Point start) const;
	Image downsample_sum(unsigned int factor) const;

  private:
	Dimensions dims_;
	std::vector<double> data_;
};

class Mask {
  public:
	Mask() = default;
	Mask(unsigned int w, unsigned int h) : dims_(w, h), data_((size_t)w * h, 0) {}
	Mask(const std::vector<bool> &data, unsigned int w, unsigned int h);
	Mask(const uint8_t *data, unsigned int w, unsigned int h);
	unsigned int getWidth() const { return dims_.x; }
	unsigned int getHeight() const { return dims_.y; }
	Dimensions getDimensions() const { return dims_; }
	size_t size() const { return data_.size(); }
	explicit operator bool() const { return (bool)dims_; }
	bool operator[](size_t i) const { return data_[i] != 0; }
	void set(size_t i, bool v) { data_[i] = v ? 1 : 0; }
	const uint8_t *bytes() const { return data_.data(); }
	std::pair<Point, Point> bounding_box() const;
	Mask upsample(unsigned int factor) const;
	Mask extend(Dimensions d, Point start) const;
	Mask expand_by(Dimensions pad) const;  // dilation on the GPU

  private:
	Dimensions dims_;
	std::vector<uint8_t> data_;
};

class Model;

// profile.h:47-74: evaluation statistics.  libprofit fills them per profile;
// here every profile of a Model is drawn in one pass of the kernels, so each
// profile reports that pass's wall time (render, sub-pixel walk and, for the
// profiles that convolve, the convolution).
using nsecs_t = uint64_t;
class ProfileStats {
  public:
	virtual ~ProfileStats() = default;
	nsecs_t total = 0;
};
struct radial_subsampling_stats {
	nsecs_t pre_subsampling = 0, new_subsampling = 0, inital_transform = 0, final_transform = 0, total = 0;
};
struct RadialProfileStats : ProfileStats {
	radial_subsampling_stats subsampling;
	nsecs_t final_image = 0;
};

class Profile {
  public:
	Profile(const std::string &name, int kind, int type);
	virtual ~Profile() = default;
	const std::string &get_name() const { return name_; }
	void parameter(const std::string &name, bool value);
	void parameter(const std::string &name, double value);
	void parameter(const std::string &name, unsigned int value);
	void parameter(const std::string &spec);  // "name = value"
	bool do_convolve() const;
	void validate() const;
	// Profile::get_stats (profile.h:198): the last evaluation's statistics
	std::shared_ptr<ProfileStats> get_stats() const { return stats(); }
	std::shared_ptr<ProfileStats> stats() const { return stats_; }
	void set_stats(const RadialProfileStats &s)
	{
		if (kind_ == RADIAL) stats_ = std::make_shared<RadialProfileStats>(s);
		else {
			stats_ = std::make_shared<ProfileStats>();
			stats_->total = s.total;
		}
	}

	// engine-facing state
	enum Kind { RADIAL = 0, SKY = 1, NULLPROF = 2, POINT = 3 };
	int kind() const { return kind_; }
	const profit_amd::ProfileParams &params() const { return params_; }
	profit_amd::ProfileParams &params() { return params_; }
	double sky_bg() const { return params_.shape[0]; }

  private:
	bool set_double(const std::string &name, double v);
	bool set_uint(const std::string &name, unsigned int v);
	bool set_bool(const std::string &name, bool v);
	std::string name_;
	int kind_;
	profit_amd::ProfileParams params_{};
	bool convolve_ = false;
	std::shared_ptr<ProfileStats> stats_;
};
using ProfilePtr = std::shared_ptr<Profile>;

enum ConvolverType { BRUTE_OLD = 0, BRUTE, OPENCL, FFT, NO_OP, BRUTE_BF16X3 = 16 };

class ConvolverCreationPreferences {
  public:
	ConvolverCreationPreferences() = default;
	Dimensions src_dims, krn_dims;
	unsigned int omp_threads = 1;
	bool reuse_krn_fft = false;
	int device = 0;
};

class Convolver {
  public:
	explicit Convolver(ConvolverType t, int device = 0) : type_(t), device_(device) {}
	virtual ~Convolver() = default;
	static Point NO_OFFSET;
	// Convolver::convolve (convolve.cpp:55-79): src smaller than krn is extended
	// first and cropped back afterwards; masked output pixels are zero.
	Image convolve(const Image &src, const Image &krn, const Mask &mask, bool crop = true,
	               Point &offset_out = NO_OFFSET);
	// Extra padding of an uncropped result (convolve.cpp:50-53, 262-267): none,
	// except for the FFT convolver's 2 * max(src, krn) frame
	virtual PointPair padding(const Dimensions &src_dims, const Dimensions &krn_dims) const;
	ConvolverType type() const { return type_; }

  protected:
	// convolve_impl of each convolver; FFT with crop = false returns the whole
	// 2 * max(src, krn) frame and its offset (convolve.cpp:283-311)
	Image convolve_impl(const Image &src, const Image &krn, const Mask &mask, bool crop, Point &offset_out);
	ConvolverType type_;
	int device_;
};
using ConvolverPtr = std::shared_ptr<Convolver>;

ConvolverPtr create_convolver(ConvolverType type,
                              const ConvolverCreationPreferences &prefs = ConvolverCreationPreferences());
ConvolverPtr create_convolver(const std::string &type,
                              const ConvolverCreationPreferences &prefs = ConvolverCreationPreferences());

class Model {
  public:
	Model() = default;
	Model(unsigned int width, unsigned int height) : requested_dimensions_(width, height) {}
	explicit Model(Dimensions d) : requested_dimensions_(d) {}
	static Point NO_OFFSET;

	ProfilePtr add_profile(const std::string &profile_name);
	bool has_profiles() const { return !profiles_.empty(); }
	Image evaluate(Point &offset_out = NO_OFFSET);
	void evaluate(Image &image, Point &offset_out = NO_OFFSET);
	Dimensions get_drawing_dimensions() const;
	// dimensions of the image evaluate() returns (after the crop, an uncropped
	// FFT frame and the downsampling); no counterpart in libprofit, whose
	// evaluate allocates its result
	Dimensions get_output_dimensions() const;
	// Model::get_stats (model.cpp:382-388): per profile name
	std::map<std::string, std::shared_ptr<ProfileStats>> get_stats() const;
	static void adjust(Mask &mask, const Dimensions &dims, const Image &psf, unsigned int finesampling);

	void set_dimensions(const Dimensions &d) { requested_dimensions_ = d; }
	void set_finesampling(unsigned int fs)
	{
		if (fs == 0) throw std::invalid_argument("finesampling must be > 0");
		finesampling_ = fs;
	}
	void set_psf(const Image &psf) { psf_ = psf.normalize(); }
	void set_image_pixel_scale(const PixelScale &s) { scale_ = s; }
	void set_psf_pixel_scale(const PixelScale &s) { psf_scale_ = s; }
	const PixelScale &get_image_pixel_scale() const { return scale_; }
	void set_magzero(double m) { magzero_ = m; }
	void set_mask(const Mask &m) { mask_ = m; }
	void set_adjust_mask(bool a) { adjust_mask_ = a; }
	void set_convolver(ConvolverPtr c) { convolver_ = std::move(c); }
	void set_crop(bool c) { crop_ = c; }
	void set_dry_run(bool d) { dry_run_ = d; }
	void set_return_finesampled(bool r) { return_finesampled_ = r; }
	void set_omp_threads(unsigned int t) { omp_threads_ = t; }
	void set_device(int d) { device_ = d; }

	struct Analysis {
		bool convolution_required = false, mask_needs_convolution = false, mask_needs_psf_padding = false,
		     mask_needs_adjustment = false;
		Dimensions psf_padding, drawing_dims;
	};

  private:
	Analysis analyze_inputs() const;
	static void analyze_expansion(const Dimensions &dims, const Mask &mask, const Image &psf, unsigned int fs,
	                              Analysis &a, bool adjust_mask);
	static void adjust_impl(Mask &mask, const Image &psf, unsigned int fs, const Analysis &a);

	Dimensions requested_dimensions_;
	unsigned int finesampling_ = 1;
	PixelScale scale_, psf_scale_;
	double magzero_ = 0;
	Image psf_;
	Mask mask_;
	ConvolverPtr convolver_;
	bool adjust_mask_ = true, crop_ = true, dry_run_ = false, return_finesampled_ = true;
	unsigned int omp_threads_ = 1;
	int device_ = 0;
	std::vector<ProfilePtr> profiles_;
};

void init();
void finish();
// drops the device set-ups that repeated Model::evaluate calls reuse
// (the role of R_profit_clear_cache, src/r_profit.cpp:822-825)
void clear_cache();
bool has_openmp();
bool has_fftw();
bool has_opencl();
std::string version();

}  // namespace profit
