// Radial-profile maths shared by host set-up and the GPU kernels.
//
// Reference (paths relative to src/libprofit/src/profit/):
//   radial.cpp:46-52    image -> profile coordinates
//   radial.cpp:136-199  RadialProfile::initial_calculations (adjust=true rewrites
//                       rscale_switch, resolution, rscale_max and acc)
//   radial.cpp:204-214  RadialProfile::validate
//   radial.h:93-102     boxy_r
//   sersic.cpp:79-360   Sersic: pow-free r_factor templates, b_n = qgamma(0.5, 2n),
//                       lumtot, fluxfrac-based adjustments, centre-pixel rule
//   moffat.cpp:52-107, ferrer.cpp:50-125, king.cpp:51-118,
//   coresersic.cpp:51-140, brokenexponential.cpp:38-149
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "special.h"

namespace profit_amd {

#define PA_HD __host__ __device__ inline
// hot-path profile math: always inlined into the render and walker kernels
#define PA_HDI __host__ __device__ __forceinline__

enum ProfileType : int32_t {
	SERSIC = 0, MOFFAT = 1, FERRER = 2, KING = 3, CORESERSIC = 4, BROKENEXP = 5,
	SKY = 100,      // libprofit "sky": a constant background (sky.cpp), bg in shape[0]
	NULLPROF = 101, // libprofit "null": draws nothing
	POINTSRC = 102  // libprofit "psf": a PSF point source (psf.cpp), parameters xcen, ycen, mag
};

PA_HD bool is_radial(int type) { return type >= SERSIC && type <= BROKENEXP; }

// shape parameters of a radial family (all five when the family is not known
// at compile time)
__host__ __device__ constexpr int n_shape(int type)
{
	return type == SERSIC || type == MOFFAT ? 2 : type == FERRER || type == KING ? 3 : type == BROKENEXP ? 4 : 5;
}

// gradient slots per component
enum GradSlot : int { G_XCEN = 0, G_YCEN, G_MAG, G_ANG, G_AXRAT, G_BOX, G_S0, G_S1, G_S2, G_S3, G_S4, NGRAD };

// The libprofit parameters of one radial profile; the C-ABI struct
// profit_amd_profile has the identical layout.
struct ProfileParams {
	int32_t type;
	int32_t convolve, rough, adjust, rescale_flux;
	int32_t reserved;
	double xcen, ycen, mag, ang, axrat, box;
	double acc, rscale_switch, rscale_max;
	uint32_t resolution, max_recursions;
	double shape[5];
};

PA_HD void default_params(int type, ProfileParams *p)
{
	p->type = type;
	p->convolve = 0;
	p->rough = 0;
	p->adjust = 1;
	p->rescale_flux = 0;
	p->reserved = 0;
	p->xcen = 0;
	p->ycen = 0;
	p->mag = 15;
	p->ang = 0;
	p->axrat = 1;
	p->box = 0;
	p->acc = 0.1;
	p->rscale_switch = 1;
	p->rscale_max = 0;
	p->resolution = 9;
	p->max_recursions = 2;
	for (int i = 0; i < 5; ++i) p->shape[i] = 0;
	switch (type) {
	case SERSIC: p->shape[0] = 1; p->shape[1] = 1; break;
	case MOFFAT: p->shape[0] = 3; p->shape[1] = 2; break;
	case FERRER: p->shape[0] = 3; p->shape[1] = 1; p->shape[2] = 1; p->acc = 1; break;
	case KING: p->shape[0] = 1; p->shape[1] = 3; p->shape[2] = 2; break;
	case CORESERSIC: p->shape[0] = 1; p->shape[1] = 1; p->shape[2] = 4; p->shape[3] = 1; p->shape[4] = 1; break;
	case BROKENEXP: p->shape[0] = 1; p->shape[1] = 1; p->shape[2] = 1; p->shape[3] = 1; break;
	case POINTSRC: p->mag = 0; break;  // psf.cpp:147-156
	default: break;                    // sky.cpp:65-71: bg = 0
	}
}

// Sersic r_factor specialisations (sersic.cpp:79-163)
enum SersicKind : int32_t { K_HALF = 0, K_ONE, K_TWO, K_THREE, K_FOUR, K_EIGHT, K_SIXTEEN, K_GENERAL };

// Device descriptor of one component after initial_calculations.
struct Comp {
	int32_t type, kind, res, maxrec;
	int32_t rough, centre_rule, boxy, has_rmax;
	int32_t convolve, instance, grad_ok, adjust;
	int32_t grad64;  // steep core: the direct-pixel adjoint sums its FP32 per-pixel derivatives in FP64
	int32_t grad_box;  // the box derivative is wanted (it costs logarithms per sample)
	double xc, yc;                    // centre after the PSF-padding shift
	double cos_ang, sin_ang, cos_over, sin_over;
	double rmax_r, rsw_r;             // rscale_max*rscale, rscale_switch*rscale
	double acc, flux_scale;
	double img_sx, img_sy;            // Model image pixel scale (centre rule)
	double box, rscale, bn, axrat;
	double s[5];
	double dbn;                       // d b_n / d nser (Sersic, core-Sersic)
	double gk;                        // Moffat: d ln rscale / d con
	double dlnfs[NGRAD];              // d ln(flux_scale) / d slot
	double lumtot, rbox, ie;
	double rscale_switch, rscale_max; // adjusted values
	// Sersic: 1/re, 1/re^2, 1/n, 1/(n B) and 1/(2n) (B = box + 2), so the
	// FP32 values and the derivatives multiply instead of dividing
	double inv_re, inv_re2, inv_n, inv_nB, inv_2n;
	double inv_axrat;
	// FP32 copies of the constants the FP32 evaluations read, so that they
	// enter FP32 arithmetic as scalar operands without conversions; aux per
	// family: King 1/rc, 1/a, C = (1 + (rt/rc)^2)^(-1/a), log2(1 + (rt/rc)^2)
	// and C / (1 + (rt/rc)^2); Moffat and Ferrer 1/rscale; core-Sersic
	// log2 rb, rb^a, re^-a, 1/(n a) and log2 re
	struct F32 {
		float cos_ang, sin_ang, cos_over, sin_over, axrat, inv_axrat;
		float inv_re, inv_re2, inv_n, inv_nB, inv_2n, bn, dbn;
		float acc, rscale, box, gk;
		float rmax2, rsw2;  // (rscale_max rscale)^2, (rscale_switch rscale)^2
		float s[5];
		float aux[5];
	} f;
};

// the FP32 copy of a component constant for T = float, the FP64 one otherwise
template <class T>
__host__ __device__ __forceinline__ T pick(double d, float f)
{
	if constexpr (sizeof(T) == 4) return f;
	else return d;
}

// validation error codes (radial.cpp:204-214 and each profile's validate())
enum SetupError : int32_t {
	SE_OK = 0,
	SE_AXRAT_LE0, SE_AXRAT_GT1, SE_BOX, SE_RE, SE_NSER, SE_FWHM, SE_CON, SE_ROUT, SE_FA, SE_FB,
	SE_RC, SE_RT, SE_KA, SE_CS_RE, SE_CS_RB, SE_CS_NSER, SE_CS_A, SE_CS_B, SE_H1, SE_H2, SE_BE_RB,
	SE_TYPE, SE_MAXREC, SE_RESOLUTION, SE_NO_PSF
};

inline const char *setup_error_message(int e)
{
	switch (e) {
	case SE_AXRAT_LE0: return "axrat <= 0, must have axrat > 0";
	case SE_AXRAT_GT1: return "axrat > 1, must have axrat <= 1";
	case SE_BOX: return "box <= -2, must have box > -2";
	case SE_RE: return "re <= 0, must have re > 0";
	case SE_NSER: return "nser <= 0, must have nser > 0";
	case SE_FWHM: return "fwhm <= 0, must have fwhm > 0";
	case SE_CON: return "con < 0, must have con >= 0";
	case SE_ROUT: return "rout <= 0, must have rout >= 0";
	case SE_FA: return "a < 0, must have a >= 0";
	case SE_FB: return "b > 2, must have b <= 2";
	case SE_RC: return "rc <= 0, must have rc > 0";
	case SE_RT: return "rt <= 0, must have rt > 0";
	case SE_KA: return "a < 0, must have a >=0";
	case SE_CS_RE: return "re <= 0, must have re > 0";
	case SE_CS_RB: return "rb <= 0, must have rb > 0";
	case SE_CS_NSER: return "nser <= 0, must have nser > 0";
	case SE_CS_A: return "a <= 0, must have a > 0";
	case SE_CS_B: return "b > 1.999, must have b < 1.999";
	case SE_H1: return "h1 <= 0, must have h1 > 0";
	case SE_H2: return "h2 <= 0, must have h2 > 0";
	case SE_BE_RB: return "rb <= 0, must have rb > 0";
	case SE_TYPE: return "unknown profile type";
	case SE_MAXREC: return "max_recursions above 30 is not supported by the GPU recursion walker";
	case SE_RESOLUTION: return "resolution above 16 is not supported by the GPU recursion walker";
	case SE_NO_PSF: return "No psf present in the model, cannot produce a psf profile";
	default: return "ok";
	}
}

PA_HD int validate_params(const ProfileParams &p)
{
	const double *s = p.shape;
	// PsfProfile::validate only asks for a PSF (checked against the galaxy);
	// SkyProfile::validate checks nothing (psf.cpp:40-45, sky.cpp:36-39)
	if (p.type == POINTSRC || p.type == SKY || p.type == NULLPROF) return SE_OK;
	if (p.axrat <= 0) return SE_AXRAT_LE0;
	if (p.axrat > 1) return SE_AXRAT_GT1;
	if (p.box <= -2) return SE_BOX;
	switch (p.type) {
	case SERSIC:
		if (s[0] <= 0) return SE_RE;
		if (s[1] <= 0) return SE_NSER;
		break;
	case MOFFAT:
		if (s[0] <= 0) return SE_FWHM;
		if (s[1] < 0) return SE_CON;
		break;
	case FERRER:
		if (s[0] <= 0) return SE_ROUT;
		if (s[1] < 0) return SE_FA;
		if (s[2] > 2) return SE_FB;
		break;
	case KING:
		if (s[0] <= 0) return SE_RC;
		if (s[1] <= 0) return SE_RT;
		if (s[2] < 0) return SE_KA;
		break;
	case CORESERSIC:
		if (s[0] <= 0) return SE_CS_RE;
		if (s[1] <= 0) return SE_CS_RB;
		if (s[2] <= 0) return SE_CS_NSER;
		if (s[3] <= 0) return SE_CS_A;
		if (s[4] > 1.999) return SE_CS_B;
		break;
	case BROKENEXP:
		if (s[0] <= 0) return SE_H1;
		if (s[1] <= 0) return SE_H2;
		if (s[2] <= 0) return SE_BE_RB;
		break;
	default: return SE_TYPE;
	}
	return SE_OK;
}

// ------------------------------------------------------------------ evaluate_at

// FP32 log2 and exp2 on the transcendental unit (v_log_f32, v_exp_f32: about
// 1 ulp for normal arguments).  The FP32 profile values and derivatives need
// a few parts in 1e7.  A subnormal argument counts as 0 (log2 -> -inf): the
// profiles take logarithms of squared radii, sums of squares and profile
// values, for which 1e-38 and 0 give the same result to FP32 precision but
// at an exact zero, which this keeps exact.
__device__ __forceinline__ float log2_fast(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float exp2_fast(float x) { return __builtin_amdgcn_exp2f(x); }
constexpr double kLog2e = 1.4426950408889634, kLn2 = 0.6931471805599453;

// FP32 exp, log and pow to a few ulps on the transcendental unit, by
// argument reduction: log2 x = k + log2 m with m = x 2^-k in [1/sqrt2, sqrt2),
// whose logarithm (|log2 m| <= 1/2) the unit gives to an ulp; 2^y = 2^n
// 2^(y - n) with n = rint(y).  The products that form y keep their rounding
// errors (fma), so the result errs by a few ulps for any |y|, where
// 2^(e log2 x) in plain FP32 errs by about |y| ulps.
__device__ __forceinline__ float exp2_hl(float hi, float lo)
{
	const float n = rintf(hi);
	return ldexpf(exp2_fast((hi - n) + lo), (int)n);
}
__device__ __forceinline__ float log2_kl(float x, float *k)
{
	int e;
	float m = frexpf(x, &e);
	if (m < 0.70710678f) {
		m *= 2.f;
		--e;
	}
	*k = (float)e;
	return log2_fast(m);
}
__device__ __forceinline__ float pow_acc(float x, float e)
{
	float k;
	const float l = log2_kl(x, &k);
	const float p = e * k;
	return exp2_hl(p, fmaf(e, l, fmaf(e, k, -p)));
}
__device__ __forceinline__ float exp_acc(float x)
{
	const float L = 1.44269502f, Llo = 1.92596299e-8f;  // log2 e = L + Llo
	const float hi = x * L;
	return exp2_hl(hi, fmaf(x, Llo, fmaf(x, L, -hi)));
}
__device__ __forceinline__ float log_acc(float x)
{
	float k;
	const float l = log2_kl(x, &k);
	// k ln 2 with ln 2 = 0.693145751953125 + 1.42860677e-6; the first product
	// is exact (11 significant bits)
	return fmaf(k, 0.693145751953125f, fmaf(k, 1.42860677e-6f, l * 0.693147182f));
}

// pow, exp and log in precision T: FP64 takes the library functions, FP32
// the forms above.  The profiles' FP32 bases are never negative but for
// rounding at a zero crossing (King's A - C, Ferrer's 1 - t^(2-b) at the
// cut-off), so x <= 0 is taken as 0: 0 for e > 0, 1 for e = 0, inf below.
template <class T>
PA_HDI T pow_t(T x, T e)
{
	if constexpr (sizeof(T) == 4) {
		if (x > T(0)) return e == T(0) ? T(1) : pow_acc(x, e);
		if (x != x) return x;
		return e > T(0) ? T(0) : (e == T(0) ? T(1) : T(INFINITY));
	}
	else {
		return ::pow(x, e);
	}
}
template <class T>
PA_HDI T exp_t(T x)
{
	if constexpr (sizeof(T) == 4) return exp_acc(x);
	else return ::exp(x);
}
template <class T>
PA_HDI T log_t(T x)
{
	if constexpr (sizeof(T) == 4) {
		if (x > T(0) && x < T(INFINITY)) return log_acc(x);
	}
	return ::log(x);
}

// a / b: FP64 divides; FP32 multiplies by the hardware reciprocal (1 ulp),
// for the derivative arithmetic, where a correctly rounded quotient buys
// nothing (the FP32 values keep the division)
template <class T>
PA_HDI T div_t(T a, T b)
{
	if constexpr (sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
	else return a / b;
}

template <class T>
PA_HDI T boxy_r(const Comp &c, T x, T y)
{
	if constexpr (sizeof(T) == 4) {
		// the hardware square root (about 1 ulp), not the correctly rounded one
		if (!c.boxy) return __builtin_amdgcn_sqrtf(x * x + y * y);
		const float b2 = c.f.box + 2.f;
		return pow_t<float>(pow_t<float>(::fabsf(x), b2) + pow_t<float>(::fabsf(y), b2), 1.f / b2);
	}
	else {
		if (!c.boxy) return ::sqrt(x * x + y * y);
		const T b2 = pick<T>(c.box, c.f.box) + T(2);
		return pow_t<T>(pow_t<T>(::fabs(x), b2) + pow_t<T>(::fabs(y), b2), T(1) / b2);
	}
}

// FP32 Sersic value without box at profile position (x, y), with the
// intermediates its derivatives share: base = (r / re)^2, l2 = log2(base)
// and rf = base^(1/(2n)).
PA_HDI float sersic_parts_f32(const Comp &c, float x, float y, float &base, float &l2, float &rf)
{
	base = (x * x + y * y) * c.f.inv_re2;
	l2 = log2_fast(base);
	rf = exp2_fast(l2 * c.f.inv_2n);
	return exp2_fast(-c.f.bn * (rf - 1.f) * float(kLog2e));
}

template <class T>
PA_HDI T sersic_at(const Comp &c, T x, T y)
{
	if constexpr (sizeof(T) == 4) {
		// base^(1/(n B)) = 2^(log2(base) / (n B)) and the exponential on the
		// transcendental unit, for every n (the FP64 path keeps libprofit's
		// special cases of n)
		if (c.boxy) {
			const float B = c.f.box + 2.f;
			const float base = pow_t<float>(::fabsf(x * c.f.inv_re), B) + pow_t<float>(::fabsf(y * c.f.inv_re), B);
			const float rf = exp2_fast(log2_fast(base) * c.f.inv_nB);
			return exp2_fast(-c.f.bn * (rf - 1.f) * float(kLog2e));
		}
		float base, l2, rf;
		return sersic_parts_f32(c, x, y, base, l2, rf);
	}
	else {
		const T re = pick<T>(c.s[0], c.f.s[0]), nser = pick<T>(c.s[1], c.f.s[1]), bn = pick<T>(c.bn, c.f.bn);
		T base, rf;
		if (c.boxy) {
			const T B = pick<T>(c.box, c.f.box) + T(2);
			base = pow_t<T>(::fabs(x / re), B) + pow_t<T>(::fabs(y / re), B);
			switch (c.kind) {
			case K_HALF: rf = base * base; break;
			case K_ONE: rf = base; break;
			case K_TWO: rf = ::sqrt(base); break;
			case K_THREE: rf = ::cbrt(base); break;
			case K_FOUR: rf = ::sqrt(::sqrt(base)); break;
			case K_EIGHT: rf = ::sqrt(::sqrt(::sqrt(base))); break;
			case K_SIXTEEN: rf = ::sqrt(::sqrt(::sqrt(::sqrt(base)))); break;
			default: rf = pow_t<T>(base, T(1) / (nser * B)); break;
			}
		}
		else {
			base = (x * x + y * y) / (re * re);
			switch (c.kind) {
			case K_HALF: rf = base; break;
			case K_ONE: rf = ::sqrt(base); break;
			case K_TWO: rf = ::sqrt(::sqrt(base)); break;
			case K_THREE: rf = ::cbrt(::sqrt(base)); break;
			case K_FOUR: rf = ::sqrt(::sqrt(::sqrt(base))); break;
			case K_EIGHT: rf = ::sqrt(::sqrt(::sqrt(::sqrt(base)))); break;
			case K_SIXTEEN: rf = ::sqrt(::sqrt(::sqrt(::sqrt(::sqrt(base))))); break;
			default: rf = pow_t<T>(::sqrt(base), T(1) / nser); break;
			}
		}
		return exp_t<T>(-bn * (rf - T(1)));
	}
}

template <class T>
PA_HDI T king_core(T r, T rc, T rt, T a)
{
	return pow_t<T>(T(1) / pow_t<T>(T(1) + pow_t<T>(r / rc, T(2)), T(1) / a) - T(1) / pow_t<T>(T(1) + pow_t<T>(rt / rc, T(2)), T(1) / a), a);
}

template <class T>
PA_HDI T coresersic_core(const Comp &c, T r)
{
	const T re = pick<T>(c.s[0], c.f.s[0]), rb = pick<T>(c.s[1], c.f.s[1]), nser = pick<T>(c.s[2], c.f.s[2]);
	const T a = pick<T>(c.s[3], c.f.s[3]), b = pick<T>(c.s[4], c.f.s[4]);
	return pow_t<T>(T(1) + pow_t<T>(r / rb, -a), b / a) *
	       exp_t<T>(-pick<T>(c.bn, c.f.bn) * pow_t<T>((pow_t<T>(r, a) + pow_t<T>(rb, a)) / pow_t<T>(re, a), T(1) / (nser * a)));
}

template <class T>
PA_HDI T brokenexp_core(T r, T h1, T h2, T rb, T a)
{
	T base = r - rb;
	const T expo = T(1) / h1 - T(1) / h2;
	if (a * base < T(40)) base = log_t<T>(T(1) + exp_t<T>(a * base)) / a;
	return exp_t<T>(-r / h1 + expo * base);
}

// d ln f / dr and d ln f / d shape of the core-Sersic profile
// f = (1 + (r/rb)^-a)^(b/a) exp(-bn ((r^a + rb^a)/re^a)^(1/(n a))),
// shape = (re, rb, nser, a, b); bn = qgamma(0.5, 2 nser), dbn its n-derivative.
template <class T>
PA_HDI T coresersic_dlog(const Comp &c, T r, T *ds)
{
	const T re = pick<T>(c.s[0], c.f.s[0]), rb = pick<T>(c.s[1], c.f.s[1]), n = pick<T>(c.s[2], c.f.s[2]), a = pick<T>(c.s[3], c.f.s[3]), b = pick<T>(c.s[4], c.f.s[4]), bn = pick<T>(c.bn, c.f.bn);
	const T t = pow_t<T>(r / rb, -a), Q = T(1) + t;
	const T ra = pow_t<T>(r, a), rba = pow_t<T>(rb, a), sum = ra + rba;
	const T S = sum / pow_t<T>(re, a);
	const T lS = log_t<T>(S);
	const T W = exp_t<T>(lS / (n * a));
	ds[0] = bn * W / (n * re);
	ds[1] = b * t / (rb * Q) - bn * W * rba / (rb * n * sum);
	ds[2] = -pick<T>(c.dbn, c.f.dbn) * W + bn * W * lS / (n * n * a);
	const T dlnS_da = (ra * log_t<T>(r) + rba * log_t<T>(rb)) / sum - log_t<T>(re);
	ds[3] = -(b / (a * a)) * log_t<T>(Q) - (b / a) * t * log_t<T>(r / rb) / Q - bn * W * (dlnS_da / (n * a) - lS / (n * a * a));
	ds[4] = log_t<T>(Q) / a;
	return -b * t / (r * Q) - bn * W * ra / (r * n * sum);
}

// d ln f / dr and d ln f / d shape of the broken exponential,
// shape = (h1, h2, rb, a); the soft break switches to its asymptote when
// a (r - rb) >= 40 as the profile itself does.
template <class T>
PA_HDI T brokenexp_dlog(T r, T h1, T h2, T rb, T a, T *ds)
{
	const T z = a * (r - rb);
	const T expo = T(1) / h1 - T(1) / h2;
	T B, sig, dBda;
	if (z < T(40)) {
		const T ez = exp_t<T>(z);
		B = log_t<T>(T(1) + ez) / a;
		sig = ez / (T(1) + ez);
		dBda = (r - rb) * sig / a - B / a;
	}
	else {
		B = r - rb;
		sig = T(1);
		dBda = T(0);
	}
	ds[0] = (r - B) / (h1 * h1);
	ds[1] = B / (h2 * h2);
	ds[2] = -expo * sig;
	ds[3] = expo * dBda;
	return -T(1) / h1 + expo * sig;
}

// The core-Sersic profile in FP32 from one logarithm of r:
// t = (r/rb)^-a, Q = 1 + t, S = (r^a + rb^a) / re^a, W = S^(1/(n a)),
// f = Q^(b/a) exp(-bn W).  With ds non-null it also gives d ln f / d shape
// (ds[0..4]) and returns d ln f / dr in *dlog, as coresersic_dlog does.
__device__ __forceinline__ float coresersic_f32(const Comp &c, float r, float *ds, float *dlog)
{
	const Comp::F32 &k = c.f;
	const float re = k.s[0], rb = k.s[1], a = k.s[3], b = k.s[4], bn = k.bn;
	const float lrb = k.aux[0], rba = k.aux[1], rea = k.aux[2], ina = k.aux[3], lre = k.aux[4];
	const float lr = log2_fast(r);
	const float t = exp2_fast(-a * (lr - lrb)), Q = 1.f + t, lQ = log2_fast(Q);
	const float ra = exp2_fast(a * lr), sum = ra + rba, S = sum * rea, lS2 = log2_fast(S);
	const float W = exp2_fast(lS2 * ina);
	const float ia = 1.f / a;
	const float e1 = b == 0.f ? 0.f : b * ia * lQ;
	const float f = exp2_fast(e1 - bn * W * float(kLog2e));
	if (ds) {
		const float ln2 = float(kLn2), in = a * ina, lS = lS2 * ln2;
		const float iQ = __builtin_amdgcn_rcpf(Q), isum = __builtin_amdgcn_rcpf(sum);
		const float irb = __builtin_amdgcn_rcpf(rb), ir = __builtin_amdgcn_rcpf(r);
		ds[0] = bn * W * in * __builtin_amdgcn_rcpf(re);
		ds[1] = (b * t * iQ - bn * W * rba * in * isum) * irb;
		ds[2] = -k.dbn * W + bn * W * lS * ina * in;
		const float dlnS_da = (ra * lr + rba * lrb) * ln2 * isum - lre * ln2;
		ds[3] = -(b * ia * ia) * lQ * ln2 - b * ia * t * (lr - lrb) * ln2 * iQ - bn * W * (dlnS_da * ina - lS * ina * ia);
		ds[4] = lQ * ln2 * ia;
		*dlog = (-b * t * iQ - bn * W * ra * in * isum) * ir;
	}
	return f;
}

// King and Ferrer values within 3% of the cut-off (rt, rout) lose relative
// precision in FP32 (A - C and 1 - (r/rout)^(2-b) cancel, and r itself is
// rounded): directly evaluated pixels there take the FP64 evaluation.
template <int TYPE = -1>
__device__ __forceinline__ bool near_cut_fp64(const Comp &c, float xp, float yp)
{
	if (TYPE >= 0 && TYPE != KING && TYPE != FERRER) return false;
	const int t = TYPE >= 0 ? TYPE : c.type;
	if (t != KING && t != FERRER) return false;
	const float cut = t == KING ? c.f.s[1] : c.f.rscale;
	return boxy_r<float>(c, xp, yp) >= 0.97f * cut;
}

// The profile in FP32 with the plain transcendental forms (2^(e log2 x),
// errors of about |e log2 x| ulps): the refinement test's second value,
// which only has to fall on the right side of a guard band of 1e-3.
template <int TYPE = -1>
__device__ __forceinline__ float eval_fast(const Comp &c, float x, float y)
{
	const Comp::F32 &f = c.f;
	switch (TYPE >= 0 ? TYPE : c.type) {
	case SERSIC: return sersic_at<float>(c, x, y);
	case MOFFAT: {
		const float rf = boxy_r<float>(c, x, y) * f.aux[0];
		return exp2_fast(-f.s[1] * log2_fast(1.f + rf * rf));
	}
	case FERRER: {
		const float rf = boxy_r<float>(c, x, y) * f.aux[0];
		if (!(rf < 1.f)) return 0.f;
		const float g = 1.f - exp2_fast((2.f - f.s[2]) * log2_fast(rf));
		return g > 0.f ? exp2_fast(f.s[1] * log2_fast(g)) : 0.f;
	}
	case KING: {
		const float r = boxy_r<float>(c, x, y);
		if (!(r < f.s[1])) return 0.f;
		const float u = r * f.aux[0];
		const float D = exp2_fast(-f.aux[1] * log2_fast(1.f + u * u)) - f.aux[2];
		return D > 0.f ? exp2_fast(f.s[2] * log2_fast(D)) : 0.f;
	}
	case CORESERSIC: return coresersic_f32(c, boxy_r<float>(c, x, y), nullptr, nullptr);
	case BROKENEXP: return brokenexp_core<float>(boxy_r<float>(c, x, y), f.s[0], f.s[1], f.s[2], f.s[3]);
	default: return 0.f;
	}
}

// TYPE >= 0 compiles the evaluation for that profile family only (the
// kernels specialised per family); -1 dispatches on c.type at run time.
template <class T, int TYPE = -1>
PA_HDI T eval_at(const Comp &c, T x, T y)
{
	if constexpr (sizeof(T) == 4) {
		// FP32: the constants' FP32 copies, multiplications by reciprocals,
		// and King's constant term (1 + (rt/rc)^2)^(-1/a) from the set-up
		const Comp::F32 &f = c.f;
		switch (TYPE >= 0 ? TYPE : c.type) {
		case SERSIC: return sersic_at<float>(c, x, y);
		case MOFFAT: {
			const float rf = boxy_r<float>(c, x, y) * f.aux[0];
			return pow_t<float>(1.f + rf * rf, -f.s[1]);
		}
		case FERRER: {
			const float rf = boxy_r<float>(c, x, y) * f.aux[0];
			if (rf < 1.f) return pow_t<float>(1.f - pow_t<float>(rf, 2.f - f.s[2]), f.s[1]);
			return 0.f;
		}
		case KING: {
			const float r = boxy_r<float>(c, x, y);
			if (!(r < f.s[1])) return 0.f;
			const float u = r * f.aux[0];
			return pow_t<float>(pow_t<float>(1.f + u * u, -f.aux[1]) - f.aux[2], f.s[2]);
		}
		case CORESERSIC: return coresersic_core<float>(c, boxy_r<float>(c, x, y));
		case BROKENEXP: return brokenexp_core<float>(boxy_r<float>(c, x, y), f.s[0], f.s[1], f.s[2], f.s[3]);
		default: return 0.f;
		}
	}
	else {
		switch (TYPE >= 0 ? TYPE : c.type) {
		case SERSIC: return sersic_at<T>(c, x, y);
		case MOFFAT: {
			const T rf = boxy_r<T>(c, x, y) / pick<T>(c.rscale, c.f.rscale);
			return pow_t<T>(T(1) + rf * rf, -pick<T>(c.s[1], c.f.s[1]));
		}
		case FERRER: {
			const T rf = boxy_r<T>(c, x, y) / pick<T>(c.rscale, c.f.rscale);
			if (rf < T(1)) return pow_t<T>(T(1) - pow_t<T>(rf, T(2) - pick<T>(c.s[2], c.f.s[2])), pick<T>(c.s[1], c.f.s[1]));
			return T(0);
		}
		case KING: {
			const T r = boxy_r<T>(c, x, y);
			if (r < pick<T>(c.s[1], c.f.s[1])) return king_core<T>(r, pick<T>(c.s[0], c.f.s[0]), pick<T>(c.s[1], c.f.s[1]), pick<T>(c.s[2], c.f.s[2]));
			return T(0);
		}
		case CORESERSIC: return coresersic_core<T>(c, boxy_r<T>(c, x, y));
		case BROKENEXP: return brokenexp_core<T>(boxy_r<T>(c, x, y), pick<T>(c.s[0], c.f.s[0]), pick<T>(c.s[1], c.f.s[1]), pick<T>(c.s[2], c.f.s[2]), pick<T>(c.s[3], c.f.s[3]));
		default: return T(0);
		}
	}
}

// ------------------------------------------------------------------ derivatives

template <class T>
PA_HDI void dr_terms(const Comp &c, T x, T y, T r, T *drx, T *dry, T *drB)
{
	*drx = *dry = *drB = T(0);
	if (!(r > T(0))) return;
	const T B = pick<T>(c.box, c.f.box) + T(2);
	const T ax = ::fabs(x), ay = ::fabs(y);
	if (!c.boxy) {
		if constexpr (sizeof(T) == 4) {
			const float ir = __builtin_amdgcn_rcpf(r);
			*drx = x * ir;
			*dry = y * ir;
		}
		else {
			*drx = x / r;
			*dry = y / r;
		}
	}
	else {
		const T rp = pow_t<T>(r, T(1) - B);
		*drx = ax > T(0) ? rp * pow_t<T>(ax, B - T(1)) * (x > T(0) ? T(1) : T(-1)) : T(0);
		*dry = ay > T(0) ? rp * pow_t<T>(ay, B - T(1)) * (y > T(0) ? T(1) : T(-1)) : T(0);
	}
	if (!c.grad_box) return;
	const T px = pow_t<T>(ax, B), py = pow_t<T>(ay, B);
	const T S = px + py;
	const T Sl = (ax > T(0) ? px * log_t<T>(ax) : T(0)) + (ay > T(0) ? py * log_t<T>(ay) : T(0));
	*drB = r * (div_t<T>(Sl, B * S) - div_t<T>(log_t<T>(S), B * B));
}

template <class T>
PA_HDI T king_dcore(T r, T rc, T rt, T a, int which)
{
	const T u2 = (r / rc) * (r / rc), w2 = (rt / rc) * (rt / rc);
	const T A = pow_t<T>(T(1) + u2, -T(1) / a), C = pow_t<T>(T(1) + w2, -T(1) / a), D = A - C;
	if (!(D > T(0))) return T(0);
	const T Da1 = a * pow_t<T>(D, a - T(1));
	switch (which) {
	case -1: return Da1 * (-(T(2) * r / (a * rc * rc)) * A / (T(1) + u2));
	case 0: return Da1 * ((T(2) * u2 / (a * rc)) * A / (T(1) + u2) - (T(2) * w2 / (a * rc)) * C / (T(1) + w2));
	case 1: return Da1 * ((T(2) * rt / (a * rc * rc)) * C / (T(1) + w2));
	default: return pow_t<T>(D, a) * (log_t<T>(D) + (A * log_t<T>(T(1) + u2) - C * log_t<T>(T(1) + w2)) / (a * D));
	}
}

// o[0] = f, o[1] = df/dx_prof, o[2] = df/dy_prof, o[3] = df/dbox, o[4..8] = df/dshape
template <class T, int TYPE = -1>
PA_HDI void sample_grad(const Comp &c, T x, T y, T *o)
{
	for (int i = 0; i < 9; ++i) o[i] = T(0);
	switch (TYPE >= 0 ? TYPE : c.type) {
	case SERSIC: {
		// f = exp(-bn (base^(1/(n B)) - 1)), base = |x/re|^B + |y/re|^B, B = box + 2.
		// Without box the powers are squares, and f follows from the same
		// logarithm as the derivatives: one log and two exps per sample
		// (plus two logs when the box derivative is wanted).
		const T bn = pick<T>(c.bn, c.f.bn), B = pick<T>(c.box, c.f.box) + T(2);
		const T ire = pick<T>(c.inv_re, c.f.inv_re), in = pick<T>(c.inv_n, c.f.inv_n), inB = pick<T>(c.inv_nB, c.f.inv_nB);
		const T u = ::fabs(x * ire), v = ::fabs(y * ire);
		T pu, pv;
		if (c.boxy) {
			pu = pow_t<T>(u, B);
			pv = pow_t<T>(v, B);
		}
		else {
			pu = u * u;
			pv = v * v;
		}
		const T base = pu + pv;
		if (!(base > T(0))) {
			o[0] = exp_t<T>(bn);
			return;
		}
		T lb, rf, f;
		if constexpr (sizeof(T) == 4) {
			const T l2 = log2_fast(base);
			rf = exp2_fast(l2 * inB);
			f = exp2_fast(-bn * (rf - T(1)) * T(kLog2e));
			lb = l2 * T(kLn2);
		}
		else {
			lb = log_t<T>(base);
			rf = exp_t<T>(lb * inB);
			f = exp_t<T>(-bn * (rf - T(1)));
		}
		o[0] = f;
		const T k = div_t<T>(-bn * f * rf * inB, base);  // df/dbase
		// d base / d x_prof = B pu / u / re (2 u / re without box)
		const T bx = c.boxy ? (u > T(0) ? B * pu / u : T(0)) : T(2) * u;
		const T by = c.boxy ? (v > T(0) ? B * pv / v : T(0)) : T(2) * v;
		o[1] = k * bx * ire * (x > T(0) ? T(1) : T(-1));
		o[2] = k * by * ire * (y > T(0) ? T(1) : T(-1));
		if (c.grad_box) {
			const T dbase_dB = (u > T(0) ? pu * log_t<T>(u) : T(0)) + (v > T(0) ? pv * log_t<T>(v) : T(0));
			o[3] = -bn * f * rf * (dbase_dB * inB / base - lb * inB / B);
		}
		o[4] = bn * f * rf * in * ire;
		o[5] = f * (-pick<T>(c.dbn, c.f.dbn) * (rf - T(1)) + bn * rf * lb * in * inB);
		return;
	}
	case MOFFAT: {
		const T r = boxy_r<T>(c, x, y), rs = pick<T>(c.rscale, c.f.rscale), con = pick<T>(c.s[1], c.f.s[1]);
		const T t = r / rs, q = T(1) + t * t;
		if constexpr (sizeof(T) == 4) {
			// one logarithm for the value and the con derivative
			const float lq = log2_fast(q), f = exp2_fast(-con * lq);
			o[0] = f;
			const float dfdr = -con * f * __builtin_amdgcn_rcpf(q) * 2.f * t * c.f.aux[0];
			float drx, dry, drB;
			dr_terms<float>(c, x, y, r, &drx, &dry, &drB);
			o[1] = dfdr * drx;
			o[2] = dfdr * dry;
			o[3] = dfdr * drB;
			const float dfdrs = -dfdr * r * c.f.aux[0];
			o[4] = dfdrs * rs * __builtin_amdgcn_rcpf(c.f.s[0]);
			o[5] = -f * lq * float(kLn2) + dfdrs * rs * c.f.gk;
			return;
		}
		const T f = pow_t<T>(q, -con);
		o[0] = f;
		const T dfdr = -con * f / q * T(2) * t / rs;
		T drx, dry, drB;
		dr_terms<T>(c, x, y, r, &drx, &dry, &drB);
		o[1] = dfdr * drx;
		o[2] = dfdr * dry;
		o[3] = dfdr * drB;
		const T dfdrs = -dfdr * r / rs;
		o[4] = dfdrs * rs / pick<T>(c.s[0], c.f.s[0]);
		o[5] = -f * log_t<T>(q) + dfdrs * rs * pick<T>(c.gk, c.f.gk);
		return;
	}
	case FERRER: {
		const T r = boxy_r<T>(c, x, y), rout = pick<T>(c.s[0], c.f.s[0]), a = pick<T>(c.s[1], c.f.s[1]), b = pick<T>(c.s[2], c.f.s[2]);
		const T t = r / rout;
		if (!(t < T(1))) return;
		if constexpr (sizeof(T) == 4) {
			// t^(2-b), g^a and their logarithms shared: g^(a-1) = f / g,
			// t^(1-b) = t^(2-b) / t
			const float lt = log2_fast(t), tb = exp2_fast((2.f - b) * lt), g = 1.f - tb;
			if (!(g > 0.f)) return;
			const float lg = log2_fast(g), f = exp2_fast(a * lg);
			o[0] = f;
			float drx, dry, drB;
			dr_terms<float>(c, x, y, r, &drx, &dry, &drB);
			const float ga1 = f * __builtin_amdgcn_rcpf(g), irout = __builtin_amdgcn_rcpf(rout);
			const float dfdr = t > 0.f ? a * ga1 * (-(2.f - b) * tb * __builtin_amdgcn_rcpf(t)) * irout : 0.f;
			o[1] = dfdr * drx;
			o[2] = dfdr * dry;
			o[3] = dfdr * drB;
			o[4] = -dfdr * r * irout;
			o[5] = f * lg * float(kLn2);
			o[6] = t > 0.f ? a * ga1 * tb * lt * float(kLn2) : 0.f;
			return;
		}
		const T g = T(1) - pow_t<T>(t, T(2) - b);
		const T f = pow_t<T>(g, a);
		o[0] = f;
		T drx, dry, drB;
		dr_terms<T>(c, x, y, r, &drx, &dry, &drB);
		const T dfdr = r > T(0) ? a * pow_t<T>(g, a - T(1)) * (-(T(2) - b) * pow_t<T>(t, T(1) - b)) / rout : T(0);
		o[1] = dfdr * drx;
		o[2] = dfdr * dry;
		o[3] = dfdr * drB;
		o[4] = -dfdr * r / rout;
		o[5] = g > T(0) ? f * log_t<T>(g) : T(0);
		o[6] = t > T(0) ? a * pow_t<T>(g, a - T(1)) * pow_t<T>(t, T(2) - b) * log_t<T>(t) : T(0);
		return;
	}
	case KING: {
		const T r = boxy_r<T>(c, x, y);
		const T rc = pick<T>(c.s[0], c.f.s[0]), rt = pick<T>(c.s[1], c.f.s[1]), a = pick<T>(c.s[2], c.f.s[2]);
		if (!(r < rt)) return;
		if constexpr (sizeof(T) == 4) {
			// A = (1 + u^2)^(-1/a), D = A - C, f = D^a, with the logarithms
			// shared among the value and the four derivatives
			const float irc = c.f.aux[0], ia = c.f.aux[1], C = c.f.aux[2];
			const float u = r * irc, q = 1.f + u * u, lq = log2_fast(q);
			const float A = exp2_fast(-ia * lq), D = A - C;
			if (!(D > 0.f)) return;
			const float lD = log2_fast(D), f = exp2_fast(a * lD);
			o[0] = f;
			const float iD = __builtin_amdgcn_rcpf(D);
			const float Da1 = a * f * iD;  // a D^(a-1)
			const float w = rt * irc;
			const float Aq = A * __builtin_amdgcn_rcpf(q), Cq = c.f.aux[4];
			const float dfdr = Da1 * (-2.f * ia * irc * u) * Aq;
			float drx, dry, drB;
			dr_terms<float>(c, x, y, r, &drx, &dry, &drB);
			o[1] = dfdr * drx;
			o[2] = dfdr * dry;
			o[3] = dfdr * drB;
			o[4] = Da1 * 2.f * ia * irc * (u * u * Aq - w * w * Cq);
			o[5] = Da1 * 2.f * ia * irc * w * Cq;
			o[6] = f * float(kLn2) * (lD + (A * lq - C * c.f.aux[3]) * ia * iD);
			return;
		}
		o[0] = king_core<T>(r, rc, rt, a);
		T drx, dry, drB;
		dr_terms<T>(c, x, y, r, &drx, &dry, &drB);
		const T dfdr = king_dcore<T>(r, rc, rt, a, -1);
		o[1] = dfdr * drx;
		o[2] = dfdr * dry;
		o[3] = dfdr * drB;
		for (int w = 0; w < 3; ++w) o[4 + w] = king_dcore<T>(r, rc, rt, a, w);
		return;
	}
	case CORESERSIC: {
		const T r = boxy_r<T>(c, x, y);
		if constexpr (sizeof(T) == 4) {
			if (!(r > 0.f)) {
				o[0] = coresersic_f32(c, r, nullptr, nullptr);
				return;
			}
			float ds[5], dlog;
			const float f = coresersic_f32(c, r, ds, &dlog);
			o[0] = f;
			const float dfdr = f * dlog;
			float drx, dry, drB;
			dr_terms<float>(c, x, y, r, &drx, &dry, &drB);
			o[1] = dfdr * drx;
			o[2] = dfdr * dry;
			o[3] = dfdr * drB;
			for (int w = 0; w < 5; ++w) o[4 + w] = f * ds[w];
			return;
		}
		const T f = coresersic_core<T>(c, r);
		o[0] = f;
		if (!(r > T(0))) return;
		T ds[5];
		const T dfdr = f * coresersic_dlog<T>(c, r, ds);
		T drx, dry, drB;
		dr_terms<T>(c, x, y, r, &drx, &dry, &drB);
		o[1] = dfdr * drx;
		o[2] = dfdr * dry;
		o[3] = dfdr * drB;
		for (int w = 0; w < 5; ++w) o[4 + w] = f * ds[w];
		return;
	}
	case BROKENEXP: {
		const T r = boxy_r<T>(c, x, y);
		const T h1 = pick<T>(c.s[0], c.f.s[0]), h2 = pick<T>(c.s[1], c.f.s[1]), rb = pick<T>(c.s[2], c.f.s[2]), a = pick<T>(c.s[3], c.f.s[3]);
		const T f = brokenexp_core<T>(r, h1, h2, rb, a);
		o[0] = f;
		T ds[4];
		const T dfdr = f * brokenexp_dlog<T>(r, h1, h2, rb, a, ds);
		T drx, dry, drB;
		dr_terms<T>(c, x, y, r, &drx, &dry, &drB);
		o[1] = dfdr * drx;
		o[2] = dfdr * dry;
		o[3] = dfdr * drB;
		for (int w = 0; w < 4; ++w) o[4 + w] = f * ds[w];
		return;
	}
	default: return;
	}
}

// ------------------------------------------------------------------ set-up

struct KingIntegrand {
	double rc, rt, a;
	int which;  // -2: r f(r); 0..2: r df/dshape
	PA_HD double operator()(double r) const
	{
		if (!(r < rt)) return 0;
		if (which == -2) return r * king_core<double>(r, rc, rt, a);
		return r * king_dcore<double>(r, rc, rt, a, which);
	}
};

struct CoreSersicIntegrand {
	const Comp *c;
	int which = -1;  // -1: r f(r); 0..4: r df/dshape
	PA_HD double operator()(double r) const
	{
		const double f = coresersic_core<double>(*c, r);
		if (which < 0) return r * f;
		if (!(r > 0)) return 0;
		double ds[5];
		coresersic_dlog<double>(*c, r, ds);
		return r * f * ds[which];
	}
};

struct BrokenExpIntegrand {
	double h1, h2, rb, a;
	int which = -1;  // -1: r f(r); 0..3: r df/dshape
	PA_HD double operator()(double r) const
	{
		const double f = brokenexp_core<double>(r, h1, h2, rb, a);
		if (which < 0) return r * f;
		double ds[4];
		brokenexp_dlog<double>(r, h1, h2, rb, a, ds);
		return r * f * ds[which];
	}
};

PA_HD bool almost_equal(double x, double y) { return ::fabs(x - y) < 1e-10; }

// steep cores: the centroid derivatives of a cusp (core-Sersic with an inner
// slope, Sersic n > 4) cancel over many large samples, so the direct-pixel
// adjoint sums their FP32 per-pixel derivatives in FP64 per thread
PA_HD bool wants_grad64(const ProfileParams &p)
{
	return (p.type == CORESERSIC && p.shape[4] != 0) || (p.type == SERSIC && p.shape[1] > 4);
}

// initial_calculations() + the evaluate()-time centre shift, in FP64.
// offx/offy: PSF padding (drawing-canvas shift); sx/sy: profile pixel scale
// (image scale / finesampling); img_sx/img_sy: image pixel scale.
PA_HD int n_lum_integrals(int type) { return type == KING ? 3 : type == CORESERSIC ? 5 : type == BROKENEXP ? 4 : 0; }

// d ln(flux_scale) / d shape slot w of the families whose luminosity is a
// QUADPACK integral (King 3 slots, core-Sersic 5, broken exponential 4):
// -(2 pi int r df/dshape_w dr) / lumtot, from a component whose set-up has run.
template <bool kWave = false>
PA_HD double lum_gradient(const Comp &c, int w, sf::QuadState *qs)
{
	const double pi = 3.14159265358979323846;
	const double *s = c.s;
	double d = 0;
	switch (c.type) {
	case KING: {
		KingIntegrand f{s[0], s[1], s[2], w};
		d = 2 * pi * sf::integrate<kWave>(f, 0.0, s[1], false, *qs);
		break;
	}
	case CORESERSIC: {
		CoreSersicIntegrand f{&c, w};
		d = 2 * pi * sf::integrate<kWave>(f, 0.0, 0.0, true, *qs);
		break;
	}
	case BROKENEXP: {
		BrokenExpIntegrand f{s[0], s[1], s[2], s[3], w};
		d = 2 * pi * sf::integrate<kWave>(f, 0.0, 0.0, true, *qs);
		break;
	}
	default: return 0;
	}
	return -d / c.lumtot;
}

// grad_integrals = false leaves the slots of lum_gradient to a separate pass.
// kWave (device code): a whole wavefront runs the set-up with the same
// arguments, and the luminosity integrals spread their Kronrod points over
// its lanes (sf::kronrod).
template <bool kWave = false>
PA_HD int setup_component(const ProfileParams &p, double magzero, double sx, double sy, double img_sx,
                          double img_sy, unsigned offx, unsigned offy, sf::QuadState *qs, Comp *c,
                          bool grad_integrals = true)
{
	const int verr = validate_params(p);
	if (verr != SE_OK) return verr;
	const double *s = p.shape;
	const double pi = 3.14159265358979323846;
	if (!is_radial(p.type)) {
		// point source and sky: no radial set-up.  xc, yc hold the footprint
		// origins' centres, which libprofit shifts by offset.x * scale.x on
		// both axes (psf.cpp:79-82); s[0], s[1] the footprint ends' centres,
		// shifted by offset.y * scale.y; s[2] = 1 / fs^2 for the sky.
		*c = Comp{};
		c->type = p.type;
		c->convolve = p.convolve;
		c->grad_ok = 1;
		c->axrat = 1;
		c->xc = p.xcen + offx * sx;
		c->yc = p.ycen + offx * sx;
		c->s[0] = p.xcen + offy * sy;
		c->s[1] = p.ycen + offy * sy;
		for (int i = 0; i < NGRAD; ++i) c->dlnfs[i] = 0;
		if (p.type == POINTSRC) {
			c->flux_scale = ::pow(10.0, -0.4 * (p.mag - magzero));  // psf.cpp:66
			c->dlnfs[G_MAG] = -0.4 * ::log(10.0);
		}
		else if (p.type == SKY) {
			// sky.cpp:42-45: bg / (finesampling * finesampling)
			const double fs = ::round(img_sx / sx);
			c->s[2] = 1.0 / (fs * fs);
			c->flux_scale = s[0] * c->s[2];
		}
		return SE_OK;
	}
	c->type = p.type;
	c->rough = p.rough;
	c->adjust = p.adjust;
	c->convolve = p.convolve;
	c->box = p.box;
	c->boxy = p.box != 0;
	c->axrat = p.axrat;
	c->img_sx = img_sx;
	c->img_sy = img_sy;
	c->grad_ok = 1;
	c->grad64 = wants_grad64(p);
	c->grad_box = 1;
	for (int i = 0; i < 5; ++i) c->s[i] = s[i];
	for (int i = 0; i < NGRAD; ++i) c->dlnfs[i] = 0;
	c->bn = 0;
	c->dbn = 0;
	c->gk = 0;
	if (p.type == MOFFAT) {
		// rscale = fwhm / (2 sqrt(2^(1/con) - 1)): d ln rscale / d con
		const double tt = ::pow(2.0, 1 / s[1]);
		c->gk = 0.5 * tt * 0.69314718055994530942 / (s[1] * s[1] * (tt - 1));
	}

	// b_n = qgamma(0.5, 2n); a Sersic component's adjust rule needs two more
	// quantiles of the same gamma law (rscale_switch, and rscale_max when it is
	// 0).  In a wavefront each of the three runs on a lane of its own.
	double q_sw = 0, q_max = 0;
	if (p.type == SERSIC || p.type == CORESERSIC) {
		const double a2n = 2 * (p.type == SERSIC ? s[1] : s[2]);
		if constexpr (kWave) {
			const int l = __lane_id();
			const bool more = p.type == SERSIC && p.adjust;
			double pq = 0.5;
			if (more && l == 1) pq = 1. - s[1] * s[1] / 2e3;
			if (more && l == 2 && p.rscale_max == 0) pq = 0.9999;
			const double q = l < 3 ? sf::qgamma(pq, a2n) : 0.0;
			c->bn = __shfl(q, 0);
			q_sw = __shfl(q, 1);
			q_max = __shfl(q, 2);
		}
		else {
			c->bn = sf::qgamma(0.5, a2n);
		}
	}
	c->inv_re = c->inv_re2 = c->inv_n = c->inv_nB = c->inv_2n = 0;
	if (p.type == SERSIC) {
		c->inv_re = 1 / s[0];
		c->inv_re2 = 1 / (s[0] * s[0]);
		c->inv_n = 1 / s[1];
		c->inv_nB = 1 / (s[1] * (p.box + 2));
		c->inv_2n = 0.5 / s[1];
	}
	switch (p.type) {
	case SERSIC: c->rscale = s[0]; break;
	case MOFFAT: c->rscale = s[0] / (2 * ::sqrt(::pow(2.0, 1 / s[1]) - 1)); break;
	case FERRER: c->rscale = s[0]; break;
	case KING: c->rscale = s[1]; break;
	case CORESERSIC: c->rscale = s[0]; break;
	default: c->rscale = s[0] * 4; break;
	}

	double lumtot = 0;
	switch (p.type) {
	case SERSIC:
		lumtot = ::pow(s[0], 2.0) * 2 * pi * s[1] * sf::gam(2 * s[1]) * ::exp(c->bn) / ::pow(c->bn, 2 * s[1]);
		break;
	case MOFFAT: lumtot = ::pow(c->rscale, 2.0) * pi / (s[1] - 1); break;
	case FERRER: lumtot = ::pow(s[0], 2.0) * pi * (s[1] * sf::beta(s[1], 1 + 2 / (2 - s[2]))); break;
	case KING: {
		KingIntegrand f{s[0], s[1], s[2], -2};
		lumtot = 2 * pi * sf::integrate<kWave>(f, 0.0, s[1], false, *qs);
		break;
	}
	case CORESERSIC: {
		CoreSersicIntegrand f{c};
		lumtot = 2 * pi * sf::integrate<kWave>(f, 0.0, 0.0, true, *qs);
		break;
	}
	default: {
		BrokenExpIntegrand f{s[0], s[1], s[2], s[3]};
		lumtot = 2 * pi * sf::integrate<kWave>(f, 0.0, 0.0, true, *qs);
		break;
	}
	}
	c->lumtot = lumtot;
	const double b2 = p.box + 2;
	const double r_box = pi * b2 / (2 * sf::beta(1 / b2, 1 / b2));
	c->rbox = r_box;
	const double ie = ::pow(10.0, -0.4 * (p.mag - magzero)) / (lumtot * p.axrat / r_box);
	c->ie = ie;

	double acc = p.acc, rsw = p.rscale_switch, rmax = p.rscale_max;
	unsigned res = p.resolution;
	if (p.adjust) {
		switch (p.type) {
		case SERSIC: {
			const double n = s[1];
			const double ff = s[0] * ::pow((kWave ? q_sw : sf::qgamma(1. - n * n / 2e3, 2 * n)) / c->bn, n);
			double v = ::ceil(ff);
			v = ::fmax(::fmin(v, 20.), 2.);
			rsw = v / s[0];
			break;
		}
		case MOFFAT: {
			const double ff = c->rscale * ::sqrt(::pow(1 - 0.9999, 1 / (1 - s[1])) - 1);
			const double v = ::fmax(ff, 1.823 * s[0]);
			rsw = ::fmax(::fmin(v, 20.), 2.) / c->rscale;
			break;
		}
		case FERRER:
		case KING: rsw = 0.5; break;
		default: rsw = 1; break;
		}
		unsigned r = (unsigned)::ceil(160 / (rsw * c->rscale));
		r += r % 2;
		res = r < 4U ? 4U : (r > 16U ? 16U : r);
		if (rmax == 0) {
			switch (p.type) {
			case SERSIC: {
				const double n = s[1];
				const double ff = s[0] * ::pow((kWave ? q_max : sf::qgamma(0.9999, 2 * n)) / c->bn, n);
				rmax = ::ceil(::fmax(ff, 2.) / s[0]);
				break;
			}
			case MOFFAT: {
				const double ff = c->rscale * ::sqrt(::pow(1 - 0.9999, 1 / (1 - s[1])) - 1);
				rmax = ::ceil(::fmax(ff, 2.) / c->rscale);
				break;
			}
			case FERRER: rmax = 1; break;
			case KING: rmax = 1.5; break;
			default: rmax = 100; break;
			}
		}
		if (p.type == SERSIC) acc = ::fmax(0.1, acc / ::sqrt(s[1])) / p.axrat;
		else if (p.type == MOFFAT) acc = 0.1 / p.axrat;
	}
	c->acc = acc;
	c->rscale_switch = rsw;
	c->rscale_max = rmax;
	c->res = (int)res;
	c->maxrec = (int)p.max_recursions;
	if (p.max_recursions > 30) return SE_MAXREC;
	if (res > 16 || res == 0) return SE_RESOLUTION;
	c->has_rmax = rmax > 0;
	c->rmax_r = rmax * c->rscale;
	c->rsw_r = rsw * c->rscale;

	const double angrad = ::fmod(p.ang + 90, 360.) * pi / 180.;
	c->cos_ang = ::cos(angrad);
	c->sin_ang = ::sin(angrad);
	c->cos_over = c->cos_ang / p.axrat;
	c->sin_over = c->sin_ang / p.axrat;
	c->inv_axrat = 1 / p.axrat;

	double flux_scale = sx * sy * ie;
	if (p.type == SERSIC && p.rescale_flux) {
		double rescale = 1;
		if (p.adjust) rescale = 1 / sf::pgamma(c->bn * ::pow(rmax / s[0], 1 / s[1]), 2 * s[1]);
		flux_scale *= rescale;
	}
	c->flux_scale = flux_scale;
	c->xc = p.xcen + offx * sx;
	c->yc = p.ycen + offy * sy;
	c->centre_rule = p.type == SERSIC && p.adjust && s[1] > 1;

	c->kind = K_GENERAL;
	if (p.type == SERSIC) {
		double ie_ = s[1];
		if (p.box != 0) ie_ *= (p.box + 2);
		if (almost_equal(ie_, 0.5)) c->kind = K_HALF;
		else if (almost_equal(ie_, 1)) c->kind = K_ONE;
		else if (almost_equal(ie_, 2)) c->kind = K_TWO;
		else if (almost_equal(ie_, 3)) c->kind = K_THREE;
		else if (almost_equal(ie_, 4)) c->kind = K_FOUR;
		else if (almost_equal(ie_, 8)) c->kind = K_EIGHT;
		else if (almost_equal(ie_, 16)) c->kind = K_SIXTEEN;
	}

	// d ln(flux_scale)/d slot for the adjoint
	c->dlnfs[G_MAG] = -0.4 * 2.302585092994045684;
	c->dlnfs[G_AXRAT] = -1 / p.axrat;
	{
		const double a = 1 / b2;
		c->dlnfs[G_BOX] = 1 / b2 + 2 * (sf::digamma(a) - sf::digamma(2 * a)) / (b2 * b2);
	}
	switch (p.type) {
	case SERSIC: {
		const double n = s[1], bn = c->bn, a = 2 * n;
		const double dens = ::exp((a - 1) * ::log(bn) - bn - sf::lgam(a));
		c->dbn = -2 * sf::dpgamma_da(bn, a) / dens;
		const double dlnL = 1 / n + 2 * sf::digamma(2 * n) + c->dbn - 2 * ::log(bn) - 2 * n * c->dbn / bn;
		c->dlnfs[G_S0] = -2 / s[0];
		c->dlnfs[G_S1] = -dlnL;
		if (p.rescale_flux && p.adjust) {
			// rescale = 1 / P(x, 2n), x = b_n (rmax / re)^(1/n) (sersic.cpp:303-311),
			// with the adjusted rmax held fixed (it is piecewise constant):
			// d ln rescale = -(P'(x) dx + 2 dP/da dn) / P
			const double x = bn * ::pow(rmax / s[0], 1 / n);
			const double P = sf::pgamma(x, a);
			if (x > 0 && P > 0) {
				const double dens_x = ::exp((a - 1) * ::log(x) - x - sf::lgam(a));
				const double dx_dre = -x / (n * s[0]);
				const double dx_dn = x * (c->dbn / bn - ::log(rmax / s[0]) / (n * n));
				c->dlnfs[G_S0] += -dens_x * dx_dre / P;
				c->dlnfs[G_S1] += -(dens_x * dx_dn + 2 * sf::dpgamma_da(x, a)) / P;
			}
		}
		break;
	}
	case MOFFAT: {
		const double con = s[1], t = ::pow(2.0, 1 / con);
		const double dlnrs = 0.5 * t * 0.69314718055994530942 / (con * con * (t - 1));
		c->dlnfs[G_S0] = -2 / s[0];
		c->dlnfs[G_S1] = -(2 * dlnrs - 1 / (con - 1));
		break;
	}
	case FERRER: {
		const double a = s[1], b = s[2], cc = 1 + 2 / (2 - b);
		c->dlnfs[G_S0] = -2 / s[0];
		c->dlnfs[G_S1] = -(1 / a + sf::digamma(a) - sf::digamma(a + cc));
		c->dlnfs[G_S2] = -((sf::digamma(cc) - sf::digamma(a + cc)) * 2 / ((2 - b) * (2 - b)));
		break;
	}
	case KING:
	case BROKENEXP: {
		if (grad_integrals)
			for (int w = 0; w < n_lum_integrals(p.type); ++w) c->dlnfs[G_S0 + w] = lum_gradient<kWave>(*c, w, qs);
		break;
	}
	case CORESERSIC: {
		const double n = s[2], bn = c->bn, a = 2 * n;
		const double dens = ::exp((a - 1) * ::log(bn) - bn - sf::lgam(a));
		c->dbn = -2 * sf::dpgamma_da(bn, a) / dens;
		if (grad_integrals)
			for (int w = 0; w < 5; ++w) c->dlnfs[G_S0 + w] = lum_gradient<kWave>(*c, w, qs);
		break;
	}
	default: break;
	}
	Comp::F32 &f = c->f;
	f.cos_ang = (float)c->cos_ang;
	f.sin_ang = (float)c->sin_ang;
	f.cos_over = (float)c->cos_over;
	f.sin_over = (float)c->sin_over;
	f.axrat = (float)c->axrat;
	f.inv_axrat = (float)c->inv_axrat;
	f.inv_re = (float)c->inv_re;
	f.inv_re2 = (float)c->inv_re2;
	f.inv_n = (float)c->inv_n;
	f.inv_nB = (float)c->inv_nB;
	f.inv_2n = (float)c->inv_2n;
	f.bn = (float)c->bn;
	f.dbn = (float)c->dbn;
	f.acc = (float)c->acc;
	f.rscale = (float)c->rscale;
	f.box = (float)c->box;
	f.gk = (float)c->gk;
	f.rmax2 = (float)(c->rmax_r * c->rmax_r);
	f.rsw2 = (float)(c->rsw_r * c->rsw_r);
	for (int i = 0; i < 5; ++i) f.s[i] = (float)c->s[i];
	for (int i = 0; i < 5; ++i) f.aux[i] = 0.f;
	if (p.type == CORESERSIC) {
		f.aux[0] = (float)::log2(s[1]);
		f.aux[1] = (float)::pow(s[1], s[3]);
		f.aux[2] = (float)::pow(s[0], -s[3]);
		f.aux[3] = (float)(1 / (s[2] * s[3]));
		f.aux[4] = (float)::log2(s[0]);
	}
	else if (p.type == KING) {
		f.aux[0] = (float)(1 / s[0]);
		f.aux[1] = (float)(1 / s[2]);
		f.aux[2] = (float)::pow(1 + (s[1] / s[0]) * (s[1] / s[0]), -1 / s[2]);
		// the truncation terms of the derivatives: log2(1 + (rt/rc)^2), C / (1 + (rt/rc)^2)
		f.aux[3] = (float)::log2(1 + (s[1] / s[0]) * (s[1] / s[0]));
		f.aux[4] = (float)(::pow(1 + (s[1] / s[0]) * (s[1] / s[0]), -1 / s[2]) / (1 + (s[1] / s[0]) * (s[1] / s[0])));
	}
	else if (p.type == MOFFAT || p.type == FERRER) {
		f.aux[0] = (float)(1 / c->rscale);
	}
	return SE_OK;
}

#undef PA_HD
#undef PA_HDI

}  // namespace profit_amd
