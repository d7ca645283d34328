/*
 * THE ORACLE -- test infrastructure only (see oracle.h).
 *
 * FP64 restatement of libprofit's CPU path.  Reference files followed
 * (paths relative to src/libprofit/src/profit/):
 *   radial.cpp:46-52     profile coordinates
 *   radial.cpp:54-134    subsample_pixel (adaptive recursion)
 *   radial.cpp:136-199   initial_calculations (adjust)
 *   radial.cpp:204-214   validate
 *   radial.cpp:291-344   evaluate_cpu
 *   radial.h:93-102      boxy_r
 *   sersic.cpp:79-360    Sersic (pow-free templates, b_n, lumtot, adjust, centre rule)
 *   moffat.cpp:52-107, ferrer.cpp:50-125, king.cpp:51-118,
 *   coresersic.cpp:51-140, brokenexponential.cpp:38-149
 *   model.cpp:152-380    analyse / adjust / evaluate / produce_image
 *   image.cpp:91-162     Mask upsample / expand_by;  image.h:287-392 extend/crop/bbox
 *   image.cpp:283-296    downsample SUM
 *   convolve.cpp:55-99   Convolver::convolve, mask_and_crop
 *   convolve.cpp:101-156 BruteForceConvolver ("brute-old")
 *   convolve.cpp:158-231 AssociativeBruteForceConvolver ("brute")
 *   convolve.cpp:233-312 FFTConvolver (the maths equals brute force; the
 *                        oracle evaluates it directly in FP64 and keeps the
 *                        FFT convolver's crop/offset semantics)
 *   R/profitLikeModel.R:244-329 likelihoods
 * Build with -ffp-contract=off so every FP64 operation rounds as libprofit's
 * own x86-64 build does.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static __thread char g_err[512];

const char *or_last_error(void) { return g_err; }

static int fail(const char *msg)
{
	snprintf(g_err, sizeof(g_err), "%s", msg);
	return -1;
}

/* ------------------------------------------------------------------ */
/* profiles                                                             */
/* ------------------------------------------------------------------ */

void or_profile_defaults(int type, or_profile *p)
{
	memset(p, 0, sizeof(*p));
	p->type = type;
	p->convolve = 0;
	p->rough = 0;
	p->adjust = 1;
	p->rescale_flux = 0;
	p->xcen = 0; p->ycen = 0; p->mag = 15; p->ang = 0; p->axrat = 1; p->box = 0;
	p->acc = 0.1; p->rscale_switch = 1; p->rscale_max = 0;
	p->resolution = 9; p->max_recursions = 2;
	switch (type) {
	case OR_SERSIC: p->shape[0] = 1; p->shape[1] = 1; break;
	case OR_MOFFAT: p->shape[0] = 3; p->shape[1] = 2; break;
	case OR_FERRER: p->shape[0] = 3; p->shape[1] = 1; p->shape[2] = 1; p->acc = 1; break;
	case OR_KING: p->shape[0] = 1; p->shape[1] = 3; p->shape[2] = 2; break;
	case OR_CORESERSIC: p->shape[0] = 1; p->shape[1] = 1; p->shape[2] = 4; p->shape[3] = 1; p->shape[4] = 1; break;
	case OR_BROKENEXP: p->shape[0] = 1; p->shape[1] = 1; p->shape[2] = 1; p->shape[3] = 1; break;
	case OR_POINT: p->mag = 0; break; /* psf.cpp:147-156 */
	case OR_SKY: break;               /* sky.cpp:65-71: bg = 0 */
	}
}

enum { K_PF = 0, K_ONE, K_TWO, K_THREE, K_FOUR, K_EIGHT, K_SIXTEEN, K_GENERAL };

typedef struct comp {
	or_profile p;          /* with adjusted acc/rscale_switch/rscale_max/resolution */
	double rscale, ie, bn, lumtot, rbox, rescale_factor, flux_scale;
	double cos_ang, sin_ang, cos_over, sin_over;
	double xc, yc;         /* centre after the PSF-padding offset */
	double img_sx, img_sy; /* Model image pixel scale (Sersic centre rule) */
	int sersic_kind;
	double dlnfs[OR_NGRAD];/* d ln(flux_scale) / d slot */
	/* point sources: the model's normalised PSF, its pixel scale and the padding offset */
	const double *psf;
	unsigned kw, kh, off_x, off_y;
	double psf_sx, psf_sy;
	double inv_fs2;        /* sky: 1 / finesampling^2 */
} comp;

static int almost_equals(double x, double y) { return fabs(x - y) < 1e-10; }

static double boxy_r(const comp *c, double x, double y)
{
	if (c->p.box == 0) return sqrt(x * x + y * y);
	double b2 = c->p.box + 2.;
	return pow(pow(fabs(x), b2) + pow(fabs(y), b2), 1. / b2);
}

static double sersic_eval(const comp *c, double x, double y)
{
	double re = c->p.shape[0], nser = c->p.shape[1], box = c->p.box, bn = c->bn;
	double B = box + 2, base, rf;
	int boxy = box != 0;
	if (boxy) base = pow(fabs(x / re), B) + pow(fabs(y / re), B);
	else base = (x * x + y * y) / (re * re);
	double invexp = boxy ? nser * B : nser;
	if (boxy) {
		switch (c->sersic_kind) {
		case K_PF: rf = base * base; break;
		case K_ONE: rf = base; break;
		case K_TWO: rf = sqrt(base); break;
		case K_THREE: rf = cbrt(base); break;
		case K_FOUR: rf = sqrt(sqrt(base)); break;
		case K_EIGHT: rf = sqrt(sqrt(sqrt(base))); break;
		case K_SIXTEEN: rf = sqrt(sqrt(sqrt(sqrt(base)))); break;
		default: rf = pow(base, 1 / invexp);
		}
	}
	else {
		switch (c->sersic_kind) {
		case K_PF: rf = base; break;
		case K_ONE: rf = sqrt(base); break;
		case K_TWO: rf = sqrt(sqrt(base)); break;
		case K_THREE: rf = cbrt(sqrt(base)); break;
		case K_FOUR: rf = sqrt(sqrt(sqrt(base))); break;
		case K_EIGHT: rf = sqrt(sqrt(sqrt(sqrt(base)))); break;
		case K_SIXTEEN: rf = sqrt(sqrt(sqrt(sqrt(sqrt(base))))); break;
		default: rf = pow(sqrt(base), 1 / invexp);
		}
	}
	return exp(-bn * (rf - 1));
}

static double king_core(double r, double rc, double rt, double a)
{
	return pow(1 / pow(1 + pow(r / rc, 2), 1 / a) - 1 / pow(1 + pow(rt / rc, 2), 1 / a), a);
}

static double coresersic_core(const comp *c, double r)
{
	double re = c->p.shape[0], rb = c->p.shape[1], nser = c->p.shape[2], a = c->p.shape[3], b = c->p.shape[4];
	return pow(1 + pow(r / rb, -a), b / a) * exp(-c->bn * pow((pow(r, a) + pow(rb, a)) / pow(re, a), 1 / (nser * a)));
}

static double brokenexp_core(double r, double h1, double h2, double rb, double a)
{
	double base = r - rb;
	double expo = 1 / h1 - 1 / h2;
	if (a * base < 40) base = log(1 + exp(a * base)) / a;
	return exp(-r / h1 + expo * base);
}

static double eval_at(const comp *c, double x, double y)
{
	switch (c->p.type) {
	case OR_SERSIC: return sersic_eval(c, x, y);
	case OR_MOFFAT: {
		double r = boxy_r(c, x, y);
		double rf = r / c->rscale;
		return pow(1 + rf * rf, -c->p.shape[1]);
	}
	case OR_FERRER: {
		double r = boxy_r(c, x, y);
		double rf = r / c->rscale;
		if (rf < 1) return pow(1 - pow(rf, 2 - c->p.shape[2]), c->p.shape[1]);
		return 0;
	}
	case OR_KING: {
		double r = boxy_r(c, x, y);
		if (r < c->p.shape[1]) return king_core(r, c->p.shape[0], c->p.shape[1], c->p.shape[2]);
		return 0;
	}
	case OR_CORESERSIC: return coresersic_core(c, boxy_r(c, x, y));
	case OR_BROKENEXP:
		return brokenexp_core(boxy_r(c, x, y), c->p.shape[0], c->p.shape[1], c->p.shape[2], c->p.shape[3]);
	}
	return 0;
}

/* integrands for the numerically integrated luminosities */
static double king_integrand(double r, void *ctx)
{
	const comp *c = ctx;
	if (r < c->p.shape[1]) return r * king_core(r, c->p.shape[0], c->p.shape[1], c->p.shape[2]);
	return 0;
}
static double coresersic_integrand(double r, void *ctx) { return r * coresersic_core(ctx, r); }
static double brokenexp_integrand(double r, void *ctx)
{
	const comp *c = ctx;
	return r * brokenexp_core(r, c->p.shape[0], c->p.shape[1], c->p.shape[2], c->p.shape[3]);
}

static int validate(const or_profile *p)
{
	const double *s = p->shape;
	if (p->type == OR_POINT || p->type == OR_SKY) return 0; /* psf.cpp:40-45 is checked in analyze_inputs */
	if (p->axrat <= 0) return fail("axrat <= 0, must have axrat > 0");
	if (p->axrat > 1) return fail("axrat > 1, must have axrat <= 1");
	if (p->box <= -2) return fail("box <= -2, must have box > -2");
	switch (p->type) {
	case OR_SERSIC:
		if (s[0] <= 0) return fail("re <= 0, must have re > 0");
		if (s[1] <= 0) return fail("nser <= 0, must have nser > 0");
		break;
	case OR_MOFFAT:
		if (s[0] <= 0) return fail("fwhm <= 0, must have fwhm > 0");
		if (s[1] < 0) return fail("con < 0, must have con >= 0");
		break;
	case OR_FERRER:
		if (s[0] <= 0) return fail("rout <= 0, must have rout >= 0");
		if (s[1] < 0) return fail("a < 0, must have a >= 0");
		if (s[2] > 2) return fail("b > 2, must have b <= 2");
		break;
	case OR_KING:
		if (s[0] <= 0) return fail("rc <= 0, must have rc > 0");
		if (s[1] <= 0) return fail("rt <= 0, must have rt > 0");
		if (s[2] < 0) return fail("a < 0, must have a >=0");
		break;
	case OR_CORESERSIC:
		if (s[0] <= 0) return fail("re <= 0, must have re > 0");
		if (s[1] <= 0) return fail("rb <= 0, must have rb > 0");
		if (s[2] <= 0) return fail("nser <= 0, must have nser > 0");
		if (s[3] <= 0) return fail("a <= 0, must have a > 0");
		if (s[4] > 1.999) return fail("b > 1.999, must have b < 1.999");
		break;
	case OR_BROKENEXP:
		if (s[0] <= 0) return fail("h1 <= 0, must have h1 > 0");
		if (s[1] <= 0) return fail("h2 <= 0, must have h2 > 0");
		if (s[2] <= 0) return fail("rb <= 0, must have rb > 0");
		break;
	default:
		return fail("unknown profile type");
	}
	return 0;
}

static double sersic_fluxfrac(const comp *c, double fraction)
{
	double re = c->p.shape[0], nser = c->p.shape[1];
	double ratio = or_qgamma(fraction, 2 * nser) / c->bn;
	return re * pow(ratio, nser);
}

static double moffat_fluxfrac(const comp *c, double fraction)
{
	double con = c->p.shape[1];
	return c->rscale * sqrt(pow(1 - fraction, 1 / (1 - con)) - 1);
}

/* d ln(lumtot) / d shape_i, fixed-tree gradient helpers */
typedef struct king_dctx { const comp *c; int which; } king_dctx;
static double king_dintegrand(double r, void *ctx);

/* RadialProfile::initial_calculations and the subclass overrides */
static void setup_comp(comp *c, const or_profile *p, double magzero, double sx, double sy,
                       double img_sx, double img_sy, unsigned off_x, unsigned off_y)
{
	memset(c, 0, sizeof(*c));
	c->p = *p;
	c->off_x = off_x;
	c->off_y = off_y;
	if (p->type == OR_POINT) {
		/* psf.cpp:66: scale = 10^(-0.4 (mag - magzero)) */
		c->flux_scale = pow(10, -0.4 * (p->mag - magzero));
		c->dlnfs[OR_G_MAG] = -0.4 * log(10.0);
		return;
	}
	if (p->type == OR_SKY) return; /* bg / fs^2 is set by prepare() */
	c->img_sx = img_sx;
	c->img_sy = img_sy;
	const double *s = p->shape;
	or_profile *q = &c->p;

	if (p->type == OR_SERSIC || p->type == OR_CORESERSIC) {
		double nser = p->type == OR_SERSIC ? s[1] : s[2];
		c->bn = or_qgamma(0.5, 2 * nser);
	}
	switch (p->type) {
	case OR_SERSIC: c->rscale = s[0]; break;
	case OR_MOFFAT: c->rscale = s[0] / (2 * sqrt(pow(2, (1 / s[1])) - 1)); break;
	case OR_FERRER: c->rscale = s[0]; break;
	case OR_KING: c->rscale = s[1]; break;
	case OR_CORESERSIC: c->rscale = s[0]; break;
	case OR_BROKENEXP: c->rscale = s[0] * 4; break;
	}

	double lumtot = 0;
	switch (p->type) {
	case OR_SERSIC: {
		double g_factor = or_gammafn(2 * s[1]);
		lumtot = pow(s[0], 2) * 2 * M_PI * s[1] * g_factor * exp(c->bn) / pow(c->bn, 2 * s[1]);
		break;
	}
	case OR_MOFFAT: lumtot = pow(c->rscale, 2) * M_PI / (s[1] - 1); break;
	case OR_FERRER: {
		double g_factor = s[1] * or_beta(s[1], 1 + 2 / (2 - s[2]));
		lumtot = pow(s[0], 2) * M_PI * g_factor;
		break;
	}
	case OR_KING: lumtot = 2 * M_PI * or_integrate_qags(king_integrand, c, 0, s[1], NULL); break;
	case OR_CORESERSIC: lumtot = 2 * M_PI * or_integrate_qagi(coresersic_integrand, c, 0, NULL); break;
	case OR_BROKENEXP: lumtot = 2 * M_PI * or_integrate_qagi(brokenexp_integrand, c, 0, NULL); break;
	}
	c->lumtot = lumtot;

	double b2 = p->box + 2;
	double r_box = M_PI * b2 / (2 * or_beta(1 / b2, 1 / b2));
	c->rbox = r_box;
	double lumtot_eff = lumtot * p->axrat / r_box;
	c->ie = pow(10, -0.4 * (p->mag - magzero)) / lumtot_eff;

	if (p->adjust) {
		double rsw = 0;
		switch (p->type) {
		case OR_SERSIC: {
			double nser = s[1];
			rsw = ceil(sersic_fluxfrac(c, 1. - nser * nser / 2e3));
			rsw = fmax(fmin(rsw, 20.), 2.);
			rsw = rsw / s[0];
			break;
		}
		case OR_MOFFAT: {
			rsw = fmax(moffat_fluxfrac(c, 0.9999), 1.823 * s[0]);
			rsw = fmax(fmin(rsw, 20.), 2.) / c->rscale;
			break;
		}
		case OR_FERRER: rsw = 0.5; break;
		case OR_KING: rsw = 0.5; break;
		case OR_CORESERSIC: rsw = 1; break;
		case OR_BROKENEXP: rsw = 1; break;
		}
		q->rscale_switch = rsw;
		unsigned res = (unsigned)ceil(160 / (q->rscale_switch * c->rscale));
		res += res % 2;
		res = res < 4U ? 4U : (res > 16U ? 16U : res);
		q->resolution = res;
		if (q->rscale_max == 0) {
			switch (p->type) {
			case OR_SERSIC: q->rscale_max = ceil(fmax(sersic_fluxfrac(c, 0.9999), 2.) / s[0]); break;
			case OR_MOFFAT: q->rscale_max = ceil(fmax(moffat_fluxfrac(c, 0.9999), 2.) / c->rscale); break;
			case OR_FERRER: q->rscale_max = 1; break;
			case OR_KING: q->rscale_max = 1.5; break;
			case OR_CORESERSIC: q->rscale_max = 100; break;
			case OR_BROKENEXP: q->rscale_max = 100; break;
			}
		}
		switch (p->type) {
		case OR_SERSIC: {
			double acc = q->acc / sqrt(s[1]);
			q->acc = fmax(0.1, acc) / p->axrat;
			break;
		}
		case OR_MOFFAT: q->acc = 0.1 / p->axrat; break;
		default: break;
		}
	}

	double angrad = fmod(p->ang + 90, 360.) * M_PI / 180.;
	c->cos_ang = cos(angrad);
	c->sin_ang = sin(angrad);
	c->cos_over = c->cos_ang / p->axrat;
	c->sin_over = c->sin_ang / p->axrat;

	c->rescale_factor = 1;
	if (p->type == OR_SERSIC && p->adjust && p->rescale_flux) {
		double flux_r = c->bn * pow(q->rscale_max / s[0], 1 / s[1]);
		flux_r = or_pgamma(flux_r, 2 * s[1]);
		c->rescale_factor = 1 / flux_r;
	}
	double pixel_area = sx * sy;
	c->flux_scale = pixel_area * c->ie;
	if (p->type == OR_SERSIC && p->rescale_flux) c->flux_scale *= c->rescale_factor;

	c->xc = p->xcen + off_x * sx;
	c->yc = p->ycen + off_y * sy;

	if (p->type == OR_SERSIC) {
		double inv_exponent = s[1];
		if (p->box != 0) inv_exponent *= (p->box + 2);
		if (almost_equals(inv_exponent, 0.5)) c->sersic_kind = K_PF;
		else if (almost_equals(inv_exponent, 1)) c->sersic_kind = K_ONE;
		else if (almost_equals(inv_exponent, 2)) c->sersic_kind = K_TWO;
		else if (almost_equals(inv_exponent, 3)) c->sersic_kind = K_THREE;
		else if (almost_equals(inv_exponent, 4)) c->sersic_kind = K_FOUR;
		else if (almost_equals(inv_exponent, 8)) c->sersic_kind = K_EIGHT;
		else if (almost_equals(inv_exponent, 16)) c->sersic_kind = K_SIXTEEN;
		else c->sersic_kind = K_GENERAL;
	}

	/* d ln(flux_scale) / d slot (used by the analytic gradient) */
	c->dlnfs[OR_G_MAG] = -0.4 * log(10.0);
	c->dlnfs[OR_G_AXRAT] = -1 / p->axrat;
	{
		double a = 1 / b2;
		c->dlnfs[OR_G_BOX] = 1 / b2 + 2 * (or_digamma(a) - or_digamma(2 * a)) / (b2 * b2);
	}
	switch (p->type) {
	case OR_SERSIC: {
		double re = s[0], n = s[1], bn = c->bn, a = 2 * n;
		double dens = exp((a - 1) * log(bn) - bn - lgamma(a));
		double dbn = -2 * or_dpgamma_da(bn, a) / dens;
		double dlnL_dn = 1 / n + 2 * or_digamma(2 * n) + dbn - 2 * log(bn) - 2 * n * dbn / bn;
		c->dlnfs[OR_G_S0] = -2 / re;
		c->dlnfs[OR_G_S1] = -dlnL_dn;
		break;
	}
	case OR_MOFFAT: {
		double fwhm = s[0], con = s[1];
		double t = pow(2, 1 / con);
		double dlnrs_dcon = 0.5 * t * log(2.0) / (con * con * (t - 1));
		c->dlnfs[OR_G_S0] = -2 / fwhm;
		c->dlnfs[OR_G_S1] = -(2 * dlnrs_dcon - 1 / (con - 1));
		break;
	}
	case OR_FERRER: {
		double rout = s[0], a = s[1], b = s[2], cc = 1 + 2 / (2 - b);
		c->dlnfs[OR_G_S0] = -2 / rout;
		c->dlnfs[OR_G_S1] = -(1 / a + or_digamma(a) - or_digamma(a + cc));
		c->dlnfs[OR_G_S2] = -((or_digamma(cc) - or_digamma(a + cc)) * 2 / ((2 - b) * (2 - b)));
		break;
	}
	case OR_KING: {
		for (int w = 0; w < 3; w++) {
			king_dctx kc = {c, w};
			double d = 2 * M_PI * or_integrate_qags(king_dintegrand, &kc, 0, s[1], NULL);
			c->dlnfs[OR_G_S0 + w] = -d / lumtot;
		}
		break;
	}
	default: break;
	}
}

/* ------------------------------------------------------------------ */
/* per-sample derivatives (fixed-tree gradient)                         */
/* ------------------------------------------------------------------ */

/* out: [0] f, [1] df/dx_prof, [2] df/dy_prof, [3] df/dbox, [4..8] df/dshape */
static void dr_dxyB(const comp *c, double x, double y, double r, double *drx, double *dry, double *drB)
{
	double B = c->p.box + 2;
	if (r == 0) { *drx = 0; *dry = 0; *drB = 0; return; }
	double ax = fabs(x), ay = fabs(y);
	if (c->p.box == 0) {
		*drx = x / r;
		*dry = y / r;
	}
	else {
		double rp = pow(r, 1 - B);
		*drx = ax > 0 ? rp * pow(ax, B - 1) * (x > 0 ? 1 : -1) : 0;
		*dry = ay > 0 ? rp * pow(ay, B - 1) * (y > 0 ? 1 : -1) : 0;
	}
	double S = pow(ax, B) + pow(ay, B);
	double Sl = (ax > 0 ? pow(ax, B) * log(ax) : 0) + (ay > 0 ? pow(ay, B) * log(ay) : 0);
	*drB = r * (Sl / (B * S) - log(S) / (B * B));
}

static double king_dcore(double r, double rc, double rt, double a, int which)
{
	double u2 = (r / rc) * (r / rc), w2 = (rt / rc) * (rt / rc);
	double A = pow(1 + u2, -1 / a), C = pow(1 + w2, -1 / a), D = A - C;
	if (!(D > 0)) return 0;
	double Da1 = a * pow(D, a - 1);
	switch (which) {
	case -1: return Da1 * (-(2 * r / (a * rc * rc)) * A / (1 + u2)); /* d/dr */
	case 0: return Da1 * ((2 * u2 / (a * rc)) * A / (1 + u2) - (2 * w2 / (a * rc)) * C / (1 + w2));
	case 1: return Da1 * ((2 * rt / (a * rc * rc)) * C / (1 + w2));
	case 2: return pow(D, a) * (log(D) + a * (A * log(1 + u2) - C * log(1 + w2)) / (a * a) / D);
	}
	return 0;
}

static double king_dintegrand(double r, void *ctx)
{
	const king_dctx *k = ctx;
	const double *s = k->c->p.shape;
	if (r < s[1]) return r * king_dcore(r, s[0], s[1], s[2], k->which);
	return 0;
}

static int sample_grad(const comp *c, double x, double y, double *o)
{
	const double *s = c->p.shape;
	for (int i = 0; i < 9; i++) o[i] = 0;
	switch (c->p.type) {
	case OR_SERSIC: {
		double re = s[0], n = s[1], bn = c->bn, B = c->p.box + 2;
		double u = fabs(x / re), v = fabs(y / re);
		double base = pow(u, B) + pow(v, B);
		double f = sersic_eval(c, x, y);
		o[0] = f;
		if (base == 0) return 0;
		double rf = pow(base, 1 / (n * B));
		double drf_dbase = rf / (n * B * base);
		double dbase_dx = u > 0 ? B * pow(u, B - 1) / re * (x > 0 ? 1 : -1) : 0;
		double dbase_dy = v > 0 ? B * pow(v, B - 1) / re * (y > 0 ? 1 : -1) : 0;
		o[1] = -bn * f * drf_dbase * dbase_dx;
		o[2] = -bn * f * drf_dbase * dbase_dy;
		double dbase_dB = (u > 0 ? pow(u, B) * log(u) : 0) + (v > 0 ? pow(v, B) * log(v) : 0);
		double drf_dB = rf * (dbase_dB / (n * B * base) - log(base) / (n * B * B));
		o[3] = -bn * f * drf_dB;
		o[4] = bn * f * rf / (n * re);
		double a = 2 * n;
		double dens = exp((a - 1) * log(bn) - bn - lgamma(a));
		double dbn = -2 * or_dpgamma_da(bn, a) / dens;
		double drf_dn = -rf * log(base) / (n * n * B);
		o[5] = f * (-dbn * (rf - 1) - bn * drf_dn);
		return 0;
	}
	case OR_MOFFAT: {
		double r = boxy_r(c, x, y), rs = c->rscale, con = s[1];
		double t = r / rs, q = 1 + t * t;
		double f = pow(q, -con);
		o[0] = f;
		double dfdr = -con * pow(q, -con - 1) * 2 * t / rs;
		double drx, dry, drB;
		dr_dxyB(c, x, y, r, &drx, &dry, &drB);
		o[1] = dfdr * drx; o[2] = dfdr * dry; o[3] = dfdr * drB;
		double dfdrs = -dfdr * r / rs;
		double tt = pow(2, 1 / con);
		double dlnrs_dcon = 0.5 * tt * log(2.0) / (con * con * (tt - 1));
		o[4] = dfdrs * rs / s[0];
		o[5] = -f * log(q) + dfdrs * rs * dlnrs_dcon;
		return 0;
	}
	case OR_FERRER: {
		double r = boxy_r(c, x, y), rout = s[0], a = s[1], b = s[2];
		double t = r / rout;
		if (!(t < 1)) return 0;
		double g = 1 - pow(t, 2 - b);
		double f = pow(g, a);
		o[0] = f;
		double drx, dry, drB;
		dr_dxyB(c, x, y, r, &drx, &dry, &drB);
		double dfdr = r > 0 ? a * pow(g, a - 1) * (-(2 - b) * pow(t, 1 - b)) / rout : 0;
		o[1] = dfdr * drx; o[2] = dfdr * dry; o[3] = dfdr * drB;
		o[4] = -dfdr * r / rout;
		o[5] = g > 0 ? f * log(g) : 0;
		o[6] = t > 0 ? a * pow(g, a - 1) * pow(t, 2 - b) * log(t) : 0;
		return 0;
	}
	case OR_KING: {
		double r = boxy_r(c, x, y);
		if (!(r < s[1])) return 0;
		o[0] = king_core(r, s[0], s[1], s[2]);
		double drx, dry, drB;
		dr_dxyB(c, x, y, r, &drx, &dry, &drB);
		double dfdr = king_dcore(r, s[0], s[1], s[2], -1);
		o[1] = dfdr * drx; o[2] = dfdr * dry; o[3] = dfdr * drB;
		for (int w = 0; w < 3; w++) o[4 + w] = king_dcore(r, s[0], s[1], s[2], w);
		return 0;
	}
	default:
		return -1;
	}
}

/* accumulate weight * d f(sample) / d slot into g[OR_NGRAD] (shape part only;
 * the flux-scale part is added per pixel) */
static void accumulate_sample_grad(const comp *c, double x, double y, double w, double *g)
{
	double xr = x - c->xc, yr = y - c->yc;
	double xp = xr * c->cos_ang + yr * c->sin_ang;
	double yp = -xr * c->sin_over + yr * c->cos_over;
	double o[9];
	sample_grad(c, xp, yp, o);
	double q = c->p.axrat;
	double fx = o[1], fy = o[2];
	g[OR_G_XCEN] += w * (fx * -c->cos_ang + fy * (c->sin_ang / q));
	g[OR_G_YCEN] += w * (fx * -c->sin_ang + fy * (-c->cos_ang / q));
	g[OR_G_ANG] += w * (fx * (q * yp) + fy * (-xp / q)) * (M_PI / 180.);
	g[OR_G_AXRAT] += w * (fy * (-yp / q));
	g[OR_G_BOX] += w * o[3];
	for (int i = 0; i < 5; i++) g[OR_G_S0 + i] += w * o[4 + i];
}

/* ------------------------------------------------------------------ */
/* rendering                                                            */
/* ------------------------------------------------------------------ */

typedef struct tape_px {
	unsigned char cls, res, maxrec;
	unsigned n, cap, pos;
	unsigned char *bits;
} tape_px;

typedef struct tape {
	tape_px *px;
	unsigned npix;
} tape;

enum { TAPE_OFF = 0, TAPE_RECORD = 1, TAPE_REPLAY = 2 };

typedef struct walk {
	const comp *c;
	tape_px *tp;
	int mode;
	int nodes;
	double *g;       /* gradient accumulator or NULL */
	double gw;       /* weight of the current level in the pixel average */
} walk;

static int tape_next(walk *w, int computed)
{
	if (w->mode == TAPE_RECORD) {
		tape_px *t = w->tp;
		if (t->n == t->cap) {
			t->cap = t->cap ? 2 * t->cap : 256;
			t->bits = realloc(t->bits, t->cap);
		}
		t->bits[t->n++] = (unsigned char)computed;
		return computed;
	}
	if (w->mode == TAPE_REPLAY) {
		tape_px *t = w->tp;
		if (t->pos < t->n) return t->bits[t->pos++];
		return computed;
	}
	return computed;
}

/* radial.cpp:54-134, with the same FP64 operation sequence */
static double subsample_pixel(walk *w, double x0, double x1, double y0, double y1,
                              unsigned recur_level, unsigned max_recursions, unsigned resolution)
{
	const comp *c = w->c;
	double xbin = (x1 - x0) / resolution;
	double ybin = (y1 - y0) / resolution;
	double half_xbin = xbin / 2.;
	double half_ybin = ybin / 2.;
	double total = 0;
	int recurse = resolution > 1 && recur_level < max_recursions;
	unsigned npts = 0;
	double *pts = NULL;
	double level_w = w->gw / ((double)resolution * resolution);
	w->nodes++;
	if (recurse) pts = malloc(sizeof(double) * 2 * resolution * resolution);

	double x = x0;
	for (unsigned i = 0; i < resolution; i++) {
		x += half_xbin;
		double y = y0;
		for (unsigned j = 0; j < resolution; j++) {
			y += half_ybin;
			double xr = x - c->xc, yr = y - c->yc;
			double x_prof = xr * c->cos_ang + yr * c->sin_ang;
			double y_prof = -xr * c->sin_over + yr * c->cos_over;
			double subval = eval_at(c, x_prof, y_prof);
			int descend = 0;
			if (recurse) {
				double delta_y_prof = -xbin * c->sin_over + ybin * c->cos_over;
				double testval = eval_at(c, fabs(x_prof), fabs(y_prof) + fabs(delta_y_prof));
				descend = tape_next(w, fabs(testval - subval) > c->p.acc * subval);
			}
			if (descend) {
				pts[2 * npts] = x;
				pts[2 * npts + 1] = y;
				npts++;
			}
			else {
				total += subval;
				if (w->g) accumulate_sample_grad(c, x, y, level_w, w->g);
			}
			y += half_ybin;
		}
		x += half_xbin;
	}
	double saved = w->gw;
	w->gw = level_w;
	for (unsigned k = 0; k < npts; k++) {
		double px = pts[2 * k], py = pts[2 * k + 1];
		total += subsample_pixel(w, px - half_xbin, px + half_xbin, py - half_ybin, py + half_ybin,
		                         recur_level + 1, max_recursions, resolution);
	}
	w->gw = saved;
	free(pts);
	return total / (resolution * resolution);
}

/*
 * RadialProfile::evaluate_cpu (radial.cpp:291-344).  img is the drawing
 * canvas (accumulated into).  When G is given, instead of rendering the
 * function accumulates sum_q G[q] * d(flux_scale * pixel_value_q)/d slot
 * into gsum[OR_NGRAD].
 */
/* bind() of psf.cpp:48-56: floor of the value, clamped to [0, max] */
static unsigned psf_bind(double v, unsigned max)
{
	double f = floor(v);
	if (f < 0) return 0;
	if (f > (double)max) return max;
	return (unsigned)f;
}

/*
 * PsfProfile::evaluate (psf.cpp:58-145): each canvas pixel of the PSF's
 * footprint receives the PSF pixels it overlaps, weighted by the overlap
 * area.  The reference's index bounds are kept as written (the x end of the
 * footprint uses offset.y, the second x index is clamped by the PSF height,
 * the first y index by its width); a flat PSF index beyond the array, which
 * only those bounds can produce for a non-square PSF, reads 0.  With G the
 * pass accumulates sum_q G(q) d(scale val(q))/d(xcen, ycen, mag) instead: the
 * overlap lengths are piecewise linear in the centre, with derivative
 * [psf_x + psx < x + sx] - [psf_x > x] along x.
 */
static void render_point(const comp *c, double *img, unsigned W, unsigned H, double sx, double sy,
                         const double *G, double *gsum)
{
	const double psx = c->psf_sx, psy = c->psf_sy;
	const unsigned kw = c->kw, kh = c->kh;
	const double scale = c->flux_scale;
	const double origin_x = c->p.xcen + c->off_x * sx - kw * psx / 2;
	const double end_x = c->p.xcen + c->off_y * sy + kw * psx / 2;
	const double origin_y = c->p.ycen + c->off_x * sx - kh * psy / 2;
	const double end_y = c->p.ycen + c->off_y * sy + kh * psy / 2;
	const unsigned x0 = psf_bind(origin_x / sx, W - 1), y0 = psf_bind(origin_y / sy, H - 1);
	const unsigned x1 = psf_bind(end_x / sx, W - 1), y1 = psf_bind(end_y / sy, H - 1);
	double g[3] = {0, 0, 0};
	for (unsigned pix_y = y0; pix_y <= y1; pix_y++) {
		const double y = pix_y * sy;
		for (unsigned pix_x = x0; pix_x <= x1; pix_x++) {
			const double x = pix_x * sx;
			const unsigned qx0 = psf_bind(floor((x - origin_x) / psx), kw - 1);
			const unsigned qx1 = psf_bind(floor((x - origin_x + sx) / psx), kh - 1);
			const unsigned qy0 = psf_bind(floor((y - origin_y) / psy), kw - 1);
			const unsigned qy1 = psf_bind(floor((y - origin_y + sy) / psy), kh - 1);
			double val = 0, dvx = 0, dvy = 0;
			for (unsigned qy = qy0; qy <= qy1; qy++) {
				const double psf_y = qy * psy + origin_y;
				const double iy = fmin(y + sy, psf_y + psy) - fmax(y, psf_y);
				const double diy = (psf_y + psy < y + sy ? 1.0 : 0.0) - (psf_y > y ? 1.0 : 0.0);
				for (unsigned qx = qx0; qx <= qx1; qx++) {
					const double psf_x = qx * psx + origin_x;
					const double ix = fmin(x + sx, psf_x + psx) - fmax(x, psf_x);
					const double dix = (psf_x + psx < x + sx ? 1.0 : 0.0) - (psf_x > x ? 1.0 : 0.0);
					const size_t k = qx + (size_t)qy * kw;
					const double w = k < (size_t)kw * kh ? c->psf[k] : 0.0;
					val += w * (ix * iy) / (psx * psy);
					dvx += w * (dix * iy) / (psx * psy);
					dvy += w * (ix * diy) / (psx * psy);
				}
			}
			const size_t idx = pix_x + (size_t)pix_y * W;
			if (G) {
				g[0] += G[idx] * scale * dvx;
				g[1] += G[idx] * scale * dvy;
				g[2] += G[idx] * scale * val;
			}
			else {
				img[idx] += val * scale;
			}
		}
	}
	if (G && gsum) {
		gsum[OR_G_XCEN] += g[0];
		gsum[OR_G_YCEN] += g[1];
		gsum[OR_G_MAG] += g[2] * c->dlnfs[OR_G_MAG];
	}
}

/* SkyProfile::evaluate (sky.cpp:47-63): bg / fs^2 on every pixel of the calc
 * mask; with G, sum_q G(q) / fs^2 into the bg slot */
static void render_sky(const comp *c, double *img, const unsigned char *mask, unsigned W, unsigned H,
                       const double *G, double *gsum)
{
	double gs = 0;
	for (size_t idx = 0; idx < (size_t)W * H; idx++) {
		if (mask && !mask[idx]) continue;
		if (G) gs += G[idx];
		else img[idx] += c->flux_scale;
	}
	if (G && gsum) gsum[OR_G_S0] += gs * c->inv_fs2;
}

static void render_comp(const comp *c, double *img, const unsigned char *mask, unsigned W,
                        unsigned H, double sx, double sy, int threads, or_diag *diag, int plane,
                        tape *tp, int tape_mode, const double *G, double *gsum)
{
	if (c->p.type == OR_POINT) {
		render_point(c, img, W, H, sx, sy, G, gsum);
		return;
	}
	if (c->p.type == OR_SKY) {
		render_sky(c, img, mask, W, H, G, gsum);
		return;
	}
	double half_xbin = sx / 2.;
	double half_ybin = sy / 2.;
	double flux_scale = c->flux_scale;
	int nthreads = threads < 1 ? 1 : threads;
	double gtot[OR_NGRAD];
	memset(gtot, 0, sizeof(gtot));

#pragma omp parallel num_threads(nthreads) if (nthreads > 1)
	{
		double gloc[OR_NGRAD];
		memset(gloc, 0, sizeof(gloc));
#pragma omp for collapse(2) schedule(dynamic, 10)
		for (unsigned j = 0; j < H; j++) {
			for (unsigned i = 0; i < W; i++) {
				size_t idx = (size_t)i + (size_t)j * W;
				if (mask && !mask[idx]) {
					if (diag && diag->cls) diag->cls[(size_t)plane * diag->cap + idx] = 0;
					continue;
				}
				if (G && G[idx] == 0) continue;
				double y = half_ybin + j * sx;
				double x = half_xbin + i * sy;
				double xr = x - c->xc, yr = y - c->yc;
				double x_prof = xr * c->cos_ang + yr * c->sin_ang;
				double y_prof = -xr * c->sin_over + yr * c->cos_over;
				double r_prof = sqrt(x_prof * x_prof + y_prof * y_prof);
				tape_px *tpx = tp ? &tp->px[idx] : NULL;
				int cls;
				unsigned res = 0, maxrec = 0;
				if (tpx && tape_mode == TAPE_REPLAY) {
					cls = tpx->cls;
					res = tpx->res;
					maxrec = tpx->maxrec;
					tpx->pos = 0;
				}
				else {
					if (c->p.rscale_max > 0 && r_prof > c->p.rscale_max * c->rscale) cls = 1;
					else if (c->p.rough || r_prof > c->p.rscale_switch * c->rscale) cls = 2;
					else {
						cls = 3;
						res = c->p.resolution;
						maxrec = c->p.max_recursions;
						if (c->p.type == OR_SERSIC) {
							int center = fabs(x - c->xc) < c->img_sx && fabs(y - c->yc) < c->img_sy;
							if (c->p.adjust && center && c->p.shape[1] > 1) {
								res = 8;
								maxrec = 10;
							}
						}
					}
					if (tpx && tape_mode == TAPE_RECORD) {
						tpx->cls = (unsigned char)cls;
						tpx->res = (unsigned char)res;
						tpx->maxrec = (unsigned char)maxrec;
						tpx->n = 0;
						tpx->pos = 0;
					}
				}
				double pixel_val = 0;
				walk wk = {c, tpx, tpx ? tape_mode : TAPE_OFF, 0, NULL, 1.0};
				double gpix[OR_NGRAD];
				if (G) {
					memset(gpix, 0, sizeof(gpix));
					wk.g = gpix;
				}
				if (cls == 1) pixel_val = 0.;
				else if (cls == 2) {
					pixel_val = eval_at(c, x_prof, y_prof);
					if (G) accumulate_sample_grad(c, x, y, 1.0, gpix);
				}
				else {
					pixel_val = subsample_pixel(&wk, x - half_xbin, x + half_xbin, y - half_ybin,
					                            y + half_ybin, 0, maxrec, res);
				}
				if (diag) {
					if (diag->cls) diag->cls[(size_t)plane * diag->cap + idx] = (unsigned char)(cls == 1 ? 1 : cls);
					if (diag->nodes) diag->nodes[(size_t)plane * diag->cap + idx] = wk.nodes;
				}
				if (G) {
					double gq = G[idx] * flux_scale;
					for (int k = 0; k < OR_NGRAD; k++) gloc[k] += gq * (gpix[k] + c->dlnfs[k] * pixel_val);
				}
				else {
					img[idx] += flux_scale * pixel_val;
				}
			}
		}
		if (G) {
#pragma omp critical
			for (int k = 0; k < OR_NGRAD; k++) gtot[k] += gloc[k];
		}
	}
	if (G && gsum)
		for (int k = 0; k < OR_NGRAD; k++) gsum[k] += gtot[k];
}

/* ------------------------------------------------------------------ */
/* masks                                                                */
/* ------------------------------------------------------------------ */

static void mask_bbox(const unsigned char *m, unsigned w, unsigned h, unsigned *lbx, unsigned *lby,
                      unsigned *ubx, unsigned *uby)
{
	unsigned lx = w, ly = h, ux = 0, uy = 0;
	int only_zeros = 1;
	for (unsigned j = 0; j < h; j++)
		for (unsigned i = 0; i < w; i++) {
			if (!m[i + (size_t)j * w]) continue;
			if (i < lx) lx = i;
			if (j < ly) ly = j;
			if (i + 1 > ux) ux = i + 1;
			if (j + 1 > uy) uy = j + 1;
			only_zeros = 0;
		}
	if (only_zeros) { lx = ly = ux = uy = 0; }
	*lbx = lx; *lby = ly; *ubx = ux; *uby = uy;
}

void or_mask_expand_by(const unsigned char *in, unsigned width, unsigned height, unsigned pad_x,
                       unsigned pad_y, unsigned char *out)
{
	memcpy(out, in, (size_t)width * height);
	const size_t src_krn_offset = pad_x + (size_t)pad_y * width;
	for (unsigned j = 0; j < height; j++) {
		for (unsigned i = 0; i < width; i++) {
			size_t output_idx = i + (size_t)j * width;
			if (out[output_idx]) continue;
			size_t input_idx = output_idx - src_krn_offset;
			unsigned l_min = 0, l_max = 2 * pad_y + 1, k_min = 0, k_max = 2 * pad_x + 1;
			if (j < pad_y) l_min = pad_y - j;
			else if ((j + pad_y) >= height) l_max = height + pad_y - j;
			if (i < pad_x) k_min = pad_x - i;
			else if ((i + pad_x) >= width) k_max = width + pad_x - i;
			input_idx += k_min + (size_t)l_min * width;
			for (size_t l = 0; l < l_max - l_min; l++) {
				const unsigned char *first = in + input_idx;
				int any = 0;
				for (unsigned k = 0; k < k_max - k_min; k++)
					if (first[k]) { any = 1; break; }
				if (any) { out[output_idx] = 1; break; }
				input_idx += width;
			}
		}
	}
}

typedef struct analysis {
	int convolution_required, mask_needs_convolution, mask_needs_psf_padding, mask_needs_adjustment;
	unsigned pad_x, pad_y, dw, dh;
} analysis;

static void analyze_expansion(unsigned dims_x, unsigned dims_y, const unsigned char *mask,
                              unsigned mw, unsigned mh, unsigned psf_w, unsigned psf_h,
                              unsigned fs, analysis *an, int adjust_mask)
{
	int has_mask = mask && mw > 0 && mh > 0;
	an->mask_needs_convolution = 0;
	an->mask_needs_psf_padding = 0;
	int model_needs_psf_padding = an->convolution_required;
	if (has_mask && !adjust_mask) {
		model_needs_psf_padding = mw > dims_x * fs && mh > dims_y * fs;
	}
	else if (has_mask && an->convolution_required) {
		unsigned lbx, lby, ubx, uby;
		mask_bbox(mask, mw, mh, &lbx, &lby, &ubx, &uby);
		lbx *= fs; lby *= fs; ubx *= fs; uby *= fs;
		unsigned pad_low_x = lbx, pad_low_y = lby;
		unsigned pad_up_x = mw * fs - ubx, pad_up_y = mh * fs - uby;
		unsigned need_x = psf_w / 2, need_y = psf_h / 2;
		int low_ok = pad_low_x >= need_x && pad_low_y >= need_y;
		int up_ok = pad_up_x >= need_x && pad_up_y >= need_y;
		model_needs_psf_padding = !low_ok || !up_ok;
		an->mask_needs_convolution = 1;
		an->mask_needs_psf_padding = model_needs_psf_padding;
	}
	if (model_needs_psf_padding) {
		an->pad_x = psf_w / 2;
		an->pad_y = psf_h / 2;
	}
	else {
		an->pad_x = an->pad_y = 0;
	}
	an->mask_needs_adjustment = has_mask && (an->mask_needs_convolution || an->mask_needs_psf_padding || fs > 1);
	an->dw = dims_x * fs + an->pad_x * 2;
	an->dh = dims_y * fs + an->pad_y * 2;
}

/* Model::adjust(mask, psf, fs, analysis); returns malloc'ed mask of drawing dims */
static unsigned char *adjust_mask_impl(const unsigned char *mask, unsigned mw, unsigned mh,
                                       unsigned psf_w, unsigned psf_h, unsigned fs,
                                       const analysis *an, unsigned *ow, unsigned *oh)
{
	unsigned w = mw, h = mh;
	unsigned char *cur = malloc((size_t)w * h);
	memcpy(cur, mask, (size_t)w * h);
	if (fs > 1) {
		unsigned uw = w * fs, uh = h * fs;
		unsigned char *up = malloc((size_t)uw * uh);
		for (unsigned r = 0; r < uh; r++)
			for (unsigned col = 0; col < uw; col++) up[col + (size_t)r * uw] = cur[col / fs + (size_t)(r / fs) * w];
		free(cur);
		cur = up; w = uw; h = uh;
	}
	if (an->mask_needs_psf_padding) {
		unsigned ew = an->dw, eh = an->dh;
		unsigned char *ext = calloc((size_t)ew * eh, 1);
		for (unsigned j = 0; j < h; j++)
			for (unsigned i = 0; i < w; i++) ext[(i + an->pad_x) + (size_t)(j + an->pad_y) * ew] = cur[i + (size_t)j * w];
		free(cur);
		cur = ext; w = ew; h = eh;
	}
	if (an->mask_needs_convolution) {
		unsigned char *exp_ = malloc((size_t)w * h);
		or_mask_expand_by(cur, w, h, psf_w / 2, psf_h / 2, exp_);
		free(cur);
		cur = exp_;
	}
	*ow = w; *oh = h;
	return cur;
}

int or_adjust_mask(const unsigned char *mask, unsigned mw, unsigned mh, unsigned img_w,
                   unsigned img_h, const double *psf, unsigned pw, unsigned ph,
                   unsigned finesampling, unsigned char *out, unsigned out_cap, unsigned *ow,
                   unsigned *oh)
{
	(void)psf;
	analysis an;
	memset(&an, 0, sizeof(an));
	an.convolution_required = (size_t)pw * ph > 0;
	analyze_expansion(img_w, img_h, mask, mw, mh, pw, ph, finesampling, &an, 1);
	if (!an.mask_needs_adjustment) {
		if ((size_t)mw * mh > out_cap) return fail("output too small");
		memcpy(out, mask, (size_t)mw * mh);
		*ow = mw; *oh = mh;
		return 0;
	}
	unsigned w, h;
	unsigned char *adj = adjust_mask_impl(mask, mw, mh, pw, ph, finesampling, &an, &w, &h);
	if ((size_t)w * h > out_cap) { free(adj); return fail("output too small"); }
	memcpy(out, adj, (size_t)w * h);
	free(adj);
	*ow = w; *oh = h;
	return 0;
}

/* ------------------------------------------------------------------ */
/* convolvers                                                           */
/* ------------------------------------------------------------------ */

static void conv_brute_assoc(const double *src, unsigned sw, unsigned sh, const double *krn,
                             unsigned kw, unsigned kh, const unsigned char *mask, double *out,
                             int threads)
{
	const unsigned khw = kw / 2, khh = kh / 2;
	size_t kn = (size_t)kw * kh;
	double *ikrn = malloc(sizeof(double) * kn);
	for (size_t t = 0; t < kn; t++) ikrn[t] = krn[kn - 1 - t];
	const size_t src_krn_offset = khw + (size_t)khh * sw;
	int nthreads = threads < 1 ? 1 : threads;
#pragma omp parallel for collapse(2) schedule(dynamic, 10) num_threads(nthreads) if (nthreads > 1)
	for (unsigned j = 0; j < sh; j++) {
		for (unsigned i = 0; i < sw; i++) {
			size_t im_idx = i + (size_t)j * sw;
			if (mask && !mask[im_idx]) { out[im_idx] = 0; continue; }
			size_t src_offset = im_idx - src_krn_offset;
			size_t krn_offset = 0;
			unsigned l_min = 0, l_max = kh, k_min = 0, k_max = kw;
			if (j < khh) l_min = khh - j;
			else if ((j + khh) >= sh) l_max = sh + khh - j;
			if (i < khw) k_min = khw - i;
			else if ((i + khw) >= sw) k_max = sw + khw - i;
			src_offset += k_min + (size_t)l_min * sw;
			krn_offset += k_min + (size_t)l_min * kw;
			double pixel = 0;
			for (size_t l = 0; l < l_max - l_min; l++) {
				double dot = 0;
				for (unsigned k = 0; k < k_max - k_min; k++) dot += src[src_offset + k] * ikrn[krn_offset + k];
				pixel += dot;
				src_offset += sw;
				krn_offset += kw;
			}
			out[im_idx] = pixel;
		}
	}
	free(ikrn);
}

static void conv_brute_old(const double *src, unsigned sw, unsigned sh, const double *krn,
                           unsigned kw, unsigned kh, const unsigned char *mask, double *out,
                           int threads)
{
	const unsigned khw = kw / 2, khh = kh / 2;
	size_t kn = (size_t)kw * kh;
	int nthreads = threads < 1 ? 1 : threads;
#pragma omp parallel for collapse(2) schedule(dynamic, 10) num_threads(nthreads) if (nthreads > 1)
	for (unsigned j = 0; j < sh; j++) {
		for (unsigned i = 0; i < sw; i++) {
			size_t im_idx = i + (size_t)j * sw;
			if (mask && !mask[im_idx]) { out[im_idx] = 0; continue; }
			double pixel = 0;
			size_t kp = kn - 1;
			for (unsigned l = 0; l < kh; l++) {
				int src_j = (int)j + (int)l - (int)khh;
				for (unsigned k = 0; k < kw; k++, kp--) {
					int src_i = (int)i + (int)k - (int)khw;
					if (src_i >= 0 && (unsigned)src_i < sw && src_j >= 0 && (unsigned)src_j < sh)
						pixel += src[src_i + (size_t)src_j * sw] * krn[kp];
				}
			}
			out[im_idx] = pixel;
		}
	}
}

/* convolve_impl for one convolver; for FFT with crop=0 the result is the
 * extended (2*max dims) image and *off receives its offset. */
static double *convolve_impl(int convolver, const double *src, unsigned sw, unsigned sh,
                             const double *krn, unsigned kw, unsigned kh, const unsigned char *mask,
                             int crop, int threads, unsigned *rw, unsigned *rh, unsigned *offx,
                             unsigned *offy)
{
	double *out = calloc((size_t)sw * sh, sizeof(double));
	*offx = *offy = 0;
	*rw = sw; *rh = sh;
	if (convolver == OR_CONV_NULL) {
		/* NullConvolver::convolve_impl (convolve.cpp:521-524) */
		memcpy(out, src, sizeof(double) * (size_t)sw * sh);
		return out;
	}
	if (convolver == OR_CONV_BRUTE_OLD) {
		conv_brute_old(src, sw, sh, krn, kw, kh, mask, out, threads);
		return out;
	}
	conv_brute_assoc(src, sw, sh, krn, kw, kh, mask, out, threads);
	if (convolver == OR_CONV_FFT && !crop) {
		/* FFTConvolver::convolve_impl with crop=false: the whole extended
		 * image (ext = 2*max(src,krn) with max() taking x for both axes) */
		unsigned mx = sw > kw ? sw : kw;
		unsigned ew = 2 * mx, eh = 2 * mx;
		unsigned ox = sw / 2, oy = sh / 2;
		if (sw % 2 == 0 || kw % 2 == 0) ox -= 1;
		if (sh % 2 == 0 || kh % 2 == 0) oy -= 1;
		/* full linear convolution placed so the cropped window sits at (ox, oy) */
		double *ext = calloc((size_t)ew * eh, sizeof(double));
		unsigned kx0 = (sw - kw) / 2, ky0 = (sh - kh) / 2;
		for (unsigned j = 0; j < sh; j++)
			for (unsigned i = 0; i < sw; i++) {
				double s = src[i + (size_t)j * sw];
				if (s == 0) continue;
				for (unsigned b = 0; b < kh; b++)
					for (unsigned a = 0; a < kw; a++) {
						size_t X = i + a + kx0, Y = j + b + ky0;
						if (X < ew && Y < eh) ext[X + Y * ew] += s * krn[a + (size_t)b * kw];
					}
			}
		if (mask) {
			for (unsigned j = 0; j < eh; j++)
				for (unsigned i = 0; i < ew; i++) {
					int in = i >= ox && i < ox + sw && j >= oy && j < oy + sh;
					if (!in || !mask[(i - ox) + (size_t)(j - oy) * sw]) ext[i + (size_t)j * ew] = 0;
				}
		}
		free(out);
		*rw = ew; *rh = eh; *offx = ox; *offy = oy;
		return ext;
	}
	return out;
}

/* Convolver::convolve (convolve.cpp:55-79) */
static double *convolver_convolve(int convolver, const double *src, unsigned sw, unsigned sh,
                                  const double *krn, unsigned kw, unsigned kh,
                                  const unsigned char *mask, int crop, int threads, unsigned *rw,
                                  unsigned *rh, unsigned *offx, unsigned *offy, int *err)
{
	*err = 0;
	if (!(sw >= kw && sh >= kh)) {
		unsigned m = sw > kw ? sw : kw; /* libprofit's max() uses the x components for both */
		unsigned nw = m, nh = m;
		if (nw < sw || nh < sh) { *err = 1; return NULL; }
		unsigned dx = nw - sw, dy = nh - sh;
		double *es = calloc((size_t)nw * nh, sizeof(double));
		unsigned char *em = NULL;
		for (unsigned j = 0; j < sh; j++) memcpy(es + (size_t)j * nw, src + (size_t)j * sw, sizeof(double) * sw);
		if (mask) {
			em = calloc((size_t)nw * nh, 1);
			for (unsigned j = 0; j < sh; j++) memcpy(em + (size_t)j * nw, mask + (size_t)j * sw, sw);
		}
		unsigned ew, eh, ox, oy;
		double *res = convolve_impl(convolver, es, nw, nh, krn, kw, kh, em, crop, threads, &ew, &eh, &ox, &oy);
		free(es);
		free(em);
		unsigned cw = ew - dx, ch = eh - dy;
		double *cr = malloc(sizeof(double) * (size_t)cw * ch);
		for (unsigned j = 0; j < ch; j++) memcpy(cr + (size_t)j * cw, res + (size_t)j * ew, sizeof(double) * cw);
		free(res);
		*rw = cw; *rh = ch; *offx = ox; *offy = oy;
		return cr;
	}
	return convolve_impl(convolver, src, sw, sh, krn, kw, kh, mask, crop, threads, rw, rh, offx, offy);
}

int or_convolve(int convolver, const double *src, unsigned sw, unsigned sh, const double *krn,
                unsigned kw, unsigned kh, const unsigned char *mask, double *out, int threads)
{
	unsigned rw, rh, ox, oy;
	int err;
	double *r = convolver_convolve(convolver, src, sw, sh, krn, kw, kh, mask, 1, threads, &rw, &rh, &ox, &oy, &err);
	if (err || !r) return fail("cannot extend source image to kernel size");
	if (rw != sw || rh != sh) { free(r); return fail("unexpected convolution size"); }
	memcpy(out, r, sizeof(double) * (size_t)sw * sh);
	free(r);
	return 0;
}

/* ------------------------------------------------------------------ */
/* model                                                                */
/* ------------------------------------------------------------------ */

void or_model_defaults(or_model *m)
{
	memset(m, 0, sizeof(*m));
	m->scale_x = m->scale_y = 1;
	m->finesampling = 1;
	m->return_finesampled = 1;
	m->crop = 1;
	m->adjust_mask = 1;
	m->convolver = OR_CONV_BRUTE;
	m->threads = 1;
	m->psf_scale_x = 1;
	m->psf_scale_y = 1;
}

static int analyze_inputs(const or_model *m, const or_profile *profiles, int nprof, analysis *an)
{
	memset(an, 0, sizeof(*an));
	if (!(m->width > 0 && m->height > 0)) return fail("Model's requested dimensions are 0");
	if (m->scale_x <= 0) return fail("Model's scale_x cannot be negative or zero");
	if (m->scale_y <= 0) return fail("Model's scale_y cannot be negative or zero");
	if (m->finesampling == 0) return fail("finesampling must be > 0");
	int has_mask = m->mask && m->mask_w > 0 && m->mask_h > 0;
	if (has_mask && m->adjust_mask && (m->mask_w != m->width || m->mask_h != m->height))
		return fail("Mask dimensions != model dimensions");
	for (int k = 0; k < nprof; k++)
		if (profiles[k].convolve) an->convolution_required = 1;
	int has_psf = m->psf && m->psf_w > 0 && m->psf_h > 0;
	if (an->convolution_required && !has_psf) return fail("No psf provided but profile(s) requested convolution");
	for (int k = 0; k < nprof; k++) {
		if (validate(&profiles[k])) return -1;
		if (profiles[k].type == OR_POINT && !has_psf)
			return fail("No psf present in the model, cannot produce a psf profile");
	}
	analyze_expansion(m->width, m->height, has_mask ? m->mask : NULL, m->mask_w, m->mask_h,
	                  has_psf ? m->psf_w : 0, has_psf ? m->psf_h : 0, m->finesampling, an, m->adjust_mask);
	return 0;
}

int or_model_drawing_dims(const or_model *m, const or_profile *profiles, int nprof,
                          unsigned *dw, unsigned *dh, unsigned *pad_x, unsigned *pad_y)
{
	analysis an;
	if (analyze_inputs(m, profiles, nprof, &an)) return -1;
	*dw = an.dw; *dh = an.dh; *pad_x = an.pad_x; *pad_y = an.pad_y;
	return 0;
}

static double *normalized_psf(const or_model *m)
{
	size_t n = (size_t)m->psf_w * m->psf_h;
	double *p = malloc(sizeof(double) * n);
	double sum = 0;
	for (size_t i = 0; i < n; i++) { p[i] = m->psf[i]; sum += p[i]; }
	if (sum > 0)
		for (size_t i = 0; i < n; i++) p[i] /= sum;
	return p;
}

typedef struct eval_ctx {
	/* inputs of a render pass, shared by evaluate and the gradient */
	analysis an;
	const unsigned char *draw_mask; /* mask passed to profiles (drawing dims) or NULL */
	unsigned char *owned_mask;
	double psx, psy;                /* profile pixel scale */
	comp *comps;
	int ncomp;
	double *psf;
} eval_ctx;

static int prepare(const or_model *m, const or_profile *profiles, int nprof, eval_ctx *e)
{
	memset(e, 0, sizeof(*e));
	if (analyze_inputs(m, profiles, nprof, &e->an)) return -1;
	int has_mask = m->mask && m->mask_w > 0 && m->mask_h > 0;
	if (has_mask && m->adjust_mask && e->an.mask_needs_adjustment) {
		unsigned w, h;
		e->owned_mask = adjust_mask_impl(m->mask, m->mask_w, m->mask_h, m->psf_w, m->psf_h,
		                                 m->finesampling, &e->an, &w, &h);
		e->draw_mask = e->owned_mask;
	}
	else {
		if (!m->adjust_mask && (m->mask_w != e->an.dw || m->mask_h != e->an.dh)) {
			snprintf(g_err, sizeof(g_err), "Mask dimensions != drawing dimensions: [%u, %u] != [%u, %u]",
			         m->mask_w, m->mask_h, e->an.dw, e->an.dh);
			return -1;
		}
		e->draw_mask = has_mask ? m->mask : NULL;
	}
	e->psx = m->scale_x;
	e->psy = m->scale_y;
	if (m->finesampling > 1) {
		e->psx /= m->finesampling;
		e->psy /= m->finesampling;
	}
	e->ncomp = nprof;
	e->comps = calloc(nprof > 0 ? nprof : 1, sizeof(comp));
	for (int k = 0; k < nprof; k++)
		setup_comp(&e->comps[k], &profiles[k], m->magzero, e->psx, e->psy, m->scale_x, m->scale_y,
		           e->an.pad_x, e->an.pad_y);
	if (m->psf && m->psf_w > 0 && m->psf_h > 0) e->psf = normalized_psf(m);
	for (int k = 0; k < nprof; k++) {
		comp *c = &e->comps[k];
		if (c->p.type == OR_POINT) {
			c->psf = e->psf;
			c->kw = m->psf_w;
			c->kh = m->psf_h;
			c->psf_sx = m->psf_scale_x;
			c->psf_sy = m->psf_scale_y;
		}
		else if (c->p.type == OR_SKY) {
			/* sky.cpp:42-45: bg = requested_bg / (finesampling * finesampling) */
			c->flux_scale = c->p.shape[0] / (double)(m->finesampling * m->finesampling);
			c->inv_fs2 = 1.0 / (double)(m->finesampling * m->finesampling);
		}
	}
	return 0;
}

static void release(eval_ctx *e)
{
	free(e->owned_mask);
	free(e->comps);
	free(e->psf);
}

static tape *tape_new(unsigned npix)
{
	tape *t = calloc(1, sizeof(tape));
	t->npix = npix;
	t->px = calloc(npix, sizeof(tape_px));
	return t;
}

static void tape_free(tape *t)
{
	if (!t) return;
	for (unsigned i = 0; i < t->npix; i++) free(t->px[i].bits);
	free(t->px);
	free(t);
}

/*
 * Model::evaluate + produce_image (model.cpp:271-380).  tapes (one per
 * profile) may be NULL.  Returns malloc'ed image.
 */
static double *model_eval(const or_model *m, eval_ctx *e, or_diag *diag, tape **tapes,
                          int tape_mode, unsigned *ow, unsigned *oh, int *offx, int *offy)
{
	const analysis *an = &e->an;
	unsigned dw = an->dw, dh = an->dh;
	size_t dn = (size_t)dw * dh;
	double *image = calloc(dn, sizeof(double));
	*offx = *offy = 0;
	if (m->dry_run) { *ow = dw; *oh = dh; return image; }

	double *to_convolve = an->convolution_required ? calloc(dn, sizeof(double)) : NULL;
	for (int k = 0; k < e->ncomp; k++) {
		double *target = e->comps[k].p.convolve ? to_convolve : image;
		render_comp(&e->comps[k], target, e->draw_mask, dw, dh, e->psx, e->psy, m->threads, diag, k,
		            tapes ? tapes[k] : NULL, tape_mode, NULL, NULL);
	}
	unsigned iw = dw, ih = dh;
	unsigned coffx = 0, coffy = 0;
	if (an->convolution_required) {
		unsigned rw, rh;
		int err;
		double *conv = convolver_convolve(m->convolver, to_convolve, dw, dh, e->psf, m->psf_w, m->psf_h,
		                                  e->draw_mask, m->crop, m->threads, &rw, &rh, &coffx, &coffy, &err);
		free(to_convolve);
		if (err || !conv) { free(image); fail("cannot extend source image to kernel size"); return NULL; }
		if (rw != dw || rh != dh) {
			double *ext = calloc((size_t)rw * rh, sizeof(double));
			for (unsigned j = 0; j < dh; j++)
				for (unsigned i = 0; i < dw; i++) ext[(i + coffx) + (size_t)(j + coffy) * rw] = image[i + (size_t)j * dw];
			free(image);
			image = ext;
			iw = rw; ih = rh;
		}
		for (size_t t = 0; t < (size_t)iw * ih; t++) image[t] += conv[t];
		free(conv);
	}
	int offset_x = (int)coffx, offset_y = (int)coffy;

	if (an->pad_x > 0 && an->pad_y > 0) {
		unsigned crop_ox = an->pad_x, crop_oy = an->pad_y;
		unsigned cw = dw - 2 * an->pad_x, ch = dh - 2 * an->pad_y;
		if (!m->crop) {
			int diff_x = 0, diff_y = 0, ddim_x = 0, ddim_y = 0;
			if (m->convolver == OR_CONV_FFT) {
				/* FFTConvolver::padding for actual and intended source dims */
				unsigned sx_[2] = {dw, dw - 2 * an->pad_x}, sy_[2] = {dh, dh - 2 * an->pad_y};
				int first_x[2], first_y[2], second_x[2], second_y[2];
				for (int t = 0; t < 2; t++) {
					unsigned mx = sx_[t] > m->psf_w ? sx_[t] : m->psf_w;
					unsigned ex = 2 * mx, ey = 2 * mx;
					int ox = (int)(sx_[t] / 2), oy = (int)(sy_[t] / 2);
					if (sx_[t] % 2 == 0 || m->psf_w % 2 == 0) ox -= 1;
					if (sy_[t] % 2 == 0 || m->psf_h % 2 == 0) oy -= 1;
					first_x[t] = ox; first_y[t] = oy;
					second_x[t] = (int)(ex - ex / 2) - ox;
					second_y[t] = (int)(ey - ey / 2) - oy;
				}
				diff_x = first_x[0] - first_x[1]; diff_y = first_y[0] - first_y[1];
				ddim_x = second_x[0] - second_x[1]; ddim_y = second_y[0] - second_y[1];
			}
			cw = iw - 2 * an->pad_x - (diff_x + ddim_x);
			ch = ih - 2 * an->pad_y - (diff_y + ddim_y);
			crop_ox += diff_x; crop_oy += diff_y;
			offset_x -= diff_x; offset_y -= diff_y;
		}
		double *cr = malloc(sizeof(double) * (size_t)cw * ch);
		for (unsigned j = 0; j < ch; j++)
			for (unsigned i = 0; i < cw; i++) cr[i + (size_t)j * cw] = image[(i + crop_ox) + (size_t)(j + crop_oy) * iw];
		free(image);
		image = cr;
		iw = cw; ih = ch;
	}

	unsigned fs = m->finesampling;
	if (fs > 1 && !m->return_finesampled) {
		unsigned ddx = (iw + fs - 1) / fs, ddy = (ih + fs - 1) / fs;
		double *down = calloc((size_t)ddx * ddy, sizeof(double));
		for (unsigned row = 0; row < ih; row++)
			for (unsigned col = 0; col < iw; col++) down[col / fs + (size_t)(row / fs) * ddx] += image[col + (size_t)row * iw];
		free(image);
		image = down;
		iw = ddx; ih = ddy;
		offset_x /= (int)fs;
		offset_y /= (int)fs;
	}
	if (e->owned_mask) {
		size_t n = (size_t)m->mask_w * m->mask_h;
		size_t tot = (size_t)iw * ih;
		for (size_t t = 0; t < tot && t < n; t++)
			if (!m->mask[t]) image[t] = 0;
	}
	*ow = iw; *oh = ih;
	*offx = offset_x; *offy = offset_y;
	return image;
}

int or_model_evaluate(const or_model *m, const or_profile *profiles, int nprof, double *out,
                      unsigned out_cap, unsigned *out_w, unsigned *out_h, int *off_x, int *off_y,
                      or_diag *diag)
{
	eval_ctx e;
	if (prepare(m, profiles, nprof, &e)) { release(&e); return -1; }
	if (diag) {
		size_t need = (size_t)e.an.dw * e.an.dh;
		if (diag->cap < need) { release(&e); return fail("diag buffers too small"); }
	}
	unsigned w, h;
	int ox, oy;
	double *img = model_eval(m, &e, diag, NULL, TAPE_OFF, &w, &h, &ox, &oy);
	release(&e);
	if (!img) return -1;
	if ((size_t)w * h > out_cap) { free(img); return fail("output buffer too small"); }
	memcpy(out, img, sizeof(double) * (size_t)w * h);
	free(img);
	*out_w = w; *out_h = h;
	if (off_x) *off_x = ox;
	if (off_y) *off_y = oy;
	return 0;
}

int or_profile_info(const or_profile *p, double magzero, double sx, double sy, or_comp_info *info)
{
	if (validate(p)) return -1;
	comp c;
	setup_comp(&c, p, magzero, sx, sy, sx, sy, 0, 0);
	info->rscale = c.rscale;
	info->ie = c.ie;
	info->bn = c.bn;
	info->lumtot = c.lumtot;
	info->rbox = c.rbox;
	info->flux_scale = c.flux_scale;
	info->acc = c.p.acc;
	info->rscale_switch = c.p.rscale_switch;
	info->rscale_max = c.p.rscale_max;
	info->resolution = c.p.resolution;
	info->max_recursions = c.p.max_recursions;
	info->cos_ang = c.cos_ang;
	info->sin_ang = c.sin_ang;
	return 0;
}

double or_profile_eval_at(const or_profile *p, double magzero, double x, double y)
{
	comp c;
	setup_comp(&c, p, magzero, 1, 1, 1, 1, 0, 0);
	return eval_at(&c, x, y);
}

/* ------------------------------------------------------------------ */
/* likelihood                                                           */
/* ------------------------------------------------------------------ */

int or_loglike_image(int likefunc, const double *data, const double *sigma,
                     const unsigned char *region, const double *model, unsigned n, double *ll,
                     double *dll)
{
	const double ln_sqrt_2pi = 0.918938533204672741780329736406;
	size_t N = 0;
	for (unsigned t = 0; t < n; t++)
		if (!region || region[t]) N++;
	if (dll) memset(dll, 0, sizeof(double) * n);
	if (N == 0) return fail("empty likelihood region");
	if (likefunc == OR_LIKE_CHISQ || likefunc == OR_LIKE_NORM || likefunc == OR_LIKE_T) {
		double chisq = 0, lsum = 0;
		for (unsigned t = 0; t < n; t++) {
			if (region && !region[t]) continue;
			double z = (data[t] - model[t]) / sigma[t];
			chisq += z * z;
			lsum += -(ln_sqrt_2pi + 0.5 * z * z);
		}
		if (likefunc == OR_LIKE_NORM) {
			*ll = lsum;
			if (dll)
				for (unsigned t = 0; t < n; t++)
					if (!region || region[t]) dll[t] = (data[t] - model[t]) / (sigma[t] * sigma[t]);
			return 0;
		}
		if (likefunc == OR_LIKE_CHISQ) {
			/* dchisq(x, N, log) = dgamma(x, N/2, scale 2, log) */
			double k = 0.5 * (double)N;
			double X = chisq;
			if (k >= 1) *ll = (k - 1) * log(X / 2) - X / 2 - lgamma(k) - log(2.0);
			else *ll = k * log(X / 2) - X / 2 - lgamma(k) - log(X);
			double dLdX = (k >= 1) ? (k - 1) / X - 0.5 : (k - 1) / X - 0.5;
			if (dll)
				for (unsigned t = 0; t < n; t++)
					if (!region || region[t]) dll[t] = dLdX * (-2) * (data[t] - model[t]) / (sigma[t] * sigma[t]);
			return 0;
		}
		/* t: dof from the sample variance of the standardised residuals */
		double mean = 0;
		for (unsigned t = 0; t < n; t++)
			if (!region || region[t]) mean += (data[t] - model[t]) / sigma[t];
		mean /= (double)N;
		double var = 0;
		for (unsigned t = 0; t < n; t++)
			if (!region || region[t]) {
				double z = (data[t] - model[t]) / sigma[t] - mean;
				var += z * z;
			}
		var /= (double)(N - 1);
		double dof = 2 * var / (var - 1);
		/* max(1, min(Inf, dof, na.rm=TRUE)): NaN (N < 2) gives Inf, below 1 gives 1 */
		int clamped = 0;
		if (N < 2 || isnan(dof)) dof = INFINITY;
		else if (!(dof > 1)) {
			dof = 1;
			clamped = 1;
		}
		double s = 0, sl = 0, sc = 0;
		for (unsigned t = 0; t < n; t++) {
			if (region && !region[t]) continue;
			double z = (data[t] - model[t]) / sigma[t];
			if (isinf(dof)) s += -(ln_sqrt_2pi + 0.5 * z * z);
			else {
				s += lgamma((dof + 1) / 2) - lgamma(dof / 2) - 0.5 * log(dof * M_PI) - (dof + 1) / 2 * log1p(z * z / dof);
				sl += log1p(z * z / dof);
				sc += z * z / (dof * (dof + z * z));
			}
		}
		*ll = s;
		if (dll) {
			/* d LL / d z_t at fixed dof, plus the dof's dependence on z through
			 * the sample variance: dof = 2 v / (v - 1), dv/dz_t = 2 (z_t - mean) / (N - 1) */
			double kdof = 0;
			if (!isinf(dof) && !clamped) {
				double dLdnu = (double)N * (0.5 * or_digamma((dof + 1) / 2) - 0.5 * or_digamma(dof / 2) - 0.5 / dof) -
				               0.5 * sl + 0.5 * (dof + 1) * sc;
				double dnudv = -2 / ((var - 1) * (var - 1));
				kdof = dLdnu * dnudv * 2 / (double)(N - 1);
			}
			for (unsigned t = 0; t < n; t++) {
				if (region && !region[t]) continue;
				double z = (data[t] - model[t]) / sigma[t];
				double dz = isinf(dof) ? -z : -(dof + 1) * z / (dof + z * z) + kdof * (z - mean);
				dll[t] = -dz / sigma[t];
			}
		}
		return 0;
	}
	if (likefunc == OR_LIKE_POIS) {
		double scale = 0;
		for (unsigned t = 0; t < n; t++)
			if (!region || region[t]) {
				double v = fabs(data[t]) / (fabs(sigma[t]) * fabs(sigma[t]));
				if (v > scale) scale = v;
			}
		double s = (scale < 0.1 || scale > 10) ? scale : 1;
		double L = 0;
		for (unsigned t = 0; t < n; t++) {
			if (region && !region[t]) continue;
			double mo = model[t] * s, im = data[t] * s;
			L += mo - im * log(mo);
			if (dll) dll[t] = -s * (1 - data[t] / model[t]);
		}
		*ll = -L;
		return 0;
	}
	return fail("unknown likelihood function");
}

int or_loglike(const or_fit *fit, const or_profile *profiles, int nprof, double *ll, double *model_out)
{
	unsigned n = fit->model.width * fit->model.height;
	double *img = malloc(sizeof(double) * n * 4);
	unsigned w, h;
	int ox, oy;
	if (or_model_evaluate(&fit->model, profiles, nprof, img, n * 4, &w, &h, &ox, &oy, NULL)) { free(img); return -1; }
	if (w * h != n) { free(img); return fail("model image size differs from data size"); }
	int rc = or_loglike_image(fit->likefunc, fit->data, fit->sigma, fit->region, img, n, ll, NULL);
	if (model_out) memcpy(model_out, img, sizeof(double) * n);
	free(img);
	return rc;
}

/* ------------------------------------------------------------------ */
/* gradients                                                            */
/* ------------------------------------------------------------------ */

static double *slot_ptr(or_profile *p, int slot)
{
	switch (slot) {
	case OR_G_XCEN: return &p->xcen;
	case OR_G_YCEN: return &p->ycen;
	case OR_G_MAG: return &p->mag;
	case OR_G_ANG: return &p->ang;
	case OR_G_AXRAT: return &p->axrat;
	case OR_G_BOX: return &p->box;
	default: return &p->shape[slot - OR_G_S0];
	}
}

static int check_grad_support(const or_fit *fit, const or_profile *profiles, int nprof)
{
	if (fit->model.finesampling != 1) return fail("gradient supports finesampling = 1 only");
	for (int k = 0; k < nprof; k++) {
		int t = profiles[k].type;
		if (!(t == OR_SERSIC || t == OR_MOFFAT || t == OR_FERRER || t == OR_KING || t == OR_POINT || t == OR_SKY))
			return fail("gradient supports sersic, moffat, ferrer, king, psf and sky profiles");
		if (t == OR_SERSIC && profiles[k].rescale_flux) return fail("gradient does not support rescale_flux");
	}
	return 0;
}

int or_loglike_grad(const or_fit *fit, const or_profile *profiles, int nprof, double *ll, double *grad)
{
	const or_model *m = &fit->model;
	if (check_grad_support(fit, profiles, nprof)) return -1;
	eval_ctx e;
	if (prepare(m, profiles, nprof, &e)) { release(&e); return -1; }
	unsigned w, h;
	int ox, oy;
	double *img = model_eval(m, &e, NULL, NULL, TAPE_OFF, &w, &h, &ox, &oy);
	if (!img) { release(&e); return -1; }
	unsigned n = m->width * m->height;
	if (w * h != n) { free(img); release(&e); return fail("model image size differs from data size"); }
	double *dll = calloc(n, sizeof(double));
	if (or_loglike_image(fit->likefunc, fit->data, fit->sigma, fit->region, img, n, ll, dll)) {
		free(img); free(dll); release(&e); return -1;
	}
	/* back through the final mask and the crop: w on the drawing canvas */
	unsigned dw = e.an.dw, dh = e.an.dh;
	double *wd = calloc((size_t)dw * dh, sizeof(double));
	for (unsigned j = 0; j < m->height; j++)
		for (unsigned i = 0; i < m->width; i++) {
			size_t t = i + (size_t)j * m->width;
			if (e.owned_mask && !m->mask[t]) continue;
			wd[(i + e.an.pad_x) + (size_t)(j + e.an.pad_y) * dw] = dll[t];
		}
	/* back through the masked convolution: G(q) = sum_o mask(o) w(o) krn(...) */
	double *Gc = NULL;
	if (e.an.convolution_required) {
		unsigned kw = m->psf_w, kh = m->psf_h;
		const unsigned khw = kw / 2, khh = kh / 2;
		Gc = calloc((size_t)dw * dh, sizeof(double));
		for (unsigned oj = 0; oj < dh; oj++)
			for (unsigned oi = 0; oi < dw; oi++) {
				size_t o = oi + (size_t)oj * dw;
				double wv = wd[o];
				if (wv == 0) continue;
				if (e.draw_mask && !e.draw_mask[o]) continue;
				for (unsigned l = 0; l < kh; l++) {
					int qj = (int)oj + (int)l - (int)khh;
					if (qj < 0 || qj >= (int)dh) continue;
					for (unsigned k = 0; k < kw; k++) {
						int qi = (int)oi + (int)k - (int)khw;
						if (qi < 0 || qi >= (int)dw) continue;
						Gc[qi + (size_t)qj * dw] += wv * e.psf[(kw - 1 - k) + (size_t)(kh - 1 - l) * kw];
					}
				}
			}
	}
	memset(grad, 0, sizeof(double) * OR_NGRAD * nprof);
	for (int k = 0; k < nprof; k++) {
		const double *G = e.comps[k].p.convolve ? Gc : wd;
		render_comp(&e.comps[k], NULL, e.draw_mask, dw, dh, e.psx, e.psy, m->threads, NULL, k, NULL,
		            TAPE_OFF, G, grad + (size_t)k * OR_NGRAD);
	}
	free(Gc);
	free(wd);
	free(dll);
	free(img);
	release(&e);
	return 0;
}

/* evaluate the likelihood with a given tape mode */
static int loglike_taped(const or_fit *fit, const or_profile *profiles, const eval_ctx *base,
                         int nprof, tape **tapes, int mode, double *ll, double *model_out)
{
	const or_model *m = &fit->model;
	eval_ctx e;
	if (prepare(m, profiles, nprof, &e)) { release(&e); return -1; }
	if (base) {
		/* hold the adjusted sub-pixel settings of the base parameters */
		for (int k = 0; k < nprof; k++) {
			e.comps[k].p.acc = base->comps[k].p.acc;
			e.comps[k].p.rscale_switch = base->comps[k].p.rscale_switch;
			e.comps[k].p.rscale_max = base->comps[k].p.rscale_max;
			e.comps[k].p.resolution = base->comps[k].p.resolution;
			e.comps[k].p.max_recursions = base->comps[k].p.max_recursions;
		}
	}
	unsigned w, h;
	int ox, oy;
	double *img = model_eval(m, &e, NULL, tapes, mode, &w, &h, &ox, &oy);
	release(&e);
	if (!img) return -1;
	unsigned n = m->width * m->height;
	if (w * h != n) { free(img); return fail("model image size differs from data size"); }
	int rc = 0;
	if (ll) rc = or_loglike_image(fit->likefunc, fit->data, fit->sigma, fit->region, img, n, ll, NULL);
	if (model_out) memcpy(model_out, img, sizeof(double) * n);
	free(img);
	return rc;
}

int or_loglike_grad_fd(const or_fit *fit, const or_profile *profiles, int nprof,
                       const double *rel_step, const unsigned char *slot_mask, int frozen,
                       double *ll, double *grad)
{
	const or_model *m = &fit->model;
	eval_ctx base;
	if (prepare(m, profiles, nprof, &base)) { release(&base); return -1; }
	unsigned npix = base.an.dw * base.an.dh;
	tape **tapes = NULL;
	if (frozen) {
		tapes = calloc(nprof, sizeof(tape *));
		for (int k = 0; k < nprof; k++) tapes[k] = tape_new(npix);
	}
	int rc = loglike_taped(fit, profiles, NULL, nprof, tapes, frozen ? TAPE_RECORD : TAPE_OFF, ll, NULL);
	or_profile *pp = malloc(sizeof(or_profile) * (nprof > 0 ? nprof : 1));
	memset(grad, 0, sizeof(double) * OR_NGRAD * nprof);
	for (int k = 0; rc == 0 && k < nprof; k++) {
		for (int s = 0; s < OR_NGRAD; s++) {
			if (slot_mask && !slot_mask[k * OR_NGRAD + s]) continue;
			memcpy(pp, profiles, sizeof(or_profile) * nprof);
			double *v = slot_ptr(&pp[k], s);
			double x0 = *v;
			double h = rel_step[k * OR_NGRAD + s] * (x0 != 0 ? fabs(x0) : 1.0);
			double lp, lm;
			*v = x0 + h;
			rc = loglike_taped(fit, pp, frozen ? &base : NULL, nprof, tapes, frozen ? TAPE_REPLAY : TAPE_OFF, &lp, NULL);
			if (rc) break;
			*v = x0 - h;
			rc = loglike_taped(fit, pp, frozen ? &base : NULL, nprof, tapes, frozen ? TAPE_REPLAY : TAPE_OFF, &lm, NULL);
			if (rc) break;
			grad[k * OR_NGRAD + s] = (lp - lm) / (2 * h);
		}
	}
	free(pp);
	if (tapes) {
		for (int k = 0; k < nprof; k++) tape_free(tapes[k]);
		free(tapes);
	}
	release(&base);
	return rc;
}

int or_model_image_frozen(const or_fit *fit, const or_profile *base_p, const or_profile *pert,
                          int nprof, double *out)
{
	const or_model *m = &fit->model;
	eval_ctx base;
	if (prepare(m, base_p, nprof, &base)) { release(&base); return -1; }
	unsigned npix = base.an.dw * base.an.dh;
	tape **tapes = calloc(nprof, sizeof(tape *));
	for (int k = 0; k < nprof; k++) tapes[k] = tape_new(npix);
	int rc = loglike_taped(fit, base_p, NULL, nprof, tapes, TAPE_RECORD, NULL, NULL);
	if (!rc) rc = loglike_taped(fit, pert, &base, nprof, tapes, TAPE_REPLAY, NULL, out);
	for (int k = 0; k < nprof; k++) tape_free(tapes[k]);
	free(tapes);
	release(&base);
	return rc;
}
