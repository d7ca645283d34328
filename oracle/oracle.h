/*
 * THE ORACLE -- test infrastructure only.
 *
 * Plain-C FP64 restatement of libprofit 1.9.3's CPU model-image path
 * (rendering with adaptive sub-pixel recursion, PSF convolution, Model
 * orchestration) and of ProFit's R-side likelihoods, plus a fixed-tree
 * analytic gradient and fixed-tree central finite differences.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this code.  It is the checker, never the thing measured or shipped.
 */
#ifndef PROFIT_ORACLE_H
#define PROFIT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- special functions (oracle_special.c) ---- */
typedef double (*or_integrand)(double x, void *ctx);
typedef struct or_quad_info {
	double abserr;
	int ier, last, neval;
} or_quad_info;

double or_gammafn(double x);
double or_lgammafn(double x);
double or_beta(double a, double b);
double or_digamma(double x);
double or_pgamma(double x, double shape);
double or_qgamma(double p, double shape);
double or_dpgamma_da(double x, double shape);
double or_integrate_qags(or_integrand f, void *ctx, double a, double b, or_quad_info *info);
double or_integrate_qagi(or_integrand f, void *ctx, double a, or_quad_info *info);
double or_integrate_qags_tol(or_integrand f, void *ctx, double a, double b, double epsabs,
                             double epsrel, int limit, or_quad_info *info);
double or_integrate_qagi_tol(or_integrand f, void *ctx, double a, double epsabs, double epsrel,
                             int limit, or_quad_info *info);
const double *or_quadpack_nodes(int which);

/* ---- profiles ---- */
enum {
	OR_SERSIC = 0,
	OR_MOFFAT = 1,
	OR_FERRER = 2,
	OR_KING = 3,
	OR_CORESERSIC = 4,
	OR_BROKENEXP = 5,
	OR_POINT = 6,   /* "psf": PSF point source (psf.cpp), parameters xcen, ycen, mag */
	OR_SKY = 7      /* "sky": constant background (sky.cpp), parameter bg in shape[0] */
};

/* Gradient slots per component. */
enum {
	OR_G_XCEN = 0, OR_G_YCEN, OR_G_MAG, OR_G_ANG, OR_G_AXRAT, OR_G_BOX,
	OR_G_S0, OR_G_S1, OR_G_S2, OR_G_S3, OR_G_S4,
	OR_NGRAD
};

/*
 * The parameters libprofit's RadialProfile and its subclasses register
 * (src/libprofit/src/profit/radial.cpp:724-752 and each profile's
 * constructor).  shape[] holds the profile-specific ones in this order:
 *   sersic     re, nser                 (rescale_flux is a flag below)
 *   moffat     fwhm, con
 *   ferrer     rout, a, b
 *   king       rc, rt, a
 *   coresersic re, rb, nser, a, b
 *   brokenexp  h1, h2, rb, a
 *   sky        bg
 */
typedef struct or_profile {
	int type;
	int convolve, rough, adjust, rescale_flux;
	double xcen, ycen, mag, ang, axrat, box;
	double acc, rscale_switch, rscale_max;
	unsigned resolution, max_recursions;
	double shape[5];
} or_profile;

void or_profile_defaults(int type, or_profile *p);

/* OR_CONV_NULL: NullConvolver (convolve.cpp:521-524), the source unchanged */
enum { OR_CONV_BRUTE = 0, OR_CONV_BRUTE_OLD = 1, OR_CONV_FFT = 2, OR_CONV_NULL = 3 };

typedef struct or_model {
	unsigned width, height;
	double scale_x, scale_y;
	double magzero;
	unsigned finesampling;
	int return_finesampled, crop, adjust_mask, dry_run;
	const double *psf;            /* may be NULL; normalised internally as set_psf does */
	unsigned psf_w, psf_h;
	const unsigned char *mask;    /* NULL = no mask */
	unsigned mask_w, mask_h;
	int convolver;
	int threads;
	double psf_scale_x, psf_scale_y; /* Model::set_psf_pixel_scale (default 1, 1) */
} or_model;

void or_model_defaults(or_model *m);

/* Per-pixel diagnostics of rendering, on the drawing canvas, one plane per
 * profile (profile-major): class (0 skipped by mask, 1 beyond rscale_max,
 * 2 direct, 3 subsampled) and the number of subsample_pixel nodes visited. */
typedef struct or_diag {
	unsigned char *cls;
	int *nodes;
	unsigned cap;                 /* pixels per plane available */
} or_diag;

const char *or_last_error(void);

/* Model::evaluate.  out receives out_w*out_h doubles (capacity out_cap). */
int or_model_evaluate(const or_model *m, const or_profile *profiles, int nprof,
                      double *out, unsigned out_cap, unsigned *out_w, unsigned *out_h,
                      int *off_x, int *off_y, or_diag *diag);

/* Drawing dimensions and PSF padding that Model::evaluate would use. */
int or_model_drawing_dims(const or_model *m, const or_profile *profiles, int nprof,
                          unsigned *dw, unsigned *dh, unsigned *pad_x, unsigned *pad_y);

/* Model::adjust(mask, dims, psf, finesampling) (static form). */
int or_adjust_mask(const unsigned char *mask, unsigned mw, unsigned mh, unsigned img_w,
                   unsigned img_h, const double *psf, unsigned pw, unsigned ph,
                   unsigned finesampling, unsigned char *out, unsigned out_cap,
                   unsigned *ow, unsigned *oh);

/* Convolver::convolve(src, krn, mask, crop=true) for the three convolvers. */
int or_convolve(int convolver, const double *src, unsigned sw, unsigned sh, const double *krn,
                unsigned kw, unsigned kh, const unsigned char *mask, double *out, int threads);

/* Mask::expand_by (box dilation by pad). */
void or_mask_expand_by(const unsigned char *in, unsigned w, unsigned h, unsigned pad_x,
                       unsigned pad_y, unsigned char *out);

/* Per-component derived scalars after initial_calculations (for tests). */
typedef struct or_comp_info {
	double rscale, ie, bn, lumtot, rbox, flux_scale;
	double acc, rscale_switch, rscale_max;
	unsigned resolution, max_recursions;
	double cos_ang, sin_ang;
} or_comp_info;
int or_profile_info(const or_profile *p, double magzero, double scale_x, double scale_y,
                    or_comp_info *info);

/* Single-point profile value evaluate_at(x_prof, y_prof) after setup. */
double or_profile_eval_at(const or_profile *p, double magzero, double x, double y);

/* ---- likelihood (R/profitLikeModel.R:244-329) ---- */
enum { OR_LIKE_NORM = 0, OR_LIKE_CHISQ = 1, OR_LIKE_T = 2, OR_LIKE_POIS = 3 };

typedef struct or_fit {
	or_model model;               /* with mask = calcregion, like profitSetupData */
	const double *data;           /* width*height */
	const double *sigma;          /* width*height */
	const unsigned char *region;  /* width*height, NULL = all pixels */
	int likefunc;
} or_fit;

int or_loglike_image(int likefunc, const double *data, const double *sigma,
                     const unsigned char *region, const double *model, unsigned n,
                     double *ll, double *dll_dmodel);

int or_loglike(const or_fit *fit, const or_profile *profiles, int nprof, double *ll,
               double *model_out);

/* Fixed-tree analytic gradient of logL w.r.t. OR_NGRAD slots per profile. */
int or_loglike_grad(const or_fit *fit, const or_profile *profiles, int nprof, double *ll,
                    double *grad);

/* Central finite differences over the selected slots (mask over OR_NGRAD per
 * profile); frozen != 0 holds the pixel classes, adjusted sub-pixel settings
 * and recursion decisions of the unperturbed parameters fixed. */
int or_loglike_grad_fd(const or_fit *fit, const or_profile *profiles, int nprof,
                       const double *rel_step, const unsigned char *slot_mask, int frozen,
                       double *ll, double *grad);

/* Jacobian-vector check helper: fixed-tree model image at perturbed params
 * with the tree recorded at base params. */
int or_model_image_frozen(const or_fit *fit, const or_profile *base, const or_profile *pert,
                          int nprof, double *out);

#ifdef __cplusplus
}
#endif

#endif
