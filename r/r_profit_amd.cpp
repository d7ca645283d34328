/*
 * R binding of ProFit's compiled layer over the profit_amd C-ABI: a drop-in
 * for src/r_profit.cpp.  It keeps the twelve .Call routines ProFit's R code
 * calls (src/r_profit.cpp:774-850) with the same names, arities, argument
 * reading and results, and adds the likelihood engine's routines (ProFit
 * computes its likelihood in R, R/profitLikeModel.R:244-342, and has no
 * gradient).
 *
 * Built by R CMD INSTALL in place of src/r_profit.cpp, linked against
 * libprofit_amd.so (Makevars: PKG_CPPFLAGS = -I<profit_amd>/include,
 * PKG_LIBS = -L<profit_amd>/profit_amd -lprofit_amd).  No R is installed where
 * this file was written; its C-ABI calls are checked by a syntax-only compile
 * against include/profit_amd.h.
 *
 * Errors: every C-ABI call returns a status and profit_amd_last_error() holds
 * the message.  C++ objects are released before Rf_error is reached (it
 * longjmps), so the error path leaks nothing.
 */
#include <cctype>
#include <cstdio>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#define R_NO_REMAP
#include <R.h>
#include <Rinternals.h>
#include <R_ext/Rdynload.h>

#include "profit_amd.h"

namespace {

// runs f; an exception becomes an R error raised after f's objects are gone
template <class F>
SEXP guarded(const char *what, F &&f)
{
	char msg[2048] = {0};
	SEXP out = R_NilValue;
	try {
		out = f();
	} catch (const std::exception &e) {
		std::snprintf(msg, sizeof msg, "%s: %s", what, e.what());
	}
	if (msg[0]) Rf_error("%s\n", msg);
	return out;
}

void check(int rc)
{
	if (rc != 0) throw std::runtime_error(profit_amd_last_error());
}

// ------------------------------------------------------------------ reading R objects

// src/r_profit.cpp:19-58: a numeric, integer or logical matrix as doubles
// (nrow = width); a REALSXP is used in place
struct ImageIn {
	const double *data = nullptr;
	std::vector<double> copy;
	uint32_t w = 0, h = 0;
};

ImageIn read_image(SEXP r)
{
	ImageIn im;
	if (r == R_NilValue) return im;
	im.w = (uint32_t)Rf_nrows(r);
	im.h = (uint32_t)Rf_ncols(r);
	const size_t n = (size_t)im.w * im.h;
	switch (TYPEOF(r)) {
	case REALSXP: im.data = REAL(r); break;
	case INTSXP: im.copy.assign(INTEGER(r), INTEGER(r) + n); break;
	case LGLSXP: im.copy.assign(LOGICAL(r), LOGICAL(r) + n); break;
	default: throw std::invalid_argument("Image not in one of the supported formats (logical, integer, double)");
	}
	if (!im.data) im.data = im.copy.data();
	return im;
}

// src/r_profit.cpp:60-67: a logical matrix as bytes
std::vector<uint8_t> read_mask(SEXP r, uint32_t *w, uint32_t *h)
{
	*w = (uint32_t)Rf_nrows(r);
	*h = (uint32_t)Rf_ncols(r);
	const int *v = LOGICAL(r);
	std::vector<uint8_t> m((size_t)*w * *h);
	for (size_t k = 0; k < m.size(); ++k) m[k] = v[k] ? 1 : 0;
	return m;
}

SEXP list_element(SEXP list, const char *name)
{
	SEXP names = Rf_getAttrib(list, R_NamesSymbol);
	if (names == R_NilValue) return R_NilValue;
	for (int i = 0; i < Rf_length(list); ++i)
		if (!std::strcmp(name, CHAR(STRING_ELT(names, i)))) return VECTOR_ELT(list, i);
	return R_NilValue;
}

// src/r_profit.cpp:80-127: element idx of a named parameter vector, when present
bool read_real(SEXP list, const char *name, unsigned idx, double *out)
{
	SEXP e = list_element(list, name);
	if (e == R_NilValue) return false;
	*out = REAL(e)[idx];
	return true;
}

bool read_int_like(SEXP list, const char *name, unsigned idx, long *out, const char *kind)
{
	SEXP e = list_element(list, name);
	if (e == R_NilValue) return false;
	switch (TYPEOF(e)) {
	case LGLSXP: *out = LOGICAL(e)[idx]; break;
	case INTSXP: *out = INTEGER(e)[idx]; break;
	case REALSXP: *out = (long)REAL(e)[idx]; break;
	default: throw std::invalid_argument(std::string("Parameter ") + name + " should be of " + kind + " type");
	}
	return true;
}

void read_bool(SEXP list, const char *name, unsigned idx, int32_t *out)
{
	long v;
	if (read_int_like(list, name, idx, &v, "boolean or integer")) *out = v != 0;
}

void read_uint(SEXP list, const char *name, unsigned idx, uint32_t *out)
{
	long v;
	if (read_int_like(list, name, idx, &v, "numeric")) *out = (uint32_t)v;
}

void read_radial(SEXP l, unsigned i, profit_amd_profile &p)
{
	// src/r_profit.cpp:128-145
	read_real(l, "xcen", i, &p.xcen);
	read_real(l, "ycen", i, &p.ycen);
	read_real(l, "mag", i, &p.mag);
	read_real(l, "ang", i, &p.ang);
	read_real(l, "axrat", i, &p.axrat);
	read_real(l, "box", i, &p.box);
	read_bool(l, "rough", i, &p.rough);
	read_bool(l, "adjust", i, &p.adjust);
	read_real(l, "acc", i, &p.acc);
	read_real(l, "rscale_switch", i, &p.rscale_switch);
	read_uint(l, "resolution", i, &p.resolution);
	read_uint(l, "max_recursions", i, &p.max_recursions);
	read_real(l, "rscale_max", i, &p.rscale_max);
}

struct Family {
	const char *name;       // profile list name in model$profiles
	const char *count_by;   // the parameter whose length gives the number of profiles
	int32_t type;
	const char *shape[5];   // parameter names of profit_amd_profile::shape
};

// the reading order of src/r_profit.cpp:620-629 and each list_to_* routine
const Family kFamilies[] = {
	{"sersic", "xcen", PROFIT_AMD_SERSIC, {"re", "nser", nullptr, nullptr, nullptr}},
	{"moffat", "xcen", PROFIT_AMD_MOFFAT, {"fwhm", "con", nullptr, nullptr, nullptr}},
	{"ferrer", "xcen", PROFIT_AMD_FERRER, {"rout", "a", "b", nullptr, nullptr}},
	{"ferrers", "xcen", PROFIT_AMD_FERRER, {"rout", "a", "b", nullptr, nullptr}},
	{"coresersic", "xcen", PROFIT_AMD_CORESERSIC, {"re", "rb", "nser", "a", "b"}},
	{"king", "xcen", PROFIT_AMD_KING, {"rc", "rt", "a", nullptr, nullptr}},
	{"brokenexp", "xcen", PROFIT_AMD_BROKENEXP, {"h1", "h2", "rb", "a", nullptr}},
	{"sky", "bg", PROFIT_AMD_SKY, {"bg", nullptr, nullptr, nullptr, nullptr}},
	{"psf", "xcen", PROFIT_AMD_PSF, {nullptr, nullptr, nullptr, nullptr, nullptr}},
	{"null", "convolve", PROFIT_AMD_NULL, {nullptr, nullptr, nullptr, nullptr, nullptr}},
};

// src/r_profit.cpp:239-292: every profile of model$profiles, family by family
std::vector<profit_amd_profile> read_profiles(SEXP profiles)
{
	std::vector<profit_amd_profile> out;
	for (const Family &f : kFamilies) {
		SEXP l = list_element(profiles, f.name);
		if (l == R_NilValue) continue;
		SEXP count = list_element(l, f.count_by);
		if (count == R_NilValue) continue;
		const unsigned n = (unsigned)Rf_length(count);
		for (unsigned i = 0; i < n; ++i) {
			profit_amd_profile p;
			check(profit_amd_profile_defaults(f.type, &p));
			read_bool(l, "convolve", i, &p.convolve);
			if (f.type == PROFIT_AMD_PSF) {
				read_real(l, "xcen", i, &p.xcen);
				read_real(l, "ycen", i, &p.ycen);
				read_real(l, "mag", i, &p.mag);
			}
			else if (f.type != PROFIT_AMD_NULL && f.type != PROFIT_AMD_SKY) {
				read_radial(l, i, p);
			}
			for (int s = 0; s < 5; ++s)
				if (f.shape[s]) read_real(l, f.shape[s], i, &p.shape[s]);
			if (f.type == PROFIT_AMD_SERSIC) read_bool(l, "rescale_flux", i, &p.rescale_flux);
			out.push_back(p);
		}
	}
	return out;
}

// ------------------------------------------------------------------ convolvers

// what R_profit_make_convolver's external pointer holds
struct ConvolverRef {
	int32_t type;
	int32_t device;
};

int32_t convolver_type(const char *name)
{
	const std::string t = name;
	if (t == "brute") return PROFIT_AMD_CONV_BRUTE;
	if (t == "brute-old") return PROFIT_AMD_CONV_BRUTE_OLD;
	if (t == "fft" || t == "fftw") return PROFIT_AMD_CONV_FFT;
	if (t == "null") return PROFIT_AMD_CONV_NULL;
	if (t == "brute-bf16x3") return PROFIT_AMD_CONV_BRUTE_BF16X3;
	throw std::invalid_argument("Convolver of type " + t + " is not supported");
}

void convolver_finalizer(SEXP ptr)
{
	delete static_cast<ConvolverRef *>(R_ExternalPtrAddr(ptr));
	R_ClearExternalPtr(ptr);
}

const ConvolverRef &unwrap_convolver(SEXP c)
{
	if (TYPEOF(c) != EXTPTRSXP) throw std::invalid_argument("Given convolver not of proper type");
	const ConvolverRef *r = static_cast<const ConvolverRef *>(R_ExternalPtrAddr(c));
	if (!r) throw std::invalid_argument("No Convolver found in convolver object");
	return *r;
}

// ------------------------------------------------------------------ model

struct ModelIn {
	profit_amd_model_desc desc;
	ImageIn psf;
	std::vector<uint8_t> mask;
	std::vector<profit_amd_profile> profiles;
};

// src/r_profit.cpp:549-629: the model list's settings and profiles
void read_model(SEXP model_list, ModelIn &m)
{
	std::memset(&m.desc, 0, sizeof(m.desc));
	profit_amd_model_desc &d = m.desc;
	SEXP dims = list_element(model_list, "dimensions");
	if (dims == R_NilValue) throw std::invalid_argument("No dimensions provided in the model");
	d.width = (uint32_t)INTEGER(dims)[0];
	d.height = (uint32_t)INTEGER(dims)[1];
	SEXP magzero = list_element(model_list, "magzero");
	if (magzero == R_NilValue) throw std::invalid_argument("No magzero provided in the model");
	d.magzero = Rf_asReal(magzero);
	d.scale_x = d.scale_y = 1;
	SEXP sx = list_element(model_list, "scale_x"), sy = list_element(model_list, "scale_y");
	if (sx != R_NilValue) d.scale_x = Rf_asReal(sx);
	if (sy != R_NilValue) d.scale_y = Rf_asReal(sy);
	SEXP fs = list_element(model_list, "finesampling");
	d.finesampling = fs != R_NilValue ? (uint32_t)Rf_asInteger(fs) : 1;
	SEXP rf = list_element(model_list, "returnfine");
	d.return_finesampled = rf != R_NilValue ? Rf_asLogical(rf) != 0 : 0;
	SEXP crop = list_element(model_list, "crop");
	d.crop = crop != R_NilValue ? Rf_asLogical(crop) != 0 : 1;
	d.adjust_mask = 1;
	d.convolver = PROFIT_AMD_CONV_BRUTE;  // Model::ensure_convolver's default
	SEXP psf = list_element(model_list, "psf");
	if (psf != R_NilValue) {
		m.psf = read_image(psf);
		d.psf = m.psf.data;
		d.psf_w = m.psf.w;
		d.psf_h = m.psf.h;
		// the PSF's pixel scale is the image's (src/r_profit.cpp:586-587)
		d.psf_scale_x = d.scale_x;
		d.psf_scale_y = d.scale_y;
	}
	SEXP calc = list_element(model_list, "calcregion");
	if (calc != R_NilValue) {
		m.mask = read_mask(calc, &d.mask_w, &d.mask_h);
		d.mask = m.mask.data();
		d.adjust_mask = Rf_asLogical(list_element(model_list, "adjust_calcregion")) != 0;
	}
	SEXP conv = list_element(model_list, "convolver");
	if (conv != R_NilValue) {
		const ConvolverRef &c = unwrap_convolver(conv);
		d.convolver = c.type;
		d.device = c.device;
	}
	// "openclenv" and "omp_threads" have no effect here: the model runs on the GPU
	SEXP profiles = list_element(model_list, "profiles");
	if (profiles == R_NilValue) throw std::invalid_argument("No profiles provided in the model");
	m.profiles = read_profiles(profiles);
	if (m.profiles.empty()) throw std::invalid_argument("No valid profiles found in profiles list");
}

SEXP make_model(SEXP model_list, SEXP previous_image)
{
	ModelIn m;
	read_model(model_list, m);
	const int32_t n = (int32_t)m.profiles.size();
	uint32_t dw, dh, ow, oh;
	check(profit_amd_drawing_dims(&m.desc, m.profiles.data(), n, &dw, &dh));
	check(profit_amd_output_dims(&m.desc, m.profiles.data(), n, &ow, &oh));
	// src/r_profit.cpp:647-670: the previous image is written in place when it
	// has the drawing dimensions and the result keeps them (no padding to crop,
	// no downsampling)
	const bool in_place = previous_image != R_NilValue && TYPEOF(previous_image) == REALSXP &&
	                      (uint32_t)Rf_nrows(previous_image) == dw && (uint32_t)Rf_ncols(previous_image) == dh &&
	                      ow == dw && oh == dh;
	int nprot = 0;
	SEXP image = previous_image;
	if (!in_place) {
		image = PROTECT(Rf_allocMatrix(REALSXP, (int)ow, (int)oh));
		++nprot;
	}
	uint32_t rw, rh;
	int32_t offx, offy;
	const int rc = profit_amd_make_model(&m.desc, m.profiles.data(), n, REAL(image), (size_t)ow * oh, &rw, &rh, &offx,
	                                     &offy);
	if (rc != 0) {
		UNPROTECT(nprot);
		throw std::runtime_error(std::string("Error while calculating model: ") + profit_amd_last_error());
	}
	SEXP offset = PROTECT(Rf_list2(Rf_ScalarInteger(offx), Rf_ScalarInteger(offy)));
	SEXP ret = PROTECT(Rf_list2(image, offset));
	UNPROTECT(nprot + 2);
	return ret;
}

// ------------------------------------------------------------------ likelihood engine

void engine_finalizer(SEXP ptr)
{
	profit_amd_engine *e = static_cast<profit_amd_engine *>(R_ExternalPtrAddr(ptr));
	if (e) profit_amd_engine_destroy(e);
	R_ClearExternalPtr(ptr);
}

profit_amd_engine *unwrap_engine(SEXP e)
{
	if (TYPEOF(e) != EXTPTRSXP) throw std::invalid_argument("Given engine not of proper type");
	profit_amd_engine *p = static_cast<profit_amd_engine *>(R_ExternalPtrAddr(e));
	if (!p) throw std::invalid_argument("No engine found in engine object");
	return p;
}

int32_t like_func(const char *name)
{
	// profitParseLikefunc (R/profitUtility.R:99-122), case-insensitive; the
	// skewed t ("st") fits its own parameters in R and is not on the engine
	std::string f = name;
	for (char &ch : f) ch = (char)std::tolower((unsigned char)ch);
	if (f == "norm" || f == "normal") return PROFIT_AMD_LIKE_NORM;
	if (f == "chisq" || f == "chi-sq") return PROFIT_AMD_LIKE_CHISQ;
	if (f == "t" || f == "student" || f == "student-t") return PROFIT_AMD_LIKE_T;
	if (f == "pois" || f == "poisson" || f == "cash" || f == "c") return PROFIT_AMD_LIKE_POIS;
	throw std::invalid_argument("Likelihood function " + f + " is not supported by the engine");
}

// the instances' profiles: a list with one model$profiles list per instance
std::vector<profit_amd_profile> read_batch_profiles(SEXP lists)
{
	std::vector<profit_amd_profile> out;
	for (int i = 0; i < Rf_length(lists); ++i) {
		const std::vector<profit_amd_profile> p = read_profiles(VECTOR_ELT(lists, i));
		out.insert(out.end(), p.begin(), p.end());
	}
	return out;
}

SEXP engine_like(SEXP engine, SEXP profile_lists, bool grad)
{
	profit_amd_engine *e = unwrap_engine(engine);
	const std::vector<profit_amd_profile> p = read_batch_profiles(profile_lists);
	const int ninst = Rf_length(profile_lists);
	SEXP ll = PROTECT(Rf_allocVector(REALSXP, ninst));
	if (!grad) {
		const int rc = profit_amd_engine_like(e, p.data(), REAL(ll));
		UNPROTECT(1);
		check(rc);
		return ll;
	}
	// d logL / d (xcen, ycen, mag, ang, axrat, box, shape[0..4]) per component
	SEXP g = PROTECT(Rf_allocMatrix(REALSXP, PROFIT_AMD_NGRAD, (int)p.size()));
	const int rc = profit_amd_engine_like_grad(e, p.data(), REAL(ll), REAL(g));
	SEXP ret = PROTECT(Rf_list2(ll, g));
	UNPROTECT(3);
	check(rc);
	return ret;
}

}  // namespace

extern "C" {

/* ------------------------------------------------------------------ src/r_profit.cpp:774-850 */

SEXP R_profit_make_model(SEXP model_list, SEXP previous_image)
{
	return guarded("R_profit_make_model", [&] { return make_model(model_list, previous_image); });
}

SEXP R_profit_convolvers()
{
	// ProFit maps "fftw" to "fft" (R/profitConvolver.R:41-44); the FFT
	// convolver is always present here, and so are the GPU's own variants
	return guarded("R_profit_convolvers", [&] {
		const char *names[] = {"brute", "fftw", "brute-old", "null", "brute-bf16x3"};
		SEXP out = PROTECT(Rf_allocVector(STRSXP, 5));
		for (int k = 0; k < 5; ++k) SET_STRING_ELT(out, k, Rf_mkChar(names[k]));
		UNPROTECT(1);
		return out;
	});
}

SEXP R_profit_make_convolver(SEXP type, SEXP image_dimensions, SEXP psf, SEXP reuse_psf_fft, SEXP fft_effort,
                             SEXP omp_threads, SEXP openclenv)
{
	// src/r_profit.cpp:504-547; the FFT's PSF spectrum is always reused, and
	// effort, threads and OpenCL have no effect on the GPU
	(void)image_dimensions;
	(void)psf;
	(void)reuse_psf_fft;
	(void)fft_effort;
	(void)omp_threads;
	(void)openclenv;
	return guarded("R_profit_make_convolver", [&] {
		ConvolverRef *r = new ConvolverRef{convolver_type(CHAR(STRING_ELT(type, 0))), 0};
		SEXP ptr = PROTECT(R_MakeExternalPtr(r, Rf_install("Convolver"), R_NilValue));
		R_RegisterCFinalizerEx(ptr, convolver_finalizer, TRUE);
		UNPROTECT(1);
		return ptr;
	});
}

SEXP R_profit_convolve(SEXP convolver, SEXP r_image, SEXP r_psf, SEXP r_mask)
{
	// src/r_profit.cpp:684-706: a vector of the source's size
	return guarded("R_profit_convolve", [&] {
		const ConvolverRef &c = unwrap_convolver(convolver);
		const ImageIn src = read_image(r_image), krn = read_image(r_psf);
		std::vector<uint8_t> mask;
		if (r_mask != R_NilValue) {
			uint32_t mw, mh;
			mask = read_mask(r_mask, &mw, &mh);
			if (mw != src.w || mh != src.h) throw std::invalid_argument("mask and image dimensions differ");
		}
		SEXP out = PROTECT(Rf_allocVector(REALSXP, (R_xlen_t)src.w * src.h));
		const int rc = profit_amd_convolve(c.type, src.data, src.w, src.h, krn.data, krn.w, krn.h,
		                                   mask.empty() ? nullptr : mask.data(), REAL(out), c.device);
		UNPROTECT(1);
		check(rc);
		return out;
	});
}

SEXP R_profit_has_openmp() { return Rf_ScalarLogical(FALSE); }

SEXP R_profit_has_fftw()
{
	// the FFT convolver is available (the GPU's own FFT, not FFTW)
	return Rf_ScalarLogical(TRUE);
}

SEXP R_profit_openclenv_info()
{
	// no OpenCL platforms: the model runs on the GPU through HIP
	return Rf_allocVector(VECSXP, 0);
}

SEXP R_profit_openclenv(SEXP plat_idx, SEXP dev_idx, SEXP use_dbl)
{
	(void)plat_idx;
	(void)dev_idx;
	(void)use_dbl;
	Rf_error("OpenCL is not supported: the model runs on the GPU through HIP\n");
	return R_NilValue;
}

SEXP R_profit_downsample(SEXP img, SEXP factor)
{
	// src/r_profit.cpp:708-729: SUM mode
	return guarded("Error while downsampling image", [&] {
		const ImageIn im = read_image(img);
		if (!im.data) return R_NilValue;
		const uint32_t f = (uint32_t)Rf_asInteger(factor);
		if (f == 0) throw std::invalid_argument("downsampling factor is 0");
		const uint32_t w = (im.w + f - 1) / f, h = (im.h + f - 1) / f;
		SEXP out = PROTECT(Rf_allocMatrix(REALSXP, (int)w, (int)h));
		uint32_t ow, oh;
		const int rc = profit_amd_downsample(im.data, im.w, im.h, f, 1, REAL(out), (size_t)w * h, &ow, &oh);
		UNPROTECT(1);
		check(rc);
		return out;
	});
}

SEXP R_profit_upsample(SEXP img, SEXP factor)
{
	// src/r_profit.cpp:731-755: COPY mode
	return guarded("Error while upsampling image", [&] {
		const ImageIn im = read_image(img);
		if (!im.data) return R_NilValue;
		const uint32_t f = (uint32_t)Rf_asInteger(factor);
		SEXP out = PROTECT(Rf_allocMatrix(REALSXP, (int)(im.w * f), (int)(im.h * f)));
		uint32_t ow, oh;
		const int rc = profit_amd_upsample(im.data, im.w, im.h, f, 0, REAL(out), (size_t)im.w * f * im.h * f, &ow, &oh);
		UNPROTECT(1);
		check(rc);
		return out;
	});
}

SEXP R_profit_adjust_mask(SEXP r_mask, SEXP r_img_dims, SEXP r_psf, SEXP r_finesampling)
{
	// src/r_profit.cpp:757-772
	return guarded("R_profit_adjust_mask", [&] {
		uint32_t mw, mh;
		const std::vector<uint8_t> mask = read_mask(r_mask, &mw, &mh);
		const ImageIn psf = read_image(r_psf);
		const uint32_t fs = (uint32_t)Rf_asInteger(r_finesampling);
		const uint32_t iw = (uint32_t)INTEGER(r_img_dims)[0], ih = (uint32_t)INTEGER(r_img_dims)[1];
		const size_t cap = (size_t)(mw * fs + psf.w) * (mh * fs + psf.h);
		std::vector<uint8_t> out(cap);
		uint32_t ow, oh;
		check(profit_amd_adjust_mask(mask.data(), mw, mh, iw, ih, psf.data, psf.w, psf.h, fs, out.data(), cap, &ow,
		                             &oh));
		SEXP r = PROTECT(Rf_allocMatrix(LGLSXP, (int)ow, (int)oh));
		int *v = LOGICAL(r);
		for (size_t k = 0; k < (size_t)ow * oh; ++k) v[k] = out[k];
		UNPROTECT(1);
		return r;
	});
}

SEXP R_profit_clear_cache()
{
	return guarded("R_profit_clear_cache", [&] {
		check(profit_amd_clear_cache());
		return R_NilValue;
	});
}

/* ------------------------------------------------------------------ the likelihood engine (new) */

SEXP R_profit_amd_engine(SEXP device)
{
	return guarded("R_profit_amd_engine", [&] {
		profit_amd_engine *e = nullptr;
		check(profit_amd_engine_create(device == R_NilValue ? 0 : Rf_asInteger(device), &e));
		SEXP ptr = PROTECT(R_MakeExternalPtr(e, Rf_install("profit_amd_engine"), R_NilValue));
		R_RegisterCFinalizerEx(ptr, engine_finalizer, TRUE);
		UNPROTECT(1);
		return ptr;
	});
}

SEXP R_profit_amd_add_galaxy(SEXP engine, SEXP data)
{
	// the fields of profitSetupData's result (R/profitSetupData.R:298-330):
	// image, sigma, psf, calcregion (already adjusted), region, like.func, magzero
	return guarded("R_profit_amd_add_galaxy", [&] {
		profit_amd_engine *e = unwrap_engine(engine);
		SEXP fine = list_element(data, "finesample");
		if (fine != R_NilValue && Rf_asInteger(fine) != 1)
			throw std::invalid_argument("the engine takes galaxies without finesampling");
		const ImageIn image = read_image(list_element(data, "image"));
		const ImageIn sigma = read_image(list_element(data, "sigma"));
		const ImageIn psf = read_image(list_element(data, "psf"));
		if (!image.data || !sigma.data) throw std::invalid_argument("image and sigma are required");
		profit_amd_galaxy_desc g;
		std::memset(&g, 0, sizeof(g));
		g.width = image.w;
		g.height = image.h;
		g.scale_x = g.scale_y = 1;
		g.magzero = Rf_asReal(list_element(data, "magzero"));
		g.psf = psf.data;
		g.psf_w = psf.w;
		g.psf_h = psf.h;
		std::vector<uint8_t> calc, region;
		SEXP r_calc = list_element(data, "calcregion");
		if (r_calc != R_NilValue) {
			calc = read_mask(r_calc, &g.calc_w, &g.calc_h);
			g.calcregion = calc.data();
		}
		SEXP r_region = list_element(data, "region");
		if (r_region != R_NilValue) {
			uint32_t rw, rh;
			region = read_mask(r_region, &rw, &rh);
			if (rw != image.w || rh != image.h) throw std::invalid_argument("region and image dimensions differ");
			g.region = region.data();
		}
		g.data = image.data;
		g.sigma = sigma.data;
		SEXP lf = list_element(data, "like.func");
		g.likefunc = lf == R_NilValue ? PROFIT_AMD_LIKE_NORM : like_func(CHAR(STRING_ELT(lf, 0)));
		int32_t id = -1;
		check(profit_amd_engine_add_galaxy(e, &g, &id));
		return Rf_ScalarInteger(id);
	});
}

SEXP R_profit_amd_set_batch(SEXP engine, SEXP galaxies, SEXP ncomp)
{
	return guarded("R_profit_amd_set_batch", [&] {
		profit_amd_engine *e = unwrap_engine(engine);
		if (Rf_length(galaxies) != Rf_length(ncomp)) throw std::invalid_argument("instance arrays differ in length");
		check(profit_amd_engine_set_batch(e, INTEGER(galaxies), INTEGER(ncomp), Rf_length(galaxies)));
		return R_NilValue;
	});
}

SEXP R_profit_amd_like(SEXP engine, SEXP profile_lists)
{
	return guarded("R_profit_amd_like", [&] { return engine_like(engine, profile_lists, false); });
}

SEXP R_profit_amd_like_grad(SEXP engine, SEXP profile_lists)
{
	return guarded("R_profit_amd_like_grad", [&] { return engine_like(engine, profile_lists, true); });
}

static const R_CallMethodDef kCallMethods[] = {
	{"R_profit_make_model", (DL_FUNC)&R_profit_make_model, 2},
	{"R_profit_convolvers", (DL_FUNC)&R_profit_convolvers, 0},
	{"R_profit_make_convolver", (DL_FUNC)&R_profit_make_convolver, 7},
	{"R_profit_convolve", (DL_FUNC)&R_profit_convolve, 4},
	{"R_profit_has_openmp", (DL_FUNC)&R_profit_has_openmp, 0},
	{"R_profit_has_fftw", (DL_FUNC)&R_profit_has_fftw, 0},
	{"R_profit_openclenv_info", (DL_FUNC)&R_profit_openclenv_info, 0},
	{"R_profit_openclenv", (DL_FUNC)&R_profit_openclenv, 3},
	{"R_profit_downsample", (DL_FUNC)&R_profit_downsample, 2},
	{"R_profit_upsample", (DL_FUNC)&R_profit_upsample, 2},
	{"R_profit_adjust_mask", (DL_FUNC)&R_profit_adjust_mask, 4},
	{"R_profit_clear_cache", (DL_FUNC)&R_profit_clear_cache, 0},
	{"R_profit_amd_engine", (DL_FUNC)&R_profit_amd_engine, 1},
	{"R_profit_amd_add_galaxy", (DL_FUNC)&R_profit_amd_add_galaxy, 2},
	{"R_profit_amd_set_batch", (DL_FUNC)&R_profit_amd_set_batch, 3},
	{"R_profit_amd_like", (DL_FUNC)&R_profit_amd_like, 2},
	{"R_profit_amd_like_grad", (DL_FUNC)&R_profit_amd_like_grad, 2},
	{NULL, NULL, 0}};

void R_init_ProFit(DllInfo *dll)
{
	// libprofit's init() and its PROFIT_HOME cache directory
	// (src/r_profit.cpp:852-897) have no counterpart: nothing is cached on disk
	R_registerRoutines(dll, NULL, kCallMethods, NULL, NULL);
	R_useDynamicSymbols(dll, FALSE);
}

void R_unload_ProFit(DllInfo *dll)
{
	(void)dll;
	profit_amd_clear_cache();
}

}  // extern "C"
