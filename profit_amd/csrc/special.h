// Special functions for the per-component set-up of radial profiles.
//
// libprofit obtains these from R's nmath (src/libprofit/src/profit/utils.cpp:193-258):
//   qgamma(p, shape)  = Rf_qgamma(p, shape, 1, 1, 0)
//   pgamma(q, shape)  = Rf_pgamma(q, shape, 1, 1, 0)
//   gammafn(x)        = Rf_gammafn(x)
//   beta(a, b)        = Rf_beta(a, b)
//   integrate_qags / integrate_qagi: QUADPACK dqags / dqagi, epsabs = epsrel = 1e-4,
//   limit = 100.
// Everything here is __host__ __device__ so that the same FP64 code sets up
// components on the host (single Model::evaluate) and on the GPU (batches).
// Algorithms: P(a,x) by its power series (x < a+1) or Legendre continued
// fraction; the quantile by bracketed Newton steps with a Halley correction;
// QUADPACK dqagse/dqagie (GK21 / transformed GK15, epsilon-algorithm
// extrapolation) as published by Piessens et al. (1983).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace profit_amd {
namespace sf {

#define PA_HD __host__ __device__ inline

PA_HD bool is_inf(double x) { return ::fabs(x) == HUGE_VAL; }
PA_HD double lgam(double x) { return ::lgamma(x); }
PA_HD double gam(double x) { return ::tgamma(x); }

PA_HD double beta(double a, double b)
{
	const double xmax = 171.61447887182298;  // Gamma overflow point used by R
	if (a < 0 || b < 0) return NAN;
	if (a == 0 || b == 0) return INFINITY;
	if (a + b < xmax) return (1 / gam(a + b)) * (gam(a) * gam(b));
	return ::exp(lgam(a) + lgam(b) - lgam(a + b));
}

PA_HD double digamma(double x)
{
	double acc = 0;
	while (x < 12) {
		acc -= 1 / x;
		x += 1;
	}
	const double z = 1 / (x * x);
	// Bernoulli-number asymptotic series
	const double series = z * (1.0 / 12 - z * (1.0 / 120 - z * (1.0 / 252 - z * (1.0 / 240 - z * (1.0 / 132)))));
	return acc + ::log(x) - 0.5 / x - series;
}

// log(x^a e^-x / Gamma(a)); lga = lgam(a), computed once by the callers
// that evaluate many x for one a
PA_HD double log_kernel(double a, double x, double lga) { return a * ::log(x) - x - lga; }

// regularised lower incomplete gamma by series
PA_HD double p_series(double a, double x, double lga)
{
	double term = 1 / a, sum = term, an = a;
	for (int k = 0; k < 2000; ++k) {
		an += 1;
		term *= x / an;
		sum += term;
		if (::fabs(term) <= ::fabs(sum) * 5e-18) break;
	}
	return sum * ::exp(log_kernel(a, x, lga));
}

// regularised upper incomplete gamma by continued fraction
PA_HD double q_cfrac(double a, double x, double lga)
{
	const double tiny = 1e-300;
	double b = x + 1 - a;
	double c = 1 / tiny;
	double d = 1 / b;
	double h = d;
	for (int i = 1; i < 2000; ++i) {
		const double an = -i * (i - a);
		b += 2;
		d = an * d + b;
		if (::fabs(d) < tiny) d = tiny;
		c = b + an / c;
		if (::fabs(c) < tiny) c = tiny;
		d = 1 / d;
		const double delta = d * c;
		h *= delta;
		if (::fabs(delta - 1) <= 5e-18) break;
	}
	return ::exp(log_kernel(a, x, lga)) * h;
}

PA_HD double pgamma(double x, double a, double lga)
{
	if (!(x > 0)) return 0;
	if (is_inf(x)) return 1;
	return x < a + 1 ? p_series(a, x, lga) : 1 - q_cfrac(a, x, lga);
}
PA_HD double pgamma(double x, double a) { return pgamma(x, a, lgam(a)); }

PA_HD double qgamma_upper(double x, double a, double lga)
{
	if (!(x > 0)) return 1;
	return x < a + 1 ? 1 - p_series(a, x, lga) : q_cfrac(a, x, lga);
}

// quantile of the unit-scale gamma distribution
PA_HD double qgamma(double p, double a)
{
	if (!(p > 0)) return 0;
	if (!(p < 1)) return INFINITY;
	double x;
	if (a >= 1) {
		// Wilson-Hilferty start with a rational normal-quantile approximation
		const double t = ::sqrt(-2 * ::log(p < 0.5 ? p : 1 - p));
		double z = t - (2.515517 + t * (0.802853 + t * 0.010328)) /
		                   (1 + t * (1.432788 + t * (0.189269 + t * 0.001308)));
		if (p < 0.5) z = -z;
		const double c = 1 / (9 * a);
		const double base = 1 - c + z * ::sqrt(c);
		x = base > 0 ? a * base * base * base : 1e-3 * a;
	}
	else {
		x = ::exp((::log(p) + lgam(a + 1)) / a);
		if (!(x > 0)) x = 1e-300;
	}
	const bool upper = p >= 0.9;
	const double lga = lgam(a);
	double lo = 0, hi = INFINITY;
	for (int it = 0; it < 100; ++it) {
		const double f = upper ? (1 - p) - qgamma_upper(x, a, lga) : pgamma(x, a, lga) - p;
		if (f == 0) break;
		if (f < 0) lo = x;
		else hi = x;
		const double dens = ::exp(log_kernel(a, x, lga)) / x;
		if (!(dens > 0)) break;
		const double newton = f / dens;
		const double curv = (a - 1) / x - 1;
		double h = 1 - 0.5 * newton * curv;
		if (!(h > 0.5)) h = 0.5;
		// upper tail (the set-up's switch and maximum radii): a step below the
		// resolution of x is convergence, also where rounding puts x - step
		// on a bracket end, which the test below would answer with bisection
		// steps (up to 49 iterations instead of 9)
		if (upper && ::fabs(newton / h) <= 2e-16 * x) break;
		double xn = x - newton / h;
		if (!(xn > lo && xn < hi)) xn = is_inf(hi) ? 2 * x : 0.5 * (lo + hi);
		const double dx = ::fabs(xn - x);
		x = xn;
		if (dx <= 2e-16 * x) break;
	}
	return x;
}

// dP(a,x)/da from the term-wise derivative of P = e^-x sum x^(a+k)/Gamma(a+k+1):
// term k is x^(a+k) e^-x / Gamma(a+k+1) * (ln x - psi(a+k+1)), both factors
// carried from k to k+1 by Gamma(z+1) = z Gamma(z) and psi(z+1) = psi(z) + 1/z
PA_HD double dpgamma_da(double x, double a)
{
	if (!(x > 0)) return 0;
	const double lx = ::log(x);
	double e = ::exp(a * lx - x - lgam(a + 1));
	double psi = digamma(a + 1);
	double sum = 0;
	for (int k = 0; k < 4000; ++k) {
		const double term = e * (lx - psi);
		sum += term;
		if (k > x && ::fabs(term) <= 1e-18 * ::fabs(sum)) break;
		const double z = a + k + 1;
		e *= x / z;
		psi += 1 / z;
	}
	return sum;
}

// ------------------------------------------------------------------ QUADPACK

struct QuadNodes {
	// 21-point Gauss-Kronrod (10-point Gauss embedded)
	static PA_HD double xk21(int i)
	{
		const double v[11] = {0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
		                      0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
		                      0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
		                      0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
		                      0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
		                      0.0};
		return v[i];
	}
	static PA_HD double wk21(int i)
	{
		const double v[11] = {0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
		                      0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
		                      0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
		                      0.123491976262065851077600525452038, 0.134709217311473325928054001771707,
		                      0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
		                      0.149445554002916905664936468389821};
		return v[i];
	}
	static PA_HD double wg10(int i)
	{
		const double v[5] = {0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
		                     0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
		                     0.295524224714752870173892994651338};
		return v[i];
	}
	// 15-point Kronrod (7-point Gauss) for the transformed infinite range
	static PA_HD double xk15(int i)
	{
		const double v[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
		                     0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
		                     0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
		                     0.207784955007898467600689403773245, 0.0};
		return v[i];
	}
	static PA_HD double wk15(int i)
	{
		const double v[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
		                     0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
		                     0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
		                     0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
		return v[i];
	}
	static PA_HD double wg7(int i)
	{
		const double v[8] = {0.0, 0.129484966168869693270611432679082, 0.0, 0.279705391489276667901467771423780,
		                     0.0, 0.381830050505118944950369775488975, 0.0, 0.417959183673469387755102040816327};
		return v[i];
	}
};

PA_HD void rule_error(double resk, double resg, double half, double resabs, double resasc, double *err)
{
	double e = ::fabs((resk - resg) * half);
	if (resasc != 0 && e != 0) {
		const double r = 200 * e / resasc;
		e = resasc * ::fmin(1.0, ::pow(r, 1.5));
	}
	if (resabs > DBL_MIN / (50 * DBL_EPSILON)) {
		const double floor_ = 50 * DBL_EPSILON * resabs;
		if (floor_ > e) e = floor_;
	}
	*err = e;
}

// Kronrod rule on [a,b] (finite) or on (a,b] of the map t -> bound + (1-t)/t.
// kWave (device code, a whole wavefront with the same arguments): the 21 or
// 15 integrand values are computed on as many lanes at once and broadcast;
// every lane then forms the sums below in the same order, so the results are
// those of the serial rule.
template <bool kWave = false, class F>
PA_HD void kronrod(const F &f, bool infinite, double bound, double a, double b, double *result,
                   double *err, double *resabs, double *resasc)
{
	const double centre = 0.5 * (a + b), half = 0.5 * (b - a);
	double fa[10], fb[10];
	// kWave: lane k holds the value at abscissa k of the serial order, k < 10
	// (finite; 7 infinite) centre - half x_k, then the centre, then centre +
	// half x_k; the node of lane k selected without a table lookup
	double fw = 0;
	if constexpr (kWave) {
		const int lane = __lane_id(), nside = infinite ? 7 : 10;
		const int k = lane < nside ? lane : lane - nside - 1;
		double xk = 0;
#pragma unroll
		for (int j = 0; j < 10; ++j)
			if (j == k) xk = infinite ? (j < 7 ? QuadNodes::xk15(j) : 0.0) : QuadNodes::xk21(j);
		if (lane <= 2 * nside) {
			const double d = half * xk;
			if (!infinite) {
				fw = f(lane == nside ? centre : (lane < nside ? centre - d : centre + d));
			}
			else {
				const double u = lane == nside ? centre : (lane < nside ? centre - d : centre + d);
				fw = lane == nside ? (f(bound + (1 - centre) / centre) / centre) / centre : (f(bound + (1 - u) / u) / u) / u;
			}
		}
	}
	if (!infinite) {
		double fc;
		if constexpr (kWave) fc = __shfl(fw, 10);
		else fc = f(centre);
		double resg = 0, resk = QuadNodes::wk21(10) * fc;
		double sabs = ::fabs(resk);
		for (int j = 0; j < 5; ++j) {  // Gauss abscissae first (odd Kronrod indices)
			const int k = 2 * j + 1;
			const double d = half * QuadNodes::xk21(k);
			double f1, f2;
			if constexpr (kWave) {
				f1 = __shfl(fw, k);
				f2 = __shfl(fw, 11 + k);
			}
			else {
				f1 = f(centre - d);
				f2 = f(centre + d);
			}
			fa[k] = f1;
			fb[k] = f2;
			resg += QuadNodes::wg10(j) * (f1 + f2);
			resk += QuadNodes::wk21(k) * (f1 + f2);
			sabs += QuadNodes::wk21(k) * (::fabs(f1) + ::fabs(f2));
		}
		for (int j = 0; j < 5; ++j) {
			const int k = 2 * j;
			const double d = half * QuadNodes::xk21(k);
			double f1, f2;
			if constexpr (kWave) {
				f1 = __shfl(fw, k);
				f2 = __shfl(fw, 11 + k);
			}
			else {
				f1 = f(centre - d);
				f2 = f(centre + d);
			}
			fa[k] = f1;
			fb[k] = f2;
			resk += QuadNodes::wk21(k) * (f1 + f2);
			sabs += QuadNodes::wk21(k) * (::fabs(f1) + ::fabs(f2));
		}
		const double mean = 0.5 * resk;
		double sasc = QuadNodes::wk21(10) * ::fabs(fc - mean);
		for (int k = 0; k < 10; ++k) sasc += QuadNodes::wk21(k) * (::fabs(fa[k] - mean) + ::fabs(fb[k] - mean));
		*result = resk * half;
		*resabs = sabs * ::fabs(half);
		*resasc = sasc * ::fabs(half);
		rule_error(resk, resg, half, *resabs, *resasc, err);
		return;
	}
	const double tc = bound + (1 - centre) / centre;
	double fc;
	if constexpr (kWave) fc = __shfl(fw, 7);
	else fc = (f(tc) / centre) / centre;
	double resg = QuadNodes::wg7(7) * fc, resk = QuadNodes::wk15(7) * fc;
	double sabs = ::fabs(resk);
	for (int j = 0; j < 7; ++j) {
		const double d = half * QuadNodes::xk15(j);
		const double u1 = centre - d, u2 = centre + d;
		double f1, f2;
		if constexpr (kWave) {
			f1 = __shfl(fw, j);
			f2 = __shfl(fw, 8 + j);
		}
		else {
			f1 = (f(bound + (1 - u1) / u1) / u1) / u1;
			f2 = (f(bound + (1 - u2) / u2) / u2) / u2;
		}
		fa[j] = f1;
		fb[j] = f2;
		resg += QuadNodes::wg7(j) * (f1 + f2);
		resk += QuadNodes::wk15(j) * (f1 + f2);
		sabs += QuadNodes::wk15(j) * (::fabs(f1) + ::fabs(f2));
	}
	const double mean = 0.5 * resk;
	double sasc = QuadNodes::wk15(7) * ::fabs(fc - mean);
	for (int j = 0; j < 7; ++j) sasc += QuadNodes::wk15(j) * (::fabs(fa[j] - mean) + ::fabs(fb[j] - mean));
	*result = resk * half;
	*resasc = sasc * half;
	*resabs = sabs * half;
	rule_error(resk, resg, half, *resabs, *resasc, err);
}

constexpr int kQuadLimit = 100;

struct QuadState {
	double lo[kQuadLimit + 2], hi[kQuadLimit + 2], val[kQuadLimit + 2], err[kQuadLimit + 2];
	int order[kQuadLimit + 2];
	double eps_table[56];
	double last3[4];
};

// keep `order` sorted by decreasing error (1-based, QUADPACK dqpsrt)
PA_HD void order_errors(QuadState &s, int limit, int last, int &maxerr, double &ermax, int &nrmax)
{
	int *ord = s.order;
	const double *e = s.err;
	if (last <= 2) {
		ord[1] = 1;
		ord[2] = 2;
	}
	else {
		const double errmax = e[maxerr];
		if (nrmax != 1) {
			const int steps = nrmax - 1;
			for (int t = 1; t <= steps; ++t) {
				const int prev = ord[nrmax - 1];
				if (errmax <= e[prev]) break;
				ord[nrmax] = prev;
				--nrmax;
			}
		}
		int top = last;
		if (last > limit / 2 + 2) top = limit + 3 - last;
		const double errmin = e[last];
		const int bottom = top - 1;
		int i = nrmax + 1;
		bool placed = false;
		for (; i <= bottom; ++i) {
			const int cur = ord[i];
			if (errmax >= e[cur]) {
				placed = true;
				break;
			}
			ord[i - 1] = cur;
		}
		if (!placed) {
			ord[bottom] = maxerr;
			ord[top] = last;
		}
		else {
			ord[i - 1] = maxerr;
			int k = bottom;
			bool done = false;
			for (int j = i; j <= bottom; ++j) {
				const int cur = ord[k];
				if (errmin < e[cur]) {
					ord[k + 1] = last;
					done = true;
					break;
				}
				ord[k + 1] = cur;
				--k;
			}
			if (!done) ord[i] = last;
		}
	}
	maxerr = ord[nrmax];
	ermax = e[maxerr];
}

// Wynn epsilon algorithm (QUADPACK dqelg), 1-based table
PA_HD void extrapolate(double *tab, int &n, double &result, double &abserr, double *last3, int &nres)
{
	const double eps = DBL_EPSILON, big = DBL_MAX;
	++nres;
	abserr = big;
	result = tab[n];
	if (n >= 3) {
		const int limexp = 50;
		tab[n + 2] = tab[n];
		const int newelm = (n - 1) / 2;
		tab[n] = big;
		const int num = n;
		int k1 = n;
		bool converged = false;
		for (int i = 1; i <= newelm; ++i) {
			const int k2 = k1 - 1, k3 = k1 - 2;
			double res = tab[k1 + 2];
			const double e0 = tab[k3], e1 = tab[k2], e2 = res;
			const double e1abs = ::fabs(e1);
			const double d2 = e2 - e1, err2 = ::fabs(d2);
			const double tol2 = ::fmax(::fabs(e2), e1abs) * eps;
			const double d3 = e1 - e0, err3 = ::fabs(d3);
			const double tol3 = ::fmax(e1abs, ::fabs(e0)) * eps;
			if (err2 <= tol2 && err3 <= tol3) {
				result = res;
				abserr = err2 + err3;
				converged = true;
				break;
			}
			const double e3 = tab[k1];
			tab[k1] = e1;
			const double d1 = e1 - e3, err1 = ::fabs(d1);
			const double tol1 = ::fmax(e1abs, ::fabs(e3)) * eps;
			if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) {
				n = i + i - 1;
				break;
			}
			const double ss = 1 / d1 + 1 / d2 - 1 / d3;
			if (!(::fabs(ss * e1) > 1e-4)) {
				n = i + i - 1;
				break;
			}
			res = e1 + 1 / ss;
			tab[k1] = res;
			k1 -= 2;
			const double error = err2 + ::fabs(res - e2) + err3;
			if (error > abserr) continue;
			abserr = error;
			result = res;
		}
		if (!converged) {
			if (n == limexp) n = 2 * (limexp / 2) - 1;
			int ib = (num % 2 == 0) ? 2 : 1;
			const int ie = newelm + 1;
			for (int i = 1; i <= ie; ++i) {
				tab[ib] = tab[ib + 2];
				ib += 2;
			}
			if (num != n) {
				int src = num - n + 1;
				for (int i = 1; i <= n; ++i) tab[i] = tab[src++];
			}
			if (nres < 4) {
				last3[nres] = result;
				abserr = big;
			}
			else {
				abserr = ::fabs(result - last3[3]) + ::fabs(result - last3[2]) + ::fabs(result - last3[1]);
				last3[1] = last3[2];
				last3[2] = last3[3];
				last3[3] = result;
			}
		}
	}
	abserr = ::fmax(abserr, 5 * eps * ::fabs(result));
}

// dqagse (infinite = false, range [a,b]) and dqagie (infinite = true, [a, inf))
template <bool kWave = false, class F>
PA_HD double integrate(const F &f, double a, double b, bool infinite, QuadState &s, double epsabs = 1e-4,
                       double epsrel = 1e-4, int limit = kQuadLimit)
{
	const double eps = DBL_EPSILON, tiny = DBL_MIN, big = DBL_MAX;
	if (limit > kQuadLimit) limit = kQuadLimit;
	double result = 0, abserr = 0, defabs = 0, resabs = 0;
	s.lo[1] = infinite ? 0.0 : a;
	s.hi[1] = infinite ? 1.0 : b;
	kronrod<kWave>(f, infinite, a, s.lo[1], s.hi[1], &result, &abserr, &defabs, &resabs);
	const double dres = ::fabs(result);
	double errbnd = ::fmax(epsabs, epsrel * dres);
	int last = 1;
	int ier = 0;
	s.val[1] = result;
	s.err[1] = abserr;
	s.order[1] = 1;
	if (abserr <= 100 * eps * defabs && abserr > errbnd) ier = 2;
	if (limit == 1) ier = 1;
	if (ier != 0 || (abserr <= errbnd && abserr != resabs) || abserr == 0) return result;

	s.eps_table[1] = result;
	double errmax = abserr, area = result, errsum = abserr;
	int maxerr = 1, nrmax = 1, numrl2 = 2, nres = 0, ktmin = 0, ierro = 0;
	int iroff1 = 0, iroff2 = 0, iroff3 = 0;
	bool extrap = false, noext = false;
	double erlarg = 0, ertest = 0, correc = 0, small = 0;
	abserr = big;
	const int ksgn = (dres >= (1 - 50 * eps) * defabs) ? 1 : -1;
	bool sum_list = false;

	for (last = 2; last <= limit; ++last) {
		const double a1 = s.lo[maxerr], b1 = 0.5 * (s.lo[maxerr] + s.hi[maxerr]);
		const double a2 = b1, b2 = s.hi[maxerr];
		const double erlast = errmax;
		double area1, area2, error1, error2, r1, r2, defab1, defab2;
		kronrod<kWave>(f, infinite, a, a1, b1, &area1, &error1, &r1, &defab1);
		kronrod<kWave>(f, infinite, a, a2, b2, &area2, &error2, &r2, &defab2);
		const double area12 = area1 + area2, erro12 = error1 + error2;
		errsum = errsum + erro12 - errmax;
		area = area + area12 - s.val[maxerr];
		if (defab1 != error1 && defab2 != error2) {
			if (::fabs(s.val[maxerr] - area12) <= 1e-5 * ::fabs(area12) && erro12 >= 0.99 * errmax) {
				if (extrap) ++iroff2;
				else ++iroff1;
			}
			if (last > 10 && erro12 > errmax) ++iroff3;
		}
		s.val[maxerr] = area1;
		s.val[last] = area2;
		errbnd = ::fmax(epsabs, epsrel * ::fabs(area));
		if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
		if (iroff2 >= 5) ierro = 3;
		if (last == limit) ier = 1;
		if (::fmax(::fabs(a1), ::fabs(b2)) <= (1 + 100 * eps) * (::fabs(a2) + 1000 * tiny)) ier = 4;
		if (error2 > error1) {
			s.lo[maxerr] = a2;
			s.lo[last] = a1;
			s.hi[last] = b1;
			s.val[maxerr] = area2;
			s.val[last] = area1;
			s.err[maxerr] = error2;
			s.err[last] = error1;
		}
		else {
			s.lo[last] = a2;
			s.hi[maxerr] = b1;
			s.hi[last] = b2;
			s.err[maxerr] = error1;
			s.err[last] = error2;
		}
		order_errors(s, limit, last, maxerr, errmax, nrmax);
		if (errsum <= errbnd) {
			sum_list = true;
			break;
		}
		if (ier != 0) break;
		if (last == 2) {
			small = infinite ? 0.375 : ::fabs(b - a) * 0.375;
			erlarg = errsum;
			ertest = errbnd;
			s.eps_table[2] = area;
			continue;
		}
		if (noext) continue;
		erlarg -= erlast;
		if (::fabs(b1 - a1) > small) erlarg += erro12;
		if (!extrap) {
			if (::fabs(s.hi[maxerr] - s.lo[maxerr]) > small) continue;
			extrap = true;
			nrmax = 2;
		}
		if (ierro != 3 && erlarg > ertest) {
			const int id = nrmax;
			int jupbnd = last;
			if (last > 2 + limit / 2) jupbnd = limit + 3 - last;
			bool large_left = false;
			for (int k = id; k <= jupbnd; ++k) {
				maxerr = s.order[nrmax];
				errmax = s.err[maxerr];
				if (::fabs(s.hi[maxerr] - s.lo[maxerr]) > small) {
					large_left = true;
					break;
				}
				++nrmax;
			}
			if (large_left) continue;
		}
		++numrl2;
		s.eps_table[numrl2] = area;
		double reseps, abseps;
		extrapolate(s.eps_table, numrl2, reseps, abseps, s.last3, nres);
		++ktmin;
		if (ktmin > 5 && abserr < 1e-3 * errsum) ier = 5;
		if (abseps < abserr) {
			ktmin = 0;
			abserr = abseps;
			result = reseps;
			correc = erlarg;
			ertest = ::fmax(epsabs, epsrel * ::fabs(reseps));
			if (abserr <= ertest) break;
		}
		if (numrl2 == 1) noext = true;
		if (ier == 5) break;
		maxerr = s.order[1];
		errmax = s.err[maxerr];
		nrmax = 1;
		extrap = false;
		small *= 0.5;
		erlarg = errsum;
	}
	if (last > limit) last = limit;
	if (!sum_list) {
		if (abserr == big) {
			sum_list = true;
		}
		else if (ier + ierro != 0) {
			if (ierro == 3) abserr += correc;
			if (ier == 0) ier = 3;
			if (result != 0 && area != 0) {
				if (abserr / ::fabs(result) > errsum / ::fabs(area)) sum_list = true;
			}
			else if (abserr > errsum) {
				sum_list = true;
			}
			else if (area == 0) {
				return result;
			}
		}
	}
	if (sum_list) {
		result = 0;
		for (int k = 1; k <= last; ++k) result += s.val[k];
	}
	return result;
}

#undef PA_HD

}  // namespace sf
}  // namespace profit_amd
