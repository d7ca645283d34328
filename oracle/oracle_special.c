/*
 * THE ORACLE -- test infrastructure only.
 *
 * FP64 special functions and QUADPACK integrators used by the CPU
 * restatement of libprofit's model-image path.  Only tests, the smoke check
 * and the bench's CPU-baseline leg may call anything in this directory; the
 * product library never links it.
 *
 * Contracts followed (libprofit calls these through R's nmath):
 *   src/libprofit/src/profit/utils.cpp:193-195  qgamma(p, shape)  = Rf_qgamma(p, shape, 1, 1, 0)
 *   src/libprofit/src/profit/utils.cpp:197-199  pgamma(q, shape)  = Rf_pgamma(q, shape, 1, 1, 0)
 *   src/libprofit/src/profit/utils.cpp:201-203  gammafn(x)        = Rf_gammafn(x)
 *   src/libprofit/src/profit/utils.cpp:205-207  beta(a, b)        = Rf_beta(a, b)
 *   src/libprofit/src/profit/utils.cpp:222-258  integrate_qags / integrate_qagi:
 *       QUADPACK dqags / dqagi with epsabs = epsrel = 1e-4, limit = 100.
 *
 * R's nmath is not present; the algorithms here are the published ones:
 *   - regularised incomplete gamma P(a,x): power series for x < a+1, Legendre
 *     continued fraction (modified Lentz) otherwise;
 *   - its inverse by safeguarded Halley iteration on P(a,x) - p;
 *   - Gamma via the C library's tgamma; Beta as R defines it
 *     (Gamma ratio below Gamma's overflow point, exp(lbeta) above);
 *   - QUADPACK dqagse/dqagie with dqk21/dqk15i, dqpsrt and the epsilon
 *     algorithm dqelg (Piessens, de Doncker-Kapenga, Ueberhuber, Kahaner 1983).
 * Parity of these is pinned by tests against scipy.special and
 * scipy.integrate.quad (which wraps the same QUADPACK routines).
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------------ */
/* Gamma family                                                         */
/* ------------------------------------------------------------------ */

double or_gammafn(double x) { return tgamma(x); }

double or_lgammafn(double x) { return lgamma(x); }

double or_beta(double a, double b)
{
	/* R: beta() uses gammafn ratios while a+b < xmax, else exp(lbeta) */
	const double xmax = 171.61447887182298;
	if (a < 0 || b < 0) return NAN;
	if (a == 0 || b == 0) return INFINITY;
	if (a + b < xmax) return (1 / tgamma(a + b)) * (tgamma(a) * tgamma(b));
	return exp(lgamma(a) + lgamma(b) - lgamma(a + b));
}

double or_digamma(double x)
{
	/* recurrence up to x >= 10, then the asymptotic series */
	double r = 0;
	while (x < 10) { r -= 1 / x; x += 1; }
	double f = 1 / (x * x);
	double t = f * (-1.0 / 12 + f * (1.0 / 120 + f * (-1.0 / 252 + f * (1.0 / 240 + f * (-1.0 / 132
	           + f * (691.0 / 32760 + f * (-1.0 / 12)))))));
	return r + log(x) - 0.5 / x + t;
}

/* log of x^a e^-x / Gamma(a) */
static double log_gamma_kernel(double a, double x)
{
	return a * log(x) - x - lgamma(a);
}

/* P(a,x) by series, valid for x < a+1 */
static double pgamma_series(double a, double x)
{
	double ap = a, sum = 1 / a, del = sum;
	for (int n = 0; n < 100000; n++) {
		ap += 1;
		del *= x / ap;
		sum += del;
		if (fabs(del) < fabs(sum) * 1e-17) break;
	}
	return sum * exp(log_gamma_kernel(a, x));
}

/* Q(a,x) by continued fraction (modified Lentz), valid for x >= a+1 */
static double qgamma_cf(double a, double x)
{
	const double tiny = 1e-300;
	double b = x + 1 - a, c = 1 / tiny, d = 1 / b, h = d;
	for (int i = 1; i < 100000; i++) {
		double an = -i * (i - a);
		b += 2;
		d = an * d + b;
		if (fabs(d) < tiny) d = tiny;
		c = b + an / c;
		if (fabs(c) < tiny) c = tiny;
		d = 1 / d;
		double del = d * c;
		h *= del;
		if (fabs(del - 1) < 1e-17) break;
	}
	return exp(log_gamma_kernel(a, x)) * h;
}

double or_pgamma(double x, double a)
{
	if (x <= 0) return 0;
	if (isinf(x)) return 1;
	if (x < a + 1) return pgamma_series(a, x);
	return 1 - qgamma_cf(a, x);
}

static double or_qgamma_upper(double x, double a)
{
	if (x <= 0) return 1;
	if (x < a + 1) return 1 - pgamma_series(a, x);
	return qgamma_cf(a, x);
}

double or_qgamma(double p, double a)
{
	if (p <= 0) return 0;
	if (p >= 1) return INFINITY;
	/* initial guess: Wilson-Hilferty for a >= 1, small-x power law otherwise */
	double x;
	if (a >= 1) {
		/* normal quantile of p by a short rational approximation, refined below */
		double t = sqrt(-2 * log(p < 0.5 ? p : 1 - p));
		double z = t - (2.30753 + t * 0.27061) / (1 + t * (0.99229 + t * 0.04481));
		if (p < 0.5) z = -z;
		double c = 1 / (9 * a);
		x = a * pow(1 - c + z * sqrt(c), 3);
		if (x <= 0) x = 1e-3 * a;
	}
	else {
		double lg = lgamma(a + 1);
		x = exp((log(p) + lg) / a);
		if (x <= 0) x = 1e-300;
	}
	/* Halley iteration on f(x) = P(a,x) - p, evaluated through Q(a,x) when
	 * p is close to 1 so that f keeps its relative accuracy */
	double lo = 0, hi = INFINITY;
	for (int it = 0; it < 200; it++) {
		double f = (p < 0.9) ? or_pgamma(x, a) - p : (1 - p) - or_qgamma_upper(x, a);
		if (f == 0) break;
		if (f < 0) lo = x; else hi = x;
		double dens = exp(log_gamma_kernel(a, x)) / x; /* x^(a-1) e^-x / Gamma(a) */
		if (!(dens > 0)) break;
		double u = f / dens;
		double corr = (a - 1) / x - 1; /* f''/f' */
		double step = u / (1 - 0.5 * fmin(1.0, u * corr));
		double xn = x - step;
		if (!(xn > lo && xn < hi)) xn = isinf(hi) ? 2 * x : 0.5 * (lo + hi);
		double dx = fabs(xn - x);
		x = xn;
		if (dx <= 2e-16 * fabs(x)) break;
	}
	return x;
}

/* dP(a,x)/da by term-wise differentiation of the series
 * P = e^-x sum_k x^(a+k) / Gamma(a+k+1). */
double or_dpgamma_da(double x, double a)
{
	if (x <= 0) return 0;
	double sum = 0;
	double lx = log(x);
	for (int k = 0; k < 100000; k++) {
		double ak = a + k + 1;
		double lt = (a + k) * lx - x - lgamma(ak);
		double term = exp(lt) * (lx - or_digamma(ak));
		sum += term;
		if (k > x && fabs(term) < 1e-18 * fabs(sum)) break;
	}
	return sum;
}

/* ------------------------------------------------------------------ */
/* QUADPACK                                                             */
/* ------------------------------------------------------------------ */

static const double xgk21[11] = {
	0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
	0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
	0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
	0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
	0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
	0.000000000000000000000000000000000};
static const double wgk21[11] = {
	0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
	0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
	0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
	0.123491976262065851077600525452038, 0.134709217311473325928054001771707,
	0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
	0.149445554002916905664936468389821};
static const double wg10[5] = {
	0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
	0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
	0.295524224714752870173892994651338};

static const double xgk15[8] = {
	0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
	0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
	0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
	0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double wgk15[8] = {
	0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
	0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
	0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
	0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double wg7i[8] = {
	0.0, 0.129484966168869693270611432679082,
	0.0, 0.279705391489276667901467771423780,
	0.0, 0.381830050505118944950369775488975,
	0.0, 0.417959183673469387755102040816327};

const double *or_quadpack_nodes(int which)
{
	switch (which) {
	case 0: return xgk21;
	case 1: return wgk21;
	case 2: return wg10;
	case 3: return xgk15;
	case 4: return wgk15;
	case 5: return wg7i;
	default: return 0;
	}
}

static void qk_finish(double resk, double resg, double hlgth, double *resabs,
                      double *resasc, double *abserr)
{
	const double epmach = DBL_EPSILON, uflow = DBL_MIN;
	*abserr = fabs((resk - resg) * hlgth);
	if (*resasc != 0 && *abserr != 0) *abserr = *resasc * fmin(1.0, pow(200 * *abserr / *resasc, 1.5));
	if (*resabs > uflow / (50 * epmach)) *abserr = fmax(epmach * 50 * *resabs, *abserr);
}

static void dqk21(or_integrand f, void *ctx, double a, double b, double *result,
                  double *abserr, double *resabs, double *resasc)
{
	double fv1[10], fv2[10];
	double centr = 0.5 * (a + b), hlgth = 0.5 * (b - a), dhlgth = fabs(hlgth);
	double resg = 0, fc = f(centr, ctx);
	double resk = wgk21[10] * fc;
	*resabs = fabs(resk);
	for (int j = 0; j < 5; j++) {
		int jtw = 2 * j + 1;
		double absc = hlgth * xgk21[jtw];
		double f1 = f(centr - absc, ctx), f2 = f(centr + absc, ctx);
		fv1[jtw] = f1; fv2[jtw] = f2;
		double fsum = f1 + f2;
		resg += wg10[j] * fsum;
		resk += wgk21[jtw] * fsum;
		*resabs += wgk21[jtw] * (fabs(f1) + fabs(f2));
	}
	for (int j = 0; j < 5; j++) {
		int jtwm1 = 2 * j;
		double absc = hlgth * xgk21[jtwm1];
		double f1 = f(centr - absc, ctx), f2 = f(centr + absc, ctx);
		fv1[jtwm1] = f1; fv2[jtwm1] = f2;
		double fsum = f1 + f2;
		resk += wgk21[jtwm1] * fsum;
		*resabs += wgk21[jtwm1] * (fabs(f1) + fabs(f2));
	}
	double reskh = resk * 0.5;
	*resasc = wgk21[10] * fabs(fc - reskh);
	for (int j = 0; j < 10; j++) *resasc += wgk21[j] * (fabs(fv1[j] - reskh) + fabs(fv2[j] - reskh));
	*result = resk * hlgth;
	*resabs *= dhlgth;
	*resasc *= dhlgth;
	qk_finish(resk, resg, hlgth, resabs, resasc, abserr);
}

static void dqk15i(or_integrand f, void *ctx, double boun, int inf, double a, double b,
                   double *result, double *abserr, double *resabs, double *resasc)
{
	double fv1[7], fv2[7];
	double dinf = (double)(inf < 1 ? inf : 1);
	double centr = 0.5 * (a + b), hlgth = 0.5 * (b - a);
	double tabsc1 = boun + dinf * (1 - centr) / centr;
	double fval1 = f(tabsc1, ctx);
	if (inf == 2) fval1 += f(-tabsc1, ctx);
	double fc = (fval1 / centr) / centr;
	double resg = wg7i[7] * fc, resk = wgk15[7] * fc;
	*resabs = fabs(resk);
	for (int j = 0; j < 7; j++) {
		double absc = hlgth * xgk15[j];
		double absc1 = centr - absc, absc2 = centr + absc;
		double t1 = boun + dinf * (1 - absc1) / absc1;
		double t2 = boun + dinf * (1 - absc2) / absc2;
		double f1 = f(t1, ctx), f2 = f(t2, ctx);
		if (inf == 2) { f1 += f(-t1, ctx); f2 += f(-t2, ctx); }
		f1 = (f1 / absc1) / absc1;
		f2 = (f2 / absc2) / absc2;
		fv1[j] = f1; fv2[j] = f2;
		double fsum = f1 + f2;
		resg += wg7i[j] * fsum;
		resk += wgk15[j] * fsum;
		*resabs += wgk15[j] * (fabs(f1) + fabs(f2));
	}
	double reskh = resk * 0.5;
	*resasc = wgk15[7] * fabs(fc - reskh);
	for (int j = 0; j < 7; j++) *resasc += wgk15[j] * (fabs(fv1[j] - reskh) + fabs(fv2[j] - reskh));
	*result = resk * hlgth;
	*resasc *= hlgth;
	*resabs *= hlgth;
	qk_finish(resk, resg, hlgth, resabs, resasc, abserr);
}

/* maintain the descending ordering of error estimates (1-based arrays) */
static void dqpsrt(int limit, int last, int *maxerr, double *ermax, const double *elist,
                   int *iord, int *nrmax)
{
	int i, j, k, isucc, jupbn, jbnd, ibeg, ido;
	double errmax, errmin;
	if (last <= 2) {
		iord[1] = 1;
		iord[2] = 2;
		goto done;
	}
	errmax = elist[*maxerr];
	if (*nrmax != 1) {
		ido = *nrmax - 1;
		for (i = 1; i <= ido; i++) {
			isucc = iord[*nrmax - 1];
			if (errmax <= elist[isucc]) break;
			iord[*nrmax] = isucc;
			(*nrmax)--;
		}
	}
	jupbn = last;
	if (last > (limit / 2 + 2)) jupbn = limit + 3 - last;
	errmin = elist[last];
	jbnd = jupbn - 1;
	ibeg = *nrmax + 1;
	if (ibeg <= jbnd) {
		for (i = ibeg; i <= jbnd; i++) {
			isucc = iord[i];
			if (errmax >= elist[isucc]) goto insert_min;
			iord[i - 1] = isucc;
		}
	}
	iord[jbnd] = *maxerr;
	iord[jupbn] = last;
	goto done;
insert_min:
	iord[i - 1] = *maxerr;
	k = jbnd;
	for (j = i; j <= jbnd; j++) {
		isucc = iord[k];
		if (errmin < elist[isucc]) {
			iord[k + 1] = last;
			goto done;
		}
		iord[k + 1] = isucc;
		k--;
	}
	iord[i] = last;
done:
	*maxerr = iord[*nrmax];
	*ermax = elist[*maxerr];
}

/* epsilon algorithm (1-based epstab of size >= 52+1, res3la size 3+1) */
static void dqelg(int *n, double *epstab, double *result, double *abserr, double *res3la, int *nres)
{
	const double epmach = DBL_EPSILON, oflow = DBL_MAX;
	int i, ib, ib2, ie, indx, k1, k2, k3, limexp, newelm, num;
	double delta1, delta2, delta3, e0, e1, e1abs, e2, e3, err1, err2, err3, error, res, ss, tol1, tol2, tol3, epsinf;
	(*nres)++;
	*abserr = oflow;
	*result = epstab[*n];
	if (*n < 3) goto out;
	limexp = 50;
	epstab[*n + 2] = epstab[*n];
	newelm = (*n - 1) / 2;
	epstab[*n] = oflow;
	num = *n;
	k1 = *n;
	for (i = 1; i <= newelm; i++) {
		k2 = k1 - 1;
		k3 = k1 - 2;
		res = epstab[k1 + 2];
		e0 = epstab[k3];
		e1 = epstab[k2];
		e2 = res;
		e1abs = fabs(e1);
		delta2 = e2 - e1;
		err2 = fabs(delta2);
		tol2 = fmax(fabs(e2), e1abs) * epmach;
		delta3 = e1 - e0;
		err3 = fabs(delta3);
		tol3 = fmax(e1abs, fabs(e0)) * epmach;
		if (!(err2 > tol2 || err3 > tol3)) {
			*result = res;
			*abserr = err2 + err3;
			goto out;
		}
		e3 = epstab[k1];
		epstab[k1] = e1;
		delta1 = e1 - e3;
		err1 = fabs(delta1);
		tol1 = fmax(e1abs, fabs(e3)) * epmach;
		if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) {
			*n = i + i - 1;
			goto shift;
		}
		ss = 1 / delta1 + 1 / delta2 - 1 / delta3;
		epsinf = fabs(ss * e1);
		if (!(epsinf > 1e-4)) {
			*n = i + i - 1;
			goto shift;
		}
		res = e1 + 1 / ss;
		epstab[k1] = res;
		k1 -= 2;
		error = err2 + fabs(res - e2) + err3;
		if (error > *abserr) continue;
		*abserr = error;
		*result = res;
	}
shift:
	if (*n == limexp) *n = 2 * (limexp / 2) - 1;
	ib = ((num / 2) * 2 == num) ? 2 : 1;
	ie = newelm + 1;
	for (i = 1; i <= ie; i++) {
		ib2 = ib + 2;
		epstab[ib] = epstab[ib2];
		ib = ib2;
	}
	if (num != *n) {
		indx = num - *n + 1;
		for (i = 1; i <= *n; i++) {
			epstab[i] = epstab[indx];
			indx++;
		}
	}
	if (*nres < 4) {
		res3la[*nres] = *result;
		*abserr = oflow;
	}
	else {
		*abserr = fabs(*result - res3la[3]) + fabs(*result - res3la[2]) + fabs(*result - res3la[1]);
		res3la[1] = res3la[2];
		res3la[2] = res3la[3];
		res3la[3] = *result;
	}
out:
	*abserr = fmax(*abserr, 5 * epmach * fabs(*result));
}

#define QP_LIMIT_MAX 100

/* Shared driver for dqagse (finite) and dqagie (semi-infinite, inf = 1). */
static double qag_driver(or_integrand f, void *ctx, double a, double b, int infinite,
                         double epsabs, double epsrel, int limit, or_quad_info *info)
{
	const double epmach = DBL_EPSILON, uflow = DBL_MIN, oflow = DBL_MAX;
	double alist[QP_LIMIT_MAX + 2], blist[QP_LIMIT_MAX + 2], rlist[QP_LIMIT_MAX + 2], elist[QP_LIMIT_MAX + 2];
	int iord[QP_LIMIT_MAX + 2];
	double rlist2[55], res3la[4];
	double result = 0, abserr = 0, defabs, resabs, area, area1, area2, area12, a1, a2, b1, b2;
	double defab1, defab2, dres, erlarg = 0, erlast, errbnd, errmax, error1, error2, erro12, errsum;
	double ertest = 0, correc = 0, small = 0, reseps, abseps;
	int ier = 0, ierro = 0, iroff1 = 0, iroff2 = 0, iroff3 = 0, k, ksgn, ktmin = 0, last = 0;
	int maxerr, nres = 0, nrmax, numrl2, extrap = 0, noext = 0, id, jupbnd;
	double boun = a;
	int inf = 1;

	if (limit > QP_LIMIT_MAX) limit = QP_LIMIT_MAX;
	alist[1] = infinite ? 0.0 : a;
	blist[1] = infinite ? 1.0 : b;
	rlist[1] = 0;
	elist[1] = 0;
	iord[1] = 0;
	if (epsabs <= 0 && epsrel < fmax(50 * epmach, 0.5e-28)) {
		ier = 6;
		goto finish;
	}
	if (infinite) dqk15i(f, ctx, boun, inf, 0.0, 1.0, &result, &abserr, &defabs, &resabs);
	else dqk21(f, ctx, a, b, &result, &abserr, &defabs, &resabs);

	dres = fabs(result);
	errbnd = fmax(epsabs, epsrel * dres);
	last = 1;
	rlist[1] = result;
	elist[1] = abserr;
	iord[1] = 1;
	if (abserr <= 100 * epmach * defabs && abserr > errbnd) ier = 2;
	if (limit == 1) ier = 1;
	if (ier != 0 || (abserr <= errbnd && abserr != resabs) || abserr == 0) goto finish;

	rlist2[1] = result;
	errmax = abserr;
	maxerr = 1;
	area = result;
	errsum = abserr;
	abserr = oflow;
	nrmax = 1;
	numrl2 = 2;
	ksgn = -1;
	if (dres >= (1 - 50 * epmach) * defabs) ksgn = 1;

	for (last = 2; last <= limit; last++) {
		a1 = alist[maxerr];
		b1 = 0.5 * (alist[maxerr] + blist[maxerr]);
		a2 = b1;
		b2 = blist[maxerr];
		erlast = errmax;
		if (infinite) {
			dqk15i(f, ctx, boun, inf, a1, b1, &area1, &error1, &resabs, &defab1);
			dqk15i(f, ctx, boun, inf, a2, b2, &area2, &error2, &resabs, &defab2);
		}
		else {
			dqk21(f, ctx, a1, b1, &area1, &error1, &resabs, &defab1);
			dqk21(f, ctx, a2, b2, &area2, &error2, &resabs, &defab2);
		}
		area12 = area1 + area2;
		erro12 = error1 + error2;
		errsum = errsum + erro12 - errmax;
		area = area + area12 - rlist[maxerr];
		if (!(defab1 == error1 || defab2 == error2)) {
			if (!(fabs(rlist[maxerr] - area12) > 1e-5 * fabs(area12) || erro12 < 0.99 * errmax)) {
				if (extrap) iroff2++;
				else iroff1++;
			}
			if (last > 10 && erro12 > errmax) iroff3++;
		}
		rlist[maxerr] = area1;
		rlist[last] = area2;
		errbnd = fmax(epsabs, epsrel * fabs(area));
		if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
		if (iroff2 >= 5) ierro = 3;
		if (last == limit) ier = 1;
		if (fmax(fabs(a1), fabs(b2)) <= (1 + 100 * epmach) * (fabs(a2) + 1000 * uflow)) ier = 4;
		if (error2 > error1) {
			alist[maxerr] = a2;
			alist[last] = a1;
			blist[last] = b1;
			rlist[maxerr] = area2;
			rlist[last] = area1;
			elist[maxerr] = error2;
			elist[last] = error1;
		}
		else {
			alist[last] = a2;
			blist[maxerr] = b1;
			blist[last] = b2;
			elist[maxerr] = error1;
			elist[last] = error2;
		}
		dqpsrt(limit, last, &maxerr, &errmax, elist, iord, &nrmax);
		if (errsum <= errbnd) goto sum_rlist;
		if (ier != 0) goto set_result;
		if (last == 2) {
			small = infinite ? 0.375 : fabs(b - a) * 0.375;
			erlarg = errsum;
			ertest = errbnd;
			rlist2[2] = area;
			continue;
		}
		if (noext) continue;
		erlarg -= erlast;
		if (fabs(b1 - a1) > small) erlarg += erro12;
		if (!extrap) {
			if (fabs(blist[maxerr] - alist[maxerr]) > small) continue;
			extrap = 1;
			nrmax = 2;
		}
		if (ierro != 3 && erlarg > ertest) {
			int jump = 0;
			id = nrmax;
			jupbnd = last;
			if (last > (2 + limit / 2)) jupbnd = limit + 3 - last;
			for (k = id; k <= jupbnd; k++) {
				maxerr = iord[nrmax];
				errmax = elist[maxerr];
				if (fabs(blist[maxerr] - alist[maxerr]) > small) { jump = 1; break; }
				nrmax++;
			}
			if (jump) continue;
		}
		numrl2++;
		rlist2[numrl2] = area;
		dqelg(&numrl2, rlist2, &reseps, &abseps, res3la, &nres);
		ktmin++;
		if (ktmin > 5 && abserr < 1e-3 * errsum) ier = 5;
		if (abseps < abserr) {
			ktmin = 0;
			abserr = abseps;
			result = reseps;
			correc = erlarg;
			ertest = fmax(epsabs, epsrel * fabs(reseps));
			if (abserr <= ertest) break;
		}
		if (numrl2 == 1) noext = 1;
		if (ier == 5) break;
		maxerr = iord[1];
		errmax = elist[maxerr];
		nrmax = 1;
		extrap = 0;
		small *= 0.5;
		erlarg = errsum;
	}
	if (last > limit) last = limit;

set_result:
	if (abserr == oflow) goto sum_rlist;
	if (ier + ierro != 0) {
		if (ierro == 3) abserr += correc;
		if (ier == 0) ier = 3;
		if (result != 0 && area != 0) {
			if (abserr / fabs(result) > errsum / fabs(area)) goto sum_rlist;
		}
		else if (abserr > errsum) goto sum_rlist;
		else if (area == 0) goto final;
	}
	if (ksgn == -1 && fmax(fabs(result), fabs(area)) <= defabs * 0.01) goto final;
	if (0.01 > (result / area) || (result / area) > 100 || errsum > fabs(area)) ier = 6;
	goto final;

sum_rlist:
	result = 0;
	for (k = 1; k <= last; k++) result += rlist[k];
	abserr = errsum;
final:
	if (ier > 2) ier--;
finish:
	if (info) {
		info->abserr = abserr;
		info->ier = ier;
		info->last = last;
		info->neval = infinite ? 30 * last - 15 : 42 * last - 21;
	}
	return result;
}

double or_integrate_qags(or_integrand f, void *ctx, double a, double b, or_quad_info *info)
{
	return qag_driver(f, ctx, a, b, 0, 1e-4, 1e-4, 100, info);
}

double or_integrate_qagi(or_integrand f, void *ctx, double a, or_quad_info *info)
{
	return qag_driver(f, ctx, a, 0, 1, 1e-4, 1e-4, 100, info);
}

double or_integrate_qags_tol(or_integrand f, void *ctx, double a, double b, double epsabs,
                             double epsrel, int limit, or_quad_info *info)
{
	return qag_driver(f, ctx, a, b, 0, epsabs, epsrel, limit, info);
}

double or_integrate_qagi_tol(or_integrand f, void *ctx, double a, double epsabs, double epsrel,
                             int limit, or_quad_info *info)
{
	return qag_driver(f, ctx, a, 0, 1, epsabs, epsrel, limit, info);
}
