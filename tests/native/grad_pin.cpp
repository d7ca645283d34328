// Host-side check of the per-sample profile derivatives and of the
// d ln(flux scale) terms that the adjoint kernels use, against central finite
// differences of the same profile functions.  Built host-only from the
// product's own headers; prints one line per checked quantity:
//   <profile> <point> <quantity> <analytic> <finite difference>
#include <cmath>
#include <cstdio>

#include "profile_math.h"

using namespace profit_amd;

static Comp setup(const ProfileParams &p)
{
	sf::QuadState qs;
	Comp c;
	if (setup_component(p, 30.0, 1.0, 1.0, 1.0, 1.0, 0, 0, &qs, &c) != SE_OK) std::printf("setup-error\n");
	return c;
}

static double f_at(const ProfileParams &p, double x, double y)
{
	const Comp c = setup(p);
	return eval_at<double>(c, x, y);
}

static void check(const char *name, const ProfileParams &p0, int nshape, const double pts[][2], int npts)
{
	const Comp c0 = setup(p0);
	for (int k = 0; k < npts; ++k) {
		const double x = pts[k][0], y = pts[k][1];
		double o[9];
		sample_grad<double>(c0, x, y, o);
		std::printf("%s %d value %.17g %.17g\n", name, k, o[0], eval_at<double>(c0, x, y));
		const double hx = 1e-6 * (1 + std::fabs(x)), hy = 1e-6 * (1 + std::fabs(y));
		std::printf("%s %d dxp %.17g %.17g\n", name, k, o[1],
		            (eval_at<double>(c0, x + hx, y) - eval_at<double>(c0, x - hx, y)) / (2 * hx));
		std::printf("%s %d dyp %.17g %.17g\n", name, k, o[2],
		            (eval_at<double>(c0, x, y + hy) - eval_at<double>(c0, x, y - hy)) / (2 * hy));
		{
			ProfileParams a = p0, b = p0;
			const double h = 1e-6;
			a.box += h;
			b.box -= h;
			std::printf("%s %d dbox %.17g %.17g\n", name, k, o[3], (f_at(a, x, y) - f_at(b, x, y)) / (2 * h));
		}
		for (int s = 0; s < nshape; ++s) {
			ProfileParams a = p0, b = p0;
			const double h = 1e-6 * std::fabs(p0.shape[s]);
			a.shape[s] += h;
			b.shape[s] -= h;
			std::printf("%s %d ds%d %.17g %.17g\n", name, k, s, o[4 + s], (f_at(a, x, y) - f_at(b, x, y)) / (2 * h));
		}
	}
	// d ln(flux scale) / d (mag, axrat, box, shape)
	auto lnfs = [](const ProfileParams &p) { return std::log(setup(p).ie); };
	{
		ProfileParams a = p0, b = p0;
		a.mag += 1e-5;
		b.mag -= 1e-5;
		std::printf("%s - lnfs_mag %.17g %.17g\n", name, c0.dlnfs[G_MAG], (lnfs(a) - lnfs(b)) / 2e-5);
		a = p0;
		b = p0;
		a.axrat += 1e-6;
		b.axrat -= 1e-6;
		std::printf("%s - lnfs_axrat %.17g %.17g\n", name, c0.dlnfs[G_AXRAT], (lnfs(a) - lnfs(b)) / 2e-6);
		a = p0;
		b = p0;
		a.box += 1e-5;
		b.box -= 1e-5;
		std::printf("%s - lnfs_box %.17g %.17g\n", name, c0.dlnfs[G_BOX], (lnfs(a) - lnfs(b)) / 2e-5);
	}
	for (int s = 0; s < nshape; ++s) {
		ProfileParams a = p0, b = p0;
		const double h = 1e-4 * std::fabs(p0.shape[s]);
		a.shape[s] += h;
		b.shape[s] -= h;
		std::printf("%s - lnfs_s%d %.17g %.17g\n", name, s, c0.dlnfs[G_S0 + s], (lnfs(a) - lnfs(b)) / (2 * h));
	}
}

int main()
{
	const double pts[][2] = {{1.3, 0.4}, {-2.7, 3.1}, {6.2, -0.9}, {0.35, -0.2}, {11.0, 7.5}};
	const int np = 5;
	ProfileParams p;
	default_params(SERSIC, &p);
	p.shape[0] = 4.3;
	p.shape[1] = 2.7;
	p.axrat = 0.7;
	p.box = 0.3;
	check("sersic", p, 2, pts, np);
	default_params(MOFFAT, &p);
	p.shape[0] = 3.1;
	p.shape[1] = 2.6;
	p.axrat = 0.8;
	p.box = -0.2;
	check("moffat", p, 2, pts, np);
	default_params(FERRER, &p);
	p.shape[0] = 14.0;
	p.shape[1] = 1.3;
	p.shape[2] = 0.6;
	p.axrat = 0.9;
	p.box = 0.2;
	check("ferrer", p, 3, pts, np);
	default_params(KING, &p);
	p.shape[0] = 2.2;
	p.shape[1] = 19.0;
	p.shape[2] = 2.3;
	p.axrat = 0.6;
	p.box = 0.1;
	check("king", p, 3, pts, np);
	default_params(CORESERSIC, &p);
	p.shape[0] = 5.0;
	p.shape[1] = 1.2;
	p.shape[2] = 2.5;
	p.shape[3] = 2.0;
	p.shape[4] = 0.5;
	p.axrat = 0.75;
	p.box = 0.15;
	check("coresersic", p, 5, pts, np);
	default_params(BROKENEXP, &p);
	p.shape[0] = 3.0;
	p.shape[1] = 1.5;
	p.shape[2] = 4.0;
	p.shape[3] = 1.7;
	p.axrat = 0.65;
	p.box = -0.1;
	check("brokenexp", p, 4, pts, np);
	return 0;
}
