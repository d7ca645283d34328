// Host build of the product's gamma functions (special.h): prints
//   <p> <a> <qgamma(p, a)> <pgamma(qgamma(p, a), a)>
// over the quantiles the Sersic set-up asks for (b_n at p = 0.5, the switch
// radius at 1 - n^2/2000, the maximum radius at 0.9999), n = a / 2.
#include <cstdio>

#include "special.h"

using namespace profit_amd;

int main()
{
	for (double n = 0.3; n < 20.0; n += 0.0731) {
		const double a = 2 * n;
		const double ps[3] = {0.5, 1.0 - n * n / 2e3, 0.9999};
		for (double p : ps) {
			if (!(p > 0 && p < 1)) continue;
			const double x = sf::qgamma(p, a);
			std::printf("%.17g %.17g %.17g %.17g\n", p, a, x, sf::pgamma(x, a));
		}
	}
	return 0;
}
