// The set-up with its QUADPACK integrals spread over a wavefront's lanes
// (setup_component<true>, lum_gradient<true>: each lane evaluates one Kronrod
// point, every lane forms the sums; a Sersic component's three gamma quantiles
// on three lanes) against the serial set-up, on the host
// stand-in of the HIP thread model: one block of 64 threads.  Prints one line
// per case: <profile> <case> <bitwise-equal set-up> <bitwise-equal gradients>
#include <cstdio>
#include <cstring>

#include "profile_math.h"

using namespace profit_amd;

static Comp g_wave;
static double g_grad[5];

static void check(const char *name, int k, const ProfileParams &p)
{
	// both results zero-filled first, so that fields a set-up leaves alone
	// (and padding) compare equal
	sf::QuadState qs;
	Comp ref;
	std::memset(&ref, 0, sizeof(ref));
	if (setup_component(p, 30.0, 1.0, 1.0, 1.0, 1.0, 10, 10, &qs, &ref, false) != SE_OK) {
		std::printf("%s %d setup-error\n", name, k);
		return;
	}
	double gref[5] = {0, 0, 0, 0, 0};
	for (int w = 0; w < n_lum_integrals(p.type); ++w) gref[w] = lum_gradient(ref, w, &qs);
	hip_cpu::launch(1, 64, [&] {
		sf::QuadState q;
		Comp out;
		std::memset(&out, 0, sizeof(out));
		setup_component<true>(p, 30.0, 1.0, 1.0, 1.0, 1.0, 10, 10, &q, &out, false);
		double g[5] = {0, 0, 0, 0, 0};
		for (int w = 0; w < n_lum_integrals(p.type); ++w) g[w] = lum_gradient<true>(out, w, &q);
		if (threadIdx.x == 0) {
			std::memcpy(&g_wave, &out, sizeof(Comp));
			for (int w = 0; w < 5; ++w) g_grad[w] = g[w];
		}
	});
	const bool same = std::memcmp(&ref, &g_wave, sizeof(Comp)) == 0;
	const bool gsame = std::memcmp(gref, g_grad, sizeof(gref)) == 0;
	std::printf("%s %d %d %d %.17g %.17g\n", name, k, same ? 1 : 0, gsame ? 1 : 0, ref.lumtot, g_wave.lumtot);
}

int main()
{
	// Sersic: b_n and the adjust rule's two quantiles, one lane each (the
	// fixed rscale_max and the unadjusted component leave lanes 1-2 idle)
	for (int k = 0; k < 4; ++k) {
		ProfileParams p;
		default_params(SERSIC, &p);
		p.xcen = 40; p.ycen = 40; p.mag = 16; p.axrat = 0.7; p.ang = 30;
		const double nser[4] = {0.5, 4, 8, 2.5};
		p.shape[0] = 3 + k;          // re
		p.shape[1] = nser[k];        // nser
		p.adjust = k != 3;
		p.rscale_max = k == 2 ? 5 : 0;
		check("sersic", k, p);
	}
	for (int k = 0; k < 3; ++k) {
		ProfileParams p;
		default_params(KING, &p);
		p.xcen = 40; p.ycen = 40; p.mag = 16; p.axrat = 0.8; p.ang = 20;
		p.shape[0] = 2 + k;          // rc
		p.shape[1] = 15 + 5 * k;     // rt
		p.shape[2] = 1.5 + 0.5 * k;  // a
		check("king", k, p);
	}
	for (int k = 0; k < 3; ++k) {
		ProfileParams p;
		default_params(CORESERSIC, &p);
		p.xcen = 40; p.ycen = 40; p.mag = 16; p.axrat = 0.7;
		p.shape[0] = 2 + k;          // re
		p.shape[1] = 8 + 2 * k;      // rb
		p.shape[2] = 2 + k;          // nser
		p.shape[3] = 3;              // a
		p.shape[4] = 0.5 * k;        // b
		check("coresersic", k, p);
	}
	for (int k = 0; k < 3; ++k) {
		ProfileParams p;
		default_params(BROKENEXP, &p);
		p.xcen = 40; p.ycen = 40; p.mag = 16; p.axrat = 0.6;
		p.shape[0] = 3 + k;          // h1
		p.shape[1] = 6 + 2 * k;      // h2
		p.shape[2] = 10 + 3 * k;     // rb
		p.shape[3] = 1 + k;          // a
		check("brokenexp", k, p);
	}
	return 0;
}
