// Forward recursion walker (libprofit's subsample_pixel, radial.cpp:54-134):
// a root pass and the wavefront walker, one kernel each per radial profile
// family; see kernel_common.h.
#include "kernel_common.h"

namespace profit_amd {

// The root pass (one lane per light entry) and then the wavefront walker
// over the heavy list and the entries the root pass left to it.
template <int TYPE>
static int fwd_launch(int nblocks, const GalaxyDev *gals, const InstDev *insts, const Comp *comps, WorkQueue wq,
                      int split, bool jac, hipStream_t st)
{
	if (jac) {
		k_subsample_root<TYPE, true><<<nblocks, 256, 0, st>>>(gals, insts, comps, wq, split);
		k_subsample<TYPE, true><<<nblocks, 256, 0, st>>>(gals, insts, comps, wq, fwd_counter(TYPE), split);
	}
	else {
		k_subsample_root<TYPE, false><<<nblocks, 256, 0, st>>>(gals, insts, comps, wq, split);
		k_subsample<TYPE, false><<<nblocks, 256, 0, st>>>(gals, insts, comps, wq, fwd_counter(TYPE), split);
	}
	return (int)hipGetLastError();
}

int launch_subsample_fwd(int type, int nblocks, const GalaxyDev *gals, const InstDev *insts, const Comp *comps,
                         WorkQueue wq, int split, bool jac, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	switch (type) {
	case SERSIC: return fwd_launch<SERSIC>(nblocks, gals, insts, comps, wq, split, jac, st);
	case MOFFAT: return fwd_launch<MOFFAT>(nblocks, gals, insts, comps, wq, split, jac, st);
	case FERRER: return fwd_launch<FERRER>(nblocks, gals, insts, comps, wq, split, jac, st);
	case KING: return fwd_launch<KING>(nblocks, gals, insts, comps, wq, split, jac, st);
	case CORESERSIC: return fwd_launch<CORESERSIC>(nblocks, gals, insts, comps, wq, split, jac, st);
	case BROKENEXP: return fwd_launch<BROKENEXP>(nblocks, gals, insts, comps, wq, split, jac, st);
	default: return (int)hipErrorInvalidValue;
	}
}

}  // namespace profit_amd
