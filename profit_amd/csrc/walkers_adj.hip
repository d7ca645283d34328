// Adjoint of the rendering: the re-walk of the recursion trees
// (k_subsample<true>, one kernel for every family, which measured faster on
// the 400x400 bulge+disk batch than one kernel per family: 3.81 against
// 4.16 ms) and the adjoint of the directly evaluated pixels (k_adj_render,
// one kernel per radial family and sum precision, FP64 sums only for the
// steep-core families Sersic and core-Sersic); see kernel_common.h.
#include "kernel_common.h"

namespace profit_amd {

template <int TYPE, int PREC>
static int adj_render_launch(const GalaxyDev *gals, const InstDev *insts, const int *grp_off, int ninst, int nblocks,
                             int grp, const Comp *comps, double *grad_acc, hipStream_t st)
{
	k_adj_render<TYPE, PREC != 0><<<nblocks, 256, 0, st>>>(gals, insts, grp_off, ninst, grp, comps, grad_acc);
	return (int)hipGetLastError();
}

#define PA_ADJ_CASES(CALL)                                            \
	if (prec == 0) {                                                  \
		switch (type) {                                               \
		case SERSIC: return CALL(SERSIC, 0);                          \
		case MOFFAT: return CALL(MOFFAT, 0);                          \
		case FERRER: return CALL(FERRER, 0);                          \
		case KING: return CALL(KING, 0);                              \
		case CORESERSIC: return CALL(CORESERSIC, 0);                  \
		case BROKENEXP: return CALL(BROKENEXP, 0);                    \
		default: return (int)hipErrorInvalidValue;                    \
		}                                                             \
	}                                                                 \
	if (type == SERSIC) return CALL(SERSIC, 1);                       \
	if (type == CORESERSIC) return CALL(CORESERSIC, 1);               \
	return (int)hipErrorInvalidValue;

int launch_subsample_adj(int nblocks, const GalaxyDev *gals, const InstDev *insts, const Comp *comps, WorkQueue wq,
                         double *grad_acc, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	k_subsample<true, -1, 2><<<nblocks, 256, 0, st>>>(gals, insts, comps, wq, kAdjCounter, 0, grad_acc);
	return (int)hipGetLastError();
}

int launch_adj_render(int type, int prec, const GalaxyDev *gals, const InstDev *insts, const int *grp_off, int ninst,
                      int nblocks, int grp, const Comp *comps, double *grad_acc, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	if (grp < 1) return (int)hipErrorInvalidValue;
#define PA_RENDER(T, P) adj_render_launch<T, P>(gals, insts, grp_off, ninst, nblocks, grp, comps, grad_acc, st)
	PA_ADJ_CASES(PA_RENDER)
#undef PA_RENDER
}

#undef PA_ADJ_CASES

}  // namespace profit_amd
