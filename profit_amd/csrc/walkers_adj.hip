// Adjoint of the rendering: the sub-sampled pixels from the Jacobians of the
// forward walk (k_adj_queue, one kernel per profile family) and the directly
// evaluated pixels (k_adj_render, one kernel per radial family and sum
// precision, FP64 sums only for the steep-core families Sersic and
// core-Sersic); see kernel_common.h.
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

template <int TYPE>
static int adj_queue_launch(int nblocks, const InstDev *insts, const Comp *comps, WorkQueue wq, double *grad_acc,
                            hipStream_t st)
{
	k_adj_queue<TYPE><<<nblocks, 256, 0, st>>>(insts, comps, wq, grad_acc);
	return (int)hipGetLastError();
}

int launch_adj_queue(int type, int nblocks, const InstDev *insts, const Comp *comps, WorkQueue wq, double *grad_acc,
                     hipStream_t st)
{
	if (nblocks <= 0) return 0;
	switch (type) {
	case SERSIC: return adj_queue_launch<SERSIC>(nblocks, insts, comps, wq, grad_acc, st);
	case MOFFAT: return adj_queue_launch<MOFFAT>(nblocks, insts, comps, wq, grad_acc, st);
	case FERRER: return adj_queue_launch<FERRER>(nblocks, insts, comps, wq, grad_acc, st);
	case KING: return adj_queue_launch<KING>(nblocks, insts, comps, wq, grad_acc, st);
	case CORESERSIC: return adj_queue_launch<CORESERSIC>(nblocks, insts, comps, wq, grad_acc, st);
	case BROKENEXP: return adj_queue_launch<BROKENEXP>(nblocks, insts, comps, wq, grad_acc, st);
	default: return (int)hipErrorInvalidValue;
	}
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
