// Device code shared by the kernel translation units: block ownership,
// wave reductions, libprofit's cell positions and refinement test, and the
// templates of the recursion walker and of the adjoint render.
// walkers_fwd.hip compiles the forward walker once per profile family and
// walkers_adj.hip the adjoint walker once per family and the adjoint render
// once per family and sum precision: a kernel's register allocation is the
// maximum over the code it holds, so one family per kernel lifts the walkers
// from two to three or four waves per SIMD and the adjoint render from two
// to seven or eight.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace profit_amd {

// The instance whose block range [off[i], off[i+1]) holds block b (off: the
// n + 1 prefix sums of the instances' block counts).
__device__ inline int owner_of(const int *__restrict__ off, int n, int b)
{
	// a first guess from the mean count per instance (exact, up to the
	// rounding that the neighbour checks absorb, when every instance has the
	// same count), then the binary search if it misses
	int gs = min(n - 1, (int)((float)b * ((float)n / (float)off[n])));
	if (gs > 0 && off[gs] > b) --gs;
	else if (gs + 1 < n && off[gs + 1] <= b) ++gs;
	if (off[gs] <= b && off[gs + 1] > b) return gs;
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (off[mid] <= b) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}

// 1 / res^2 and 1 / res for the recursion's resolutions 1 .. 16 (setup
// rejects higher ones), correctly rounded like the divisions they replace
static __constant__ double kInvCells[17] = {0.0,        1.0 / 1,   1.0 / 4,   1.0 / 9,   1.0 / 16,  1.0 / 25,
                                            1.0 / 36,   1.0 / 49,  1.0 / 64,  1.0 / 81,  1.0 / 100, 1.0 / 121,
                                            1.0 / 144,  1.0 / 169, 1.0 / 196, 1.0 / 225, 1.0 / 256};
static __constant__ float kInvRes[17] = {0.f,       1.f / 1,  1.f / 2,  1.f / 3,  1.f / 4,  1.f / 5,
                                         1.f / 6,   1.f / 7,  1.f / 8,  1.f / 9,  1.f / 10, 1.f / 11,
                                         1.f / 12,  1.f / 13, 1.f / 14, 1.f / 15, 1.f / 16};

// sersic.cpp:331-347: pixels within one image pixel of the centre of an
// nser > 1 Sersic profile are sub-sampled at resolution 8, depth 10
__device__ __forceinline__ bool centre_pixel(const Comp &c, double x, double y)
{
	return c.centre_rule && fabs(x - c.xc) < c.img_sx && fabs(y - c.yc) < c.img_sy;
}

struct WalkLevel {
	double x0, x1, y0, y1, weight;
	unsigned long long mask[4];
};

__device__ __forceinline__ double cell_centre(double lo, double half, int idx)
{
	// the same sequence of additions as libprofit's loops (radial.cpp:86-106)
	double v = lo;
	for (int t = 0; t < idx; ++t) {
		v += half;
		v += half;
	}
	return v + half;
}

// d(f)/d slot of one sample in precision T, chain rule to the component
// slots, accumulated with weight wgt (FP64) into gs; returns f.  TYPE >= 0
// compiles the one profile family (see eval_at).
template <class T, int TYPE = -1>
__device__ __forceinline__ double sample_accumulate(const Comp &c, double xp, double yp, double wgt, double *gs)
{
	T o[9];
	const T fxp = (T)xp, fyp = (T)yp;
	sample_grad<T, TYPE>(c, fxp, fyp, o);
	// the rotation's chain rule with the per-component sin/q, cos/q and 1/q
	const T q = (T)c.axrat, iq = (T)c.inv_axrat;
	const T ca = (T)c.cos_ang, sa = (T)c.sin_ang;
	const T sq = (T)c.sin_over, cq = (T)c.cos_over;
	const T fx = o[1], fy = o[2];
	gs[G_XCEN] += wgt * (double)(fx * -ca + fy * sq);
	gs[G_YCEN] += wgt * (double)(fx * -sa + fy * -cq);
	gs[G_ANG] += wgt * (double)((fx * (q * fyp) + fy * (-fxp * iq)) * (T(3.14159265358979323846) / T(180)));
	gs[G_AXRAT] += wgt * (double)(fy * (-fyp * iq));
	if (c.grad_box) gs[G_BOX] += wgt * (double)o[3];
#pragma unroll
	for (int s = 0; s < n_shape(TYPE); ++s) gs[G_S0 + s] += wgt * (double)o[4 + s];
	return (double)o[0];
}

// Where a King or Ferrer cell lies against its profile's cut-off (rt, rout),
// from the FP32 radii of the cell centre (x, y) and of the test point
// (ax, ay), which lies at least as far out: 0 when the centre lies beyond
// 1.01 x the cut (the value and the test are then exactly 0 in either
// precision, so the cell is a leaf of value 0); 1 when the centre lies in
// [0.97, 1.01] x the cut or the test point within 1% of it, where the value is
// a difference of nearly equal terms (King: A - C; Ferrer: 1 - t^(2-b)) whose
// FP32 relative error grows without bound and a point can fall on either side
// of the cut (the FP64 test decides, and the leaf value is FP64); -1
// otherwise, and for the other families.
template <int TYPE>
__device__ __forceinline__ int cut_zone(const Comp &c, float x, float y, float ax, float ay)
{
	if (TYPE >= 0 && TYPE != KING && TYPE != FERRER) return -1;
	const int t = TYPE >= 0 ? TYPE : c.type;
	if (t != KING && t != FERRER) return -1;
	const float cut = t == KING ? c.f.s[1] : c.f.rscale;
	const float rc = boxy_r<float>(c, x, y), rt = boxy_r<float>(c, ax, ay);
	if (rc > 1.01f * cut) return 0;
	const bool band_c = rc >= 0.97f * cut, band_t = rt >= 0.99f * cut && rt <= 1.01f * cut;
	return band_c || band_t ? 1 : -1;
}

// libprofit's refinement test |test - val| > acc * val (radial.cpp:92-100),
// evaluated in FP32 at FP32 cell positions: 1 refine, 0 leaf, 2 undecided.
// The decision is taken only when the FP32 margin lies outside a guard band
// of 1e-3 * val around the threshold; the FP32 evaluation errors (values and
// positions, a few parts in 1e7 of val) are far inside it, so a decision taken
// here equals the FP64 one.  King and Ferrer cells beyond their cut-off are
// leaves; those near it are undecided (cut_zone).  Undecided cells take
// refine_cell64 at libprofit's own FP64 cell position.
template <int TYPE = -1>
__device__ __forceinline__ int refine_cell32(const Comp &c, float xp, float yp, float dyp, float v32)
{
	const float ax = fabsf(xp), ay = fabsf(yp) + fabsf(dyp);
	const int zone = cut_zone<TYPE>(c, xp, yp, ax, ay);
	if (zone == 0) return 0;
	if (zone == 1) return 2;
	const float t32 = eval_fast<TYPE>(c, ax, ay);
	const float margin = fabsf(t32 - v32) - c.f.acc * v32;
	if (v32 > 1e-30f && isfinite(margin) && fabsf(margin) > 1e-3f * v32) return margin > 0.f ? 1 : 0;
	return 2;
}

// the FP64 refinement test; *v64 receives the FP64 value
template <int TYPE = -1>
__device__ __forceinline__ bool refine_cell64(const Comp &c, double xp, double yp, double dyp, double *v64)
{
	const double v = eval_at<double, TYPE>(c, xp, yp);
	const double t = eval_at<double, TYPE>(c, fabs(xp), fabs(yp) + fabs(dyp));
	*v64 = v;
	return fabs(t - v) > c.acc * v;
}

// Profile coordinates of cell (ci, cj) in FP32: the node's lower corner
// relative to the component centre (ox, oy, formed once per node in FP64 and
// rounded) plus the cell's offset.  Relative to the distance from the centre
// the position errs by a few parts in 1e8, which moves a profile value by
// |d ln f / d ln r| times that: far inside the refinement test's guard band.
__device__ __forceinline__ void cell_pos32(const Comp &c, float ox, float oy, float fxbin, float fybin, int ci, int cj,
                                           float *xp, float *yp)
{
	const float dx = ox + ((float)ci + 0.5f) * fxbin;
	const float dy = oy + ((float)cj + 0.5f) * fybin;
	*xp = dx * c.f.cos_ang + dy * c.f.sin_ang;
	*yp = -dx * c.f.sin_over + dy * c.f.cos_over;
}

// libprofit's FP64 cell position (its sequence of additions, radial.cpp:86-106)
// in profile coordinates
__device__ __forceinline__ void cell_pos64(const Comp &c, double nx0, double ny0, double hx, double hy, int ci, int cj,
                                           double *xp, double *yp)
{
	const double x = cell_centre(nx0, hx, ci), y = cell_centre(ny0, hy, cj);
	const double xr = x - c.xc, yr = y - c.yc;
	*xp = xr * c.cos_ang + yr * c.sin_ang;
	*yp = -xr * c.sin_over + yr * c.cos_over;
}

// d(f)/d slot of one sample in FP32, chain rule to the component slots,
// accumulated with weight wgt into FP32 sums gs; returns f.
template <int TYPE>
__device__ __forceinline__ float pixel_accumulate_f32(const Comp &c, float xp, float yp, float wgt, float *gs)
{
	// FP32 derivatives only: products and sums may fuse into FMAs
#pragma clang fp contract(fast)
	float o[9];
	sample_grad<float, TYPE>(c, xp, yp, o);
	// the rotation's chain rule with the per-component sin/q, cos/q and 1/q
	const float q = c.f.axrat, iq = c.f.inv_axrat;
	const float ca = c.f.cos_ang, sa = c.f.sin_ang;
	const float sq = c.f.sin_over, cq = c.f.cos_over;
	const float fx = o[1], fy = o[2];
	gs[G_XCEN] += wgt * (fx * -ca + fy * sq);
	gs[G_YCEN] += wgt * (fx * -sa + fy * -cq);
	gs[G_ANG] += wgt * ((fx * (q * yp) + fy * (-xp * iq)) * (3.14159265358979323846f / 180.f));
	gs[G_AXRAT] += wgt * (fy * (-yp * iq));
	if (c.grad_box) gs[G_BOX] += wgt * o[3];
#pragma unroll
	for (int s = 0; s < n_shape(TYPE); ++s) gs[G_S0 + s] += wgt * o[4 + s];
	return o[0];
}

// The same for a Sersic sample without box from the intermediates of its
// value (sersic_parts_f32): f, base = (r / re)^2, l2 = log2(base), rf =
// base^(1/(2n)), so that a leaf pays no second logarithm or exponential.
__device__ __forceinline__ void sersic_jac_f32(const Comp &c, float xp, float yp, float f, float base, float l2,
                                               float rf, float wgt, float *gs)
{
#pragma clang fp contract(fast)
	if (!(base > 0.f)) return;  // the centre: every derivative is 0
	const float bn = c.f.bn, in = c.f.inv_n;
	const float bfr = bn * f * rf * in;
	// df/dxp = k xp, df/dyp = k yp with k = -bn f rf / (n re^2 base)
	const float k = -bfr * c.f.inv_re2 * __builtin_amdgcn_rcpf(base);
	const float fx = k * xp, fy = k * yp;
	const float q = c.f.axrat, iq = c.f.inv_axrat;
	gs[G_XCEN] += wgt * (fx * -c.f.cos_ang + fy * c.f.sin_over);
	gs[G_YCEN] += wgt * (fx * -c.f.sin_ang + fy * -c.f.cos_over);
	gs[G_ANG] += wgt * ((fx * (q * yp) + fy * (-xp * iq)) * (3.14159265358979323846f / 180.f));
	gs[G_AXRAT] += wgt * (fy * (-yp * iq));
	gs[G_S0] += wgt * (bfr * c.f.inv_re);
	gs[G_S1] += wgt * (f * (-c.f.dbn * (rf - 1.f) + bn * rf * (l2 * float(kLn2)) * in * c.f.inv_2n));
}

// A node of the recursion: its FP64 bounds (libprofit's), the resolution,
// and the FP32 cell size, test offset and lower corner relative to the
// component centre that the FP32 cell positions use.
struct Node {
	double x0, x1, y0, y1;
	int res;
	float inv_res, xbin, ybin, dyp, ox, oy;
};

__device__ __forceinline__ Node make_node(const Comp &c, double x0, double x1, double y0, double y1, int res)
{
	Node n;
	n.x0 = x0;
	n.x1 = x1;
	n.y0 = y0;
	n.y1 = y1;
	n.res = res;
	n.inv_res = kInvRes[res];
	n.xbin = (float)(x1 - x0) * n.inv_res;
	n.ybin = (float)(y1 - y0) * n.inv_res;
	n.dyp = -n.xbin * c.f.sin_over + n.ybin * c.f.cos_over;
	n.ox = (float)(x0 - c.xc);
	n.oy = (float)(y0 - c.yc);
	return n;
}

// One cell of a node: its FP32 position, value and refinement test (in FP64
// at libprofit's own cell centre when the FP32 test is undecided); a leaf
// adds weight * value to sum and, with kJac, its weighted FP32 derivatives
// to gj.  `fuse`: a Sersic without box, whose derivatives reuse the value's
// logarithm and exponentials.  Returns whether the cell is to be refined.
template <int TYPE, bool kJac>
__device__ __forceinline__ bool cell_step(const Comp &c, const Node &n, int cell, bool recurse, bool fp64, bool fuse,
                                          double weight, double &sum, float *gj)
{
	// libprofit's FP64 cell size and test offset, for the cells that need an
	// FP64 position
	auto node64 = [&](double &hx, double &hy, double &dyp) {
		const double xbin = (n.x1 - n.x0) / n.res;
		const double ybin = (n.y1 - n.y0) / n.res;
		hx = xbin / 2.;
		hy = ybin / 2.;
		dyp = -xbin * c.sin_over + ybin * c.cos_over;
	};
	// cell -> (ci, cj) by an FP32 reciprocal: (cell + 0.5) / res lies at least
	// 0.5 / res from an integer, far beyond the product's rounding, so the
	// truncation is the exact quotient
	const int ci = (int)(((float)cell + 0.5f) * n.inv_res), cj = cell - ci * n.res;
	float xp, yp;
	cell_pos32(c, n.ox, n.oy, n.xbin, n.ybin, ci, cj, &xp, &yp);
	float v32, s_base = 0.f, s_l2 = 0.f, s_rf = 0.f;
	if (fuse) v32 = sersic_parts_f32(c, xp, yp, s_base, s_l2, s_rf);
	else v32 = eval_at<float, TYPE>(c, xp, yp);
	int dec = recurse ? refine_cell32<TYPE>(c, xp, yp, n.dyp, v32) : 0;
	double v64 = 0;
	bool have64 = false;
	if (dec == 2) {
		double hx, hy, dyp, xp64, yp64;
		node64(hx, hy, dyp);
		cell_pos64(c, n.x0, n.y0, hx, hy, ci, cj, &xp64, &yp64);
		dec = refine_cell64<TYPE>(c, xp64, yp64, dyp, &v64) ? 1 : 0;
		have64 = true;
	}
	if (dec == 1) return true;
	// leaf values in FP64 or the FP32 value already at hand
	if (!have64) {
		if (fp64) {
			double hx, hy, dyp, xp64, yp64;
			node64(hx, hy, dyp);
			cell_pos64(c, n.x0, n.y0, hx, hy, ci, cj, &xp64, &yp64);
			v64 = eval_at<double, TYPE>(c, xp64, yp64);
		}
		else {
			v64 = (double)v32;
		}
	}
	sum += v64 * weight;
	if (fuse) sersic_jac_f32(c, xp, yp, v32, s_base, s_l2, s_rf, (float)weight, gj);
	else if (kJac) pixel_accumulate_f32<TYPE>(c, xp, yp, (float)weight, gj);
	return false;
}

// Walks the recursion tree of one pixel of a TYPE profile.  Cell positions,
// values and refinement tests are FP32; libprofit's FP64 cell position is
// formed only for an undecided test and for FP64 leaf values (`fp64`).  The
// descent to a refined cell follows libprofit's FP64 node bounds, so the
// trees are its own.  Every lane returns its share of sum_s weight_s f(s) in
// `sum`; with kJac (the forward walk of a gradient evaluation) every leaf
// also adds its weighted FP32 derivatives to gj (per lane), for the pixel's
// Jacobian.
template <int TYPE, bool kJac>
__device__ __forceinline__ void walk_pixel(const Comp &c, WalkLevel *L, int lane, double x0, double x1, double y0,
                                           double y1, int res, int maxrec, bool fp64, double &sum, float *gj)
{
	// a Sersic Jacobian without box (or box derivative) from the value's
	// own logarithm and exponentials
	const bool fuse = kJac && TYPE == SERSIC && !c.boxy && !c.grad_box;
	const int ncell = res * res;
	const int nchunk = (ncell + 63) >> 6;
	// 1 / res^2 from a table (the quotient a division rounds to)
	const double inv_cells = kInvCells[res];
	int level = 0;
	double nx0 = x0, nx1 = x1, ny0 = y0, ny1 = y1;
	double weight = inv_cells;
	for (;;) {
		const bool recurse = res > 1 && level < maxrec;
		const Node n = make_node(c, nx0, nx1, ny0, ny1, res);
		// the cells in chunks of 64 (one per lane); lane 0 stores each chunk's
		// refine mask with the node, and `any` gathers them.  A node with no
		// cell to refine is not stored: the search below starts under it.
		unsigned long long any = 0ull;
#pragma unroll 1
		for (int k = 0; k < nchunk; ++k) {
			const int cell = k * 64 + lane;
			const bool fail = cell < ncell && cell_step<TYPE, kJac>(c, n, cell, recurse, fp64, fuse, weight, sum, gj);
			const unsigned long long mk = __ballot(fail);
			any |= mk;
			if (lane == 0) L[level].mask[k] = mk;
		}
		if (any != 0ull) {
			if (lane == 0) {
				L[level].x0 = nx0;
				L[level].x1 = nx1;
				L[level].y0 = ny0;
				L[level].y1 = ny1;
				L[level].weight = weight;
				for (int k = nchunk; k < 4; ++k) L[level].mask[k] = 0ull;
			}
		}
		else if (level == 0) {
			break;  // a leaf-only root: the tree is done
		}
		__builtin_amdgcn_wave_barrier();
		// next node: lowest pending cell of the deepest level that has one
		int lv = any != 0ull ? level : level - 1;
		int cell = -1;
		for (;;) {
			int found = -1;
#pragma unroll
			for (int k = 0; k < 4; ++k) {
				const unsigned long long mk = L[lv].mask[k];
				if (found < 0 && mk) {
					found = k * 64 + (__ffsll((long long)mk) - 1);
				}
			}
			if (found >= 0) {
				cell = found;
				break;
			}
			if (lv == 0) break;
			--lv;
		}
		if (cell < 0) break;
		__builtin_amdgcn_wave_barrier();
		if (lane == 0) L[lv].mask[cell >> 6] &= ~(1ull << (cell & 63));
		const double px0 = L[lv].x0, px1 = L[lv].x1, py0 = L[lv].y0, py1 = L[lv].y1;
		const double pxbin = (px1 - px0) / res, pybin = (py1 - py0) / res;
		const double phx = pxbin / 2., phy = pybin / 2.;
		const int ci = cell / res, cj = cell - ci * res;
		const double cx = cell_centre(px0, phx, ci);
		const double cy = cell_centre(py0, phy, cj);
		nx0 = cx - phx;
		nx1 = cx + phx;
		ny0 = cy - phy;
		ny1 = cy + phy;
		weight = L[lv].weight * inv_cells;
		level = lv + 1;
		__builtin_amdgcn_wave_barrier();
	}
}

// One wavefront per queued pixel of a TYPE component: adds flux_scale * value
// to the canvas (or the aux canvas, or the FP64 canvas) and, with kJac, writes
// the entry's Jacobian and value to wq.ljac / hjac for k_adj_queue.
// `counter`: the pair of hand-out counters of this launch (heavy, light).
// Held to the registers of four waves per SIMD (three for the Sersic, Moffat
// and broken-exponential walks with kJac, whose code needs more).
template <int TYPE, bool kJac>
__global__ __launch_bounds__(256, kJac && (TYPE == SERSIC || TYPE == MOFFAT || TYPE == BROKENEXP) ? 3 : 4) void k_subsample(const GalaxyDev *__restrict__ gals,
                                                                 const InstDev *__restrict__ insts,
                                                                 const Comp *__restrict__ comps, WorkQueue wq,
                                                                 int counter, int split)
{
	__shared__ WalkLevel levels[4][kMaxLevels];
	// the Jacobian's slots of a pixel (with kJac) and its value: per
	// wavefront, one row of 64 lane sums per slot (rows kJacPitch floats
	// apart, so that the 16 lanes reading one quarter of each row fall on
	// distinct banks)
	constexpr int kSlots = G_S0 + n_shape(TYPE), kJacPitch = 68;
	__shared__ __align__(16) float jred[4][(kJac ? kSlots + 1 : 1) * kJacPitch];
	// the wavefront index is uniform: make that visible to the compiler so queue
	// entries and component fields are read through the scalar cache
	const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	WalkLevel *L = levels[wv];
	static_assert(TYPE >= 0 && TYPE < kFamilies, "one walker launch per profile family");
	static_assert(kSlots < kJacStride && kSlots + 1 <= 16, "Jacobian slots and the value");
	const int lb = wq.lbase[TYPE];
	// the light entries the root pass left to this walker (its deep list)
	const int nl = min(wq.count[deep_count(TYPE)], wq.lbase[TYPE + 1] - lb), nh = min(wq.count[1], wq.hcap);
	// per-lane Jacobian sums
	float gj[NGRAD];
#pragma unroll
	for (int s = 0; s < NGRAD; ++s) gj[s] = 0.f;
	// dynamic hand-out, one entry at a time: the centre-pixel trees first,
	// then the light entries the root pass left (each a tree of several
	// nodes), so that no wavefront is left with a long tail
	bool heavy_left = nh > 0;
	for (;;) {
		bool heavy = false;
		int h = nh;
		if (heavy_left) {
			if (lane == 0) h = atomicAdd(&wq.count[counter], 1);
			h = __builtin_amdgcn_readfirstlane(h);
		}
		int e;  // the entry's index in its list (heavy, or the family's light list)
		if (h < nh) {
			e = h;
			heavy = true;
		}
		else {
			heavy_left = false;
			int l = 0;
			if (lane == 0) l = atomicAdd(&wq.count[counter + 1], 1);
			l = __builtin_amdgcn_readfirstlane(l);
			if (l >= nl) break;
			e = wq.deep[lb + l];
		}
		{
			const QEntry q = heavy ? wq.heavy[e] : wq.light[lb + e];
			const Comp &c = comps[q.comp];
			if (c.type != TYPE) continue;
			const InstDev &I = insts[c.instance];
			const GalaxyDev &g = gals[I.galaxy];
			const int i = q.pix % g.DW, j = q.pix / g.DW;
			const double hx = g.sx / 2., hy = g.sy / 2.;
			const double y = hy + j * g.sx;
			const double x = hx + i * g.sy;
			int res = c.res, maxrec = c.maxrec;
			if (centre_pixel(c, x, y)) {
				res = 8;
				maxrec = 10;
			}
			double s_local = 0;
			walk_pixel<TYPE, kJac>(c, L, lane, x - hx, x + hx, y - hy, y + hy, res, maxrec, g.render_fp64 != 0, s_local,
			                       gj);
			const bool to_aux = split && !c.convolve;
			// on the FP32 canvas the value is summed through the wavefront's
			// LDS rows (as one more row of the Jacobian's), the same way with
			// and without the Jacobian, so that a pixel's value does not depend
			// on whether gradients are asked for; the FP64 canvas sums it in
			// FP64 across the wave
			const bool fold = !g.canvas64;
			double tot = 0.0;
			if (!fold) {
				tot = wave_sum(s_local);
				if (lane == 0) {
					if (!to_aux) atomicAdd(&I.canvas64[q.pix], c.flux_scale * tot);
					else atomicAdd(&I.aux[q.pix], (float)(c.flux_scale * tot));
				}
			}
			if (kJac || fold) {
				// the lane sums go to the wavefront's rows (the Jacobian's
				// slots, then the value); lane l adds quarter l >> 4 of row
				// l & 15 in FP64 and two exchanges add the quarters
				constexpr int kRows = kJac ? kSlots : 0;  // the value's row
				float *R = jred[wv];
#pragma unroll
				for (int s = 0; s < kRows; ++s) {
					R[s * kJacPitch + lane] = gj[s];
					gj[s] = 0.f;
				}
				if (fold) R[kRows * kJacPitch + lane] = (float)s_local;
				__builtin_amdgcn_wave_barrier();
				const int slot = lane & 15, qt = lane >> 4;
				double t = 0.0;
				if (slot < kRows + (fold ? 1 : 0)) {
					const float4 *row = reinterpret_cast<const float4 *>(R + slot * kJacPitch + qt * 16);
#pragma unroll
					for (int k = 0; k < 4; ++k) {
						const float4 v = row[k];
						t += ((double)v.x + (double)v.y) + ((double)v.z + (double)v.w);
					}
				}
				__builtin_amdgcn_wave_barrier();
				t += __shfl_xor(t, 16);
				t += __shfl_xor(t, 32);
				float *J = nullptr;
				if (kJac) {
					J = (heavy ? wq.hjac : wq.ljac + (long)lb * kJacStride) + (long)e * kJacStride;
					if (lane < kSlots) J[lane] = (float)t;
					if (!fold && lane == 0) J[kJacStride - 1] = (float)tot;
				}
				if (fold && lane == kRows) {
					if (kJac) J[kJacStride - 1] = (float)t;
					atomicAdd(&(to_aux ? I.aux : I.canvas)[q.pix], (float)(c.flux_scale * t));
				}
			}
		}
	}
}

// The root pass of the sub-sampled pixels, one lane per light entry of a
// TYPE family: the lane evaluates its pixel's root cells one after another
// with the walker's cell step and, when none is to be refined (most pixels:
// 98% on a 400x400 bulge+disk), adds the value to the canvas and, with kJac,
// writes the entry's Jacobian and value itself, with no exchange between
// lanes.  An entry with a cell to refine goes to the family's deep list,
// whose trees the wavefront walker (k_subsample) walks from the root.
// entries from which the root pass works a lane per entry: 2^17, two
// wavefronts of 64 entries for each SIMD of the chip
constexpr int kRootMinEntries = 1 << 17;

template <int TYPE, bool kJac>
__global__ __launch_bounds__(256) void k_subsample_root(const GalaxyDev *__restrict__ gals,
                                                         const InstDev *__restrict__ insts,
                                                         const Comp *__restrict__ comps, WorkQueue wq, int split)
{
	const int lane = threadIdx.x & 63;
	const int lb = wq.lbase[TYPE];
	const int nl = min(wq.count[light_count(TYPE)], wq.lbase[TYPE + 1] - lb);
	// a lane per entry pays off when the entries fill the GPU; a shorter list
	// (a single image, say) goes whole to the walker, whose wavefronts share
	// each pixel's cells and so finish sooner
	const bool lanes = nl >= kRootMinEntries;
	for (;;) {
		int base = 0;
		if (lane == 0) base = atomicAdd(&wq.count[root_counter(TYPE)], 64);
		base = __shfl(base, 0);
		if (base >= nl) break;
		const int e = base + lane;
		bool deep = false;
		if (e < nl) {
			const QEntry q = wq.light[lb + e];
			const Comp &c = comps[q.comp];
			const InstDev &I = insts[c.instance];
			const GalaxyDev &g = gals[I.galaxy];
			const int i = q.pix % g.DW, j = q.pix / g.DW;
			const double hx = g.sx / 2., hy = g.sy / 2.;
			const double y = hy + j * g.sx;
			const double x = hx + i * g.sy;
			// centre-rule pixels belong to the heavy list: any found here go
			// to the walker, which applies the rule
			deep = !lanes || centre_pixel(c, x, y);
			if (c.type == TYPE && !deep) {
				const int res = c.res, ncell = res * res;
				const bool recurse = res > 1 && c.maxrec > 0;
				const bool fuse = kJac && TYPE == SERSIC && !c.boxy && !c.grad_box;
				const Node n = make_node(c, x - hx, x + hx, y - hy, y + hy, res);
				const double weight = kInvCells[res];
				const bool fp64 = g.render_fp64 != 0;
				double sum = 0;
				float gj[NGRAD];
#pragma unroll
				for (int s = 0; s < NGRAD; ++s) gj[s] = 0.f;
				for (int cell = 0; cell < ncell && !deep; ++cell)
					deep = cell_step<TYPE, kJac>(c, n, cell, recurse, fp64, fuse, weight, sum, gj);
				if (!deep) {
					const bool to_aux = split && !c.convolve;
					if (g.canvas64 && !to_aux) atomicAdd(&I.canvas64[q.pix], c.flux_scale * sum);
					else atomicAdd(&(to_aux ? I.aux : I.canvas)[q.pix], (float)(c.flux_scale * sum));
					if (kJac) {
						float *J = wq.ljac + (long)(lb + e) * kJacStride;
#pragma unroll
						for (int s = 0; s < G_S0 + n_shape(TYPE); ++s) J[s] = gj[s];
						J[kJacStride - 1] = (float)sum;
					}
				}
			}
		}
		// the deep entries appended, one atomic per wave
		const unsigned long long m = __ballot(deep);
		if (m) {
			const int leader = __ffsll((long long)m) - 1;
			int at = 0;
			if (lane == leader) at = atomicAdd(&wq.count[deep_count(TYPE)], __popcll(m));
			at = __shfl(at, leader);
			if (deep) wq.deep[lb + at + __popcll(m & ((1ull << lane) - 1ull))] = e;
		}
	}
}

// The adjoint of the sub-sampled pixels from their Jacobians (written by the
// forward walk, k_subsample<TYPE, true>): per entry (component c,
// pixel q), G(q) fs (J_s + dlnfs_s V) for every slot s.  Lanes take
// consecutive entries of the family's lists (the heavy list first, Sersic
// only); the lanes of one component are summed across the wave, one
// component at a time, in FP64, and added with one atomic per slot.
template <int TYPE>
__global__ __launch_bounds__(256) void k_adj_queue(const InstDev *__restrict__ insts, const Comp *__restrict__ comps,
                                                   WorkQueue wq, double *grad_acc)
{
	constexpr int kSlots = G_S0 + n_shape(TYPE);
	const int lane = threadIdx.x & 63;
	const int lb = wq.lbase[TYPE];
	const long nl = min(wq.count[light_count(TYPE)], wq.lbase[TYPE + 1] - lb);
	const long nh = TYPE == SERSIC ? min(wq.count[1], wq.hcap) : 0;
	const long total = nl + nh;
	const long stride = (long)gridDim.x * blockDim.x;
	for (long base = (long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total; base += stride) {
		const long e = base + lane;
		int comp = -1;
		double cs[kSlots];
#pragma unroll
		for (int s = 0; s < kSlots; ++s) cs[s] = 0;
		if (e < total) {
			const bool heavy = e < nh;
			const long ei = heavy ? e : e - nh;
			const QEntry q = heavy ? wq.heavy[ei] : wq.light[lb + ei];
			const float *J = heavy ? wq.hjac + ei * kJacStride : wq.ljac + (lb + ei) * kJacStride;
			const Comp &c = comps[q.comp];
			const float G = insts[c.instance].G[q.pix];
			if (c.type == TYPE && G != 0.f) {
				comp = q.comp;
				const double gw = (double)G * c.flux_scale, V = (double)J[kJacStride - 1];
#pragma unroll
				for (int s = 0; s < kSlots; ++s) cs[s] = gw * ((double)J[s] + c.dlnfs[s] * V);
			}
		}
		unsigned long long left = __ballot(comp >= 0);
		while (left) {
			const int leader = __ffsll((long long)left) - 1;
			const int lc = __builtin_amdgcn_readlane(comp, leader);
			const bool mine = comp == lc;
			left &= ~__ballot(mine);
#pragma unroll
			for (int s = 0; s < kSlots; ++s) {
				const double t = wave_sum(mine ? cs[s] : 0.0);
				if (lane == leader && t != 0) atomicAdd(&grad_acc[(long)lc * NGRAD + s], t);
			}
		}
	}
}

// ------------------------------------------------------------------ adjoint render

// Adjoint of the directly evaluated pixels: sum_q G(q) fs d f(q) / d slot over
// the pixels the forward render marked direct (so no classification is
// repeated), FP32 per-pixel derivatives, per-thread sums in FP32 (in FP64 for
// steep cores, G64: their centroid and shape terms cancel over many pixels,
// which FP32 sums do not survive), the flux-scale term applied once per
// component from sum_q G(q) fs f(q).  A block takes `grp` consecutive canvas
// tiles of one instance (block offsets per instance in grp_off), so that the
// per-component reduction and its atomics are paid once per grp tiles.
// Compiled per profile family TYPE and G64; components of other kinds are
// left to the launch made for them.
template <int TYPE, bool G64>
__global__ __launch_bounds__(256) void k_adj_render(const GalaxyDev *__restrict__ gals,
                                                    const InstDev *__restrict__ insts,
                                                    const int *__restrict__ grp_off, int ninst, int grp,
                                                    const Comp *__restrict__ comps, double *grad_acc)
{
	__shared__ double red[4][NGRAD];
	const int b = blockIdx.x;
	const int inst = owner_of(grp_off, ninst, b);
	const InstDev I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int rows = g.conv_rows;
	const int t0 = (b - grp_off[inst]) * grp, t1 = min(t0 + grp, g.n_atiles);
	const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int wpr = (g.DW + 63) >> 6;
	const double hx = g.sx / 2., hy = g.sy / 2.;
	for (int k = 0; k < I.ncomp; ++k) {
		const Comp &c = comps[I.comp0 + k];
		if (c.type != TYPE || (bool)c.grad64 != G64) continue;
		double gl[NGRAD];
#pragma unroll
		for (int s = 0; s < NGRAD; ++s) gl[s] = 0;
		double vsum = 0;
		if (G64) {
			for (int ti = t0; ti < t1; ++ti) {
				const int2 t = g.atiles[ti];
				const int i = t.x * kTileW + lane;
				const int j0 = t.y * 4 * rows + wv * rows;
				const double x = hx + i * g.sy;
				for (int r = 0; r < rows; ++r) {
					const int j = j0 + r;
					if (j >= g.DH) break;
					const unsigned long long m = I.dmask[((long)k * g.DH + j) * wpr + t.x];
					if (!((m >> lane) & 1ull)) continue;
					const float Gq = I.G[i + j * g.DW];
					if (Gq == 0.f) continue;
					const double y = hy + j * g.sx;
					const double xr = x - c.xc, yr = y - c.yc;
					const double xp = xr * c.cos_ang + yr * c.sin_ang;
					const double yp = -xr * c.sin_over + yr * c.cos_over;
					const double wq = (double)Gq * c.flux_scale;
					vsum += wq * sample_accumulate<float, TYPE>(c, xp, yp, wq, gl);
				}
			}
		}
		else {
			float gf[NGRAD];
#pragma unroll
			for (int s = 0; s < NGRAD; ++s) gf[s] = 0.f;
			float vf = 0.f;
			const float fs = (float)c.flux_scale;
			// a Sersic without box (or box derivative): the derivatives from
			// the value's own logarithm and exponentials
			const bool fuse = TYPE == SERSIC && !c.boxy && !c.grad_box;
			for (int ti = t0; ti < t1; ++ti) {
				const int2 t = g.atiles[ti];
				const int i = t.x * kTileW + lane;
				const int j0 = t.y * 4 * rows + wv * rows;
				const double x = hx + i * g.sy;
				for (int r = 0; r < rows; ++r) {
					const int j = j0 + r;
					if (j >= g.DH) break;
					const unsigned long long m = I.dmask[((long)k * g.DH + j) * wpr + t.x];
					if (!m) continue;
					if (!((m >> lane) & 1ull)) continue;
					const float Gq = I.G[i + j * g.DW];
					if (Gq == 0.f) continue;
					const double y = hy + j * g.sx;
					const float xr = (float)(x - c.xc), yr = (float)(y - c.yc);
					const float xp = xr * c.f.cos_ang + yr * c.f.sin_ang;
					const float yp = -xr * c.f.sin_over + yr * c.f.cos_over;
					const float wq = Gq * fs;
					if (fuse) {
						float base, l2, rf;
						const float v = sersic_parts_f32(c, xp, yp, base, l2, rf);
						sersic_jac_f32(c, xp, yp, v, base, l2, rf, wq, gf);
						vf += wq * v;
					}
					else {
						vf += wq * pixel_accumulate_f32<TYPE>(c, xp, yp, wq, gf);
					}
				}
			}
#pragma unroll
			for (int s = 0; s < NGRAD; ++s) gl[s] = (double)gf[s];
			vsum = (double)vf;
		}
		// flux-scale dependence: d fs / d slot = fs * dlnfs
#pragma unroll
		for (int s = 0; s < G_S0 + n_shape(TYPE); ++s) {
			const double v = wave_sum(gl[s] + vsum * c.dlnfs[s]);
			if (lane == 0) red[wv][s] = v;
		}
		__syncthreads();
		if (threadIdx.x < G_S0 + n_shape(TYPE)) {
			const int s = threadIdx.x;
			const double v = (red[0][s] + red[1][s]) + (red[2][s] + red[3][s]);
			if (v != 0) atomicAdd(&grad_acc[(long)(I.comp0 + k) * NGRAD + s], v);
		}
		__syncthreads();
	}
}

}  // namespace profit_amd
