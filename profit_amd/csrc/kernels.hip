// HIP kernels for the model-image hot path on CDNA4 (gfx950).
//
//   k_setup          per-component FP64 initial_calculations (radial.cpp:136-199)
//   k_render         pixel classification (FP64, bit-exact with libprofit) and
//                    direct FP32 evaluation (radial.cpp:291-344); pixels that need
//                    sub-pixel integration are appended to a work queue
//   k_subsample      one wavefront per queued pixel walks libprofit's adaptive
//                    recursion (radial.cpp:54-134) depth-first, FP32 cell positions
//                    and tests with FP64 re-tests in a guard band; the walk keeps
//                    one node and a 256-bit "cells to refine" mask per level
//   k_conv           LDS-tiled brute-force convolution (convolve.cpp:158-231) with
//                    three epilogues: fused masked likelihood, model image, and
//                    the adjoint correlation
//   k_like_finalize  compensated FP64 reduction of the likelihood partials
//   k_adj_render / k_adj_queue
//                    fixed-tree adjoint of the rendering: sum_q G(q) d pixel_q / d theta
//                    (the sub-sampled pixels' derivatives come from the forward walk)
//   k_dilate_rows/cols  Mask::expand_by (image.cpp:113-162) as a separable dilation
//
// Compiled with -ffp-contract=off: every FP64 operation of the classification
// and of the recursion rounds exactly as libprofit's own build does.
#include <hip/hip_runtime.h>

#include "kernel_common.h"

namespace profit_amd {

// ------------------------------------------------------------------ set-up

// The components' set-up in `order` (grouped by profile family), one
// wavefront per component: the set-up runs alike on every lane and the
// Kronrod points of its luminosity integrals spread over the lanes; the
// QUADPACK luminosity gradients are left to k_setup_grad.
__global__ __launch_bounds__(256, 2) void k_setup(const ProfileParams *__restrict__ params,
                                                  const int *__restrict__ order, const int *__restrict__ comp_inst,
                                                  const InstDev *__restrict__ insts,
                                                  const GalaxyDev *__restrict__ gals, Comp *comps, int *errors,
                                                  int ncomp)
{
	const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
	if (t >= ncomp) return;
	const int c = order[t];
	const int inst = comp_inst[c];
	const GalaxyDev &g = gals[insts[inst].galaxy];
	sf::QuadState qs;
	Comp out;
	int e = setup_component<true>(params[c], g.magzero, g.sx, g.sy, g.img_sx, g.img_sy, g.padx, g.pady, &qs, &out,
	                              false);
	if (e == SE_OK && params[c].type == POINTSRC && !g.psf_pt) e = SE_NO_PSF;
	out.grad_box = g.grad_box;
	out.instance = inst;
	// every lane holds the same result; lane 0 stores it
	if ((threadIdx.x & 63) == 0) {
		comps[c] = out;
		errors[c] = e;
	}
}

// d ln(flux_scale) / d shape of the King, core-Sersic and broken-exponential
// components (lum_gradient): one wavefront per (component, slot), the
// integral's Kronrod points spread over its lanes; components in `order`,
// after k_setup.
__global__ __launch_bounds__(256, 4) void k_setup_grad(const int *__restrict__ order, Comp *comps,
                                                    const int *__restrict__ errors, int ncomp)
{
	const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
	if (t >= ncomp * 5) return;
	const int c = order[t / 5], w = t % 5;
	if (errors[c] != SE_OK || w >= n_lum_integrals(comps[c].type)) return;
	sf::QuadState qs;
	const double v = lum_gradient<true>(comps[c], w, &qs);
	if ((threadIdx.x & 63) == 0) comps[c].dlnfs[G_S0 + w] = v;
}

__global__ void k_fill_zero_d(double *p, long n)
{
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0;
}

__global__ void k_fill_zero_i(int *p, long n)
{
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ classification

struct PixelClass {
	double x, y, xp, yp;
	int cls;  // 1 zero (beyond rscale_max), 2 direct, 3 subsample
};

__device__ inline PixelClass classify(const Comp &c, double x, double y)
{
	PixelClass pc;
	pc.x = x;
	pc.y = y;
	const double xr = x - c.xc, yr = y - c.yc;
	pc.xp = xr * c.cos_ang + yr * c.sin_ang;
	pc.yp = -xr * c.sin_over + yr * c.cos_over;
	const double r = sqrt(pc.xp * pc.xp + pc.yp * pc.yp);
	if (c.has_rmax && r > c.rmax_r) pc.cls = 1;
	else if (c.rough || r > c.rsw_r) pc.cls = 2;
	else pc.cls = 3;
	return pc;
}

// The class of a pixel from FP32 profile coordinates (xp, yp) when r^2 lies
// further than 1e-5 (relative) from both thresholds: the FP32 rotation errs
// by a few parts in 1e7 of r^2, so such a class equals the FP64 one.  -1:
// too close to a threshold, libprofit's FP64 classification decides.
__device__ __forceinline__ int classify32(const Comp &c, float xp, float yp)
{
	const float r2 = xp * xp + yp * yp;
	if (c.has_rmax && fabsf(r2 - c.f.rmax2) <= 1e-5f * c.f.rmax2) return -1;
	if (c.has_rmax && r2 > c.f.rmax2) return 1;
	if (c.rough) return 2;
	if (fabsf(r2 - c.f.rsw2) <= 1e-5f * c.f.rsw2) return -1;
	return r2 > c.f.rsw2 ? 2 : 3;
}

// ------------------------------------------------------------------ forward render

// wave-aggregated append of the lanes with `take` to a work list
__device__ __forceinline__ void append_entry(bool take, int lane, QEntry *list, int *count, int cap, QEntry e)
{
	const unsigned long long m = __ballot(take);
	if (!m) return;
	const int leader = __ffsll((long long)m) - 1;
	int base = 0;
	if (lane == leader) base = atomicAdd(count, __popcll(m));
	base = __shfl(base, leader);
	if (take) {
		const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
		if (slot < cap) list[slot] = e;
	}
}

// ------------------------------------------------------------------ point sources

// bind() of psf.cpp:48-56: the floor of a value clamped to [0, max]
__device__ __forceinline__ int psf_bind(double v, int max)
{
	const double f = floor(v);
	if (f < 0) return 0;
	if (f > (double)max) return max;
	return (int)f;
}

// Footprint of a point source on the canvas (psf.cpp:79-92): the PSF's
// origin and the clamped pixel range it covers.  The x end uses offset.y and
// the y origin offset.x, as the reference does (Comp::s[0..1], xc, yc).
struct PointFoot {
	double ox, oy;
	int x0, x1, y0, y1;
};

__device__ __forceinline__ PointFoot point_footprint(const Comp &c, const GalaxyDev &g)
{
	PointFoot f;
	f.ox = c.xc - g.pkx * g.psf_sx / 2;
	f.oy = c.yc - g.pky * g.psf_sy / 2;
	const double ex = c.s[0] + g.pkx * g.psf_sx / 2, ey = c.s[1] + g.pky * g.psf_sy / 2;
	f.x0 = psf_bind(f.ox / g.sx, g.DW - 1);
	f.y0 = psf_bind(f.oy / g.sy, g.DH - 1);
	f.x1 = psf_bind(ex / g.sx, g.DW - 1);
	f.y1 = psf_bind(ey / g.sy, g.DH - 1);
	return f;
}

// psf.cpp:94-141 at canvas pixel (px, py) of the footprint, before the flux
// scale: the PSF pixels the image pixel overlaps, weighted by the overlap
// area (the reference's index clamps kept; a flat index past the PSF array,
// reachable only through them for a non-square PSF, reads 0).  dvx, dvy: the
// derivatives along xcen and ycen, from overlap lengths that are piecewise
// linear in the centre.
__device__ double point_value(const GalaxyDev &g, const PointFoot &f, int px, int py, double *dvx, double *dvy)
{
	const double sx = g.sx, sy = g.sy, psx = g.psf_sx, psy = g.psf_sy;
	const int kw = g.pkx, kh = g.pky;
	const double x = px * sx, y = py * sy;
	const int qx0 = psf_bind(floor((x - f.ox) / psx), kw - 1);
	const int qx1 = psf_bind(floor((x - f.ox + sx) / psx), kh - 1);
	const int qy0 = psf_bind(floor((y - f.oy) / psy), kw - 1);
	const int qy1 = psf_bind(floor((y - f.oy + sy) / psy), kh - 1);
	const long nk = (long)kw * kh;
	double val = 0, gx = 0, gy = 0;
	for (int qy = qy0; qy <= qy1; ++qy) {
		const double psf_y = qy * psy + f.oy;
		const double iy = fmin(y + sy, psf_y + psy) - fmax(y, psf_y);
		const double diy = (psf_y + psy < y + sy ? 1.0 : 0.0) - (psf_y > y ? 1.0 : 0.0);
		for (int qx = qx0; qx <= qx1; ++qx) {
			const double psf_x = qx * psx + f.ox;
			const double ix = fmin(x + sx, psf_x + psx) - fmax(x, psf_x);
			const double dix = (psf_x + psx < x + sx ? 1.0 : 0.0) - (psf_x > x ? 1.0 : 0.0);
			const long k = qx + (long)qy * kw;
			const double w = k < nk ? (double)g.psf_pt[k] : 0.0;
			val += w * (ix * iy) / (psx * psy);
			gx += w * (dix * iy) / (psx * psy);
			gy += w * (ix * diy) / (psx * psy);
		}
	}
	*dvx = gx;
	*dvy = gy;
	return val;
}

// Tiles are 64 x (4*rows) pixels; thread (tx, ty) covers column tx and rows
// ty*rows .. ty*rows+rows-1.  filter: -1 every component, 0/1 only those whose
// convolve flag equals it.  target: 0 canvas, 1 the aux canvas.  With filter
// -1, point sources and sky drawn without convolution go to the aux canvas
// when the instance has one in use (InstDev::has_aux); radial profiles are
// always convolved there.  F64: the kernel holds the FP64 value path (the
// FP64 model modes, GalaxyDev::render_fp64 / canvas64, and King or Ferrer
// pixels near their cut-off); without it a Sersic render needs 48 VGPRs
// instead of 109, eight waves per SIMD instead of four.
template <int TYPE, bool F64>
__global__ __launch_bounds__(256) void k_render(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                const int *__restrict__ tile_off, int ninst,
                                                const Comp *__restrict__ comps, WorkQueue wq, int filter, int target,
                                                int *diag_cls)
{
	const int b = blockIdx.x;
	const int inst = owner_of(tile_off, ninst, b);
	const InstDev I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int rows = g.conv_rows;
	const int2 t = g.atiles[b - tile_off[inst]];
	const int lane = threadIdx.x & 63;
	const int i = t.x * kTileW + lane;
	const int j0 = t.y * 4 * rows + (threadIdx.x >> 6) * rows;
	const double hx = g.sx / 2., hy = g.sy / 2.;
	float *out = target ? I.aux : I.canvas;
	for (int r = 0; r < rows; ++r) {
		const int j = j0 + r;
		const bool inb = i < g.DW && j < g.DH;
		const int pix = i + j * g.DW;
		const bool active = inb && (!g.has_calc || g.calc[pix]);
		// radial.cpp:313-314: the x centre uses scale.y and the y centre scale.x
		const double y = hy + j * g.sx;
		const double x = hx + i * g.sy;
		float acc = 0.f, aux_acc = 0.f;
		double acc64 = 0.0;  // the FP64 canvas's sum (GalaxyDev::canvas64)
		for (int k = 0; k < I.ncomp; ++k) {
			const Comp &c = comps[I.comp0 + k];
			if (filter >= 0 && c.convolve != filter) continue;
			if (TYPE < 0 && !is_radial(c.type)) {
				// sky.cpp:47-63 (calc pixels only) and psf.cpp:58-145 (mask ignored)
				float v = 0.f;
				if (c.type == SKY) {
					if (active) v = (float)c.flux_scale;
				}
				else if (c.type == POINTSRC && inb) {
					const PointFoot f = point_footprint(c, g);
					if (i >= f.x0 && i <= f.x1 && j >= f.y0 && j <= f.y1) {
						double dvx, dvy;
						v = (float)(point_value(g, f, i, j, &dvx, &dvy) * c.flux_scale);
					}
				}
				if (filter < 0 && !c.convolve && I.has_aux) aux_acc += v;
				else if (F64 && g.canvas64 && !target) acc64 += (double)v;
				else acc += v;
				if (diag_cls && inb) diag_cls[(long)(I.comp0 + k) * g.DW * g.DH + pix] = 0;
				if (I.dmask && lane == 0 && j < g.DH) I.dmask[((long)k * g.DH + j) * ((g.DW + 63) >> 6) + t.x] = 0ull;
				continue;
			}
			bool sub = false, heavy = false;
			int cls = 0;
			if (active) {
				// FP32 profile coordinates from the FP64 offsets; FP64 ones
				// (libprofit's) only where the class or the value needs them
				const float fxr = (float)(x - c.xc), fyr = (float)(y - c.yc);
				const float fxp = fxr * c.f.cos_ang + fyr * c.f.sin_ang;
				const float fyp = -fxr * c.f.sin_over + fyr * c.f.cos_over;
				cls = classify32(c, fxp, fyp);
				if (cls < 0) cls = classify(c, x, y).cls;
				if (cls == 2) {
					const bool c64 = F64 && g.canvas64 && !target;
					if (F64 && (c64 || g.render_fp64 || near_cut_fp64<TYPE>(c, fxp, fyp))) {
						const PixelClass pc = classify(c, x, y);
						const double v = c.flux_scale * eval_at<double, TYPE>(c, pc.xp, pc.yp);
						if (c64) acc64 += v;
						else acc += (float)v;
					}
					else {
						acc += (float)c.flux_scale * eval_at<float, TYPE>(c, fxp, fyp);
					}
				}
				else if (cls == 3) {
					// the deep centre-pixel trees (sersic.cpp:331-347) go to their own list
					heavy = centre_pixel(c, x, y);
					sub = !heavy;
				}
			}
			if (diag_cls && inb) diag_cls[(long)(I.comp0 + k) * g.DW * g.DH + pix] = cls;
			// direct-pixel mask for the adjoint render
			const unsigned long long dm = __ballot(cls == 2);
			if (I.dmask && lane == 0 && j < g.DH)
				I.dmask[((long)k * g.DH + j) * ((g.DW + 63) >> 6) + t.x] = dm;
			const int fam = TYPE >= 0 ? TYPE : c.type;
			append_entry(sub, lane, wq.light + wq.lbase[fam], wq.count + light_count(fam),
			             wq.lbase[fam + 1] - wq.lbase[fam], QEntry{I.comp0 + k, pix});
			append_entry(heavy, lane, wq.heavy, wq.count + 1, wq.hcap, QEntry{I.comp0 + k, pix});
		}
		if (inb) {
			if (F64 && g.canvas64 && !target) I.canvas64[pix] = acc64;
			else out[pix] = acc;
		}
		if (inb && filter < 0 && I.has_aux) I.aux[pix] = aux_acc;
	}
}

// ------------------------------------------------------------------ recursion diagnostics

// Counts the nodes visited by the recursion of each queued pixel (diagnostics
// for decision parity with libprofit).
__global__ __launch_bounds__(256) void k_subsample_nodes(const GalaxyDev *__restrict__ gals,
                                                         const InstDev *__restrict__ insts,
                                                         const Comp *__restrict__ comps,
                                                         WorkQueue wq, int *nodes_out)
{
	__shared__ WalkLevel levels[4][kMaxLevels];
	// the wavefront index is uniform: make that visible to the compiler so queue
	// entries and component fields are read through the scalar cache
	const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	WalkLevel *L = levels[wv];
	const int gw = blockIdx.x * 4 + wv, nw = gridDim.x * 4;
	const int nh = min(wq.count[1], wq.hcap);
	int nlf[kFamilies], total = nh;
	for (int f = 0; f < kFamilies; ++f) {
		nlf[f] = min(wq.count[light_count(f)], wq.lbase[f + 1] - wq.lbase[f]);
		total += nlf[f];
	}
	for (int e = gw; e < total; e += nw) {
		QEntry q;
		if (e < nh) {
			q = wq.heavy[e];
		}
		else {
			int r = e - nh, f = 0;
			while (r >= nlf[f]) r -= nlf[f++];
			q = wq.light[wq.lbase[f] + r];
		}
		const Comp &c = comps[q.comp];
		const InstDev &I = insts[c.instance];
		const GalaxyDev &g = gals[I.galaxy];
		const int i = q.pix % g.DW, j = q.pix / g.DW;
		const double hx = g.sx / 2., hy = g.sy / 2.;
		const double y = hy + j * g.sx, x = hx + i * g.sy;
		int res = c.res, maxrec = c.maxrec;
		if (c.centre_rule && fabs(x - c.xc) < c.img_sx && fabs(y - c.yc) < c.img_sy) {
			res = 8;
			maxrec = 10;
		}
		// count nodes: one per processed node, i.e. 1 + number of refined cells
		const int ncell = res * res, nchunk = (ncell + 63) >> 6;
		int nodes = 0;
		int level = 0;
		double nx0 = x - hx, nx1 = x + hx, ny0 = y - hy, ny1 = y + hy;
		for (;;) {
			++nodes;
			const double xbin = (nx1 - nx0) / res, ybin = (ny1 - ny0) / res;
			const double chx = xbin / 2., chy = ybin / 2.;
			const bool recurse = res > 1 && level < maxrec;
			unsigned long long m[4] = {0ull, 0ull, 0ull, 0ull};
			for (int k = 0; k < nchunk; ++k) {
				const int cell = k * 64 + lane;
				bool fail = false;
				if (cell < ncell && recurse) {
					// the production decision path (the walker's FP32 test at FP32
					// cell positions, FP64 at libprofit's position when undecided)
					const int ci = cell / res, cj = cell - ci * res;
					const double dyp = -xbin * c.sin_over + ybin * c.cos_over;
					float xp, yp;
					cell_pos32(c, (float)(nx0 - c.xc), (float)(ny0 - c.yc), (float)xbin, (float)ybin, ci, cj, &xp, &yp);
					const float v32 = eval_at<float>(c, xp, yp);
					int dec = refine_cell32(c, xp, yp, (float)dyp, v32);
					if (dec == 2) {
						double xp64, yp64, v64;
						cell_pos64(c, nx0, ny0, chx, chy, ci, cj, &xp64, &yp64);
						dec = refine_cell64(c, xp64, yp64, dyp, &v64) ? 1 : 0;
					}
					fail = dec == 1;
				}
				m[k] = __ballot(fail);
			}
			if (lane == 0) {
				L[level].x0 = nx0;
				L[level].x1 = nx1;
				L[level].y0 = ny0;
				L[level].y1 = ny1;
#pragma unroll
				for (int k = 0; k < 4; ++k) L[level].mask[k] = m[k];
			}
			__builtin_amdgcn_wave_barrier();
			int lv = level, cell = -1;
			for (;;) {
				int found = -1;
#pragma unroll
				for (int k = 0; k < 4; ++k) {
					const unsigned long long mk = L[lv].mask[k];
					if (found < 0 && mk) found = k * 64 + (__ffsll((long long)mk) - 1);
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
			const double phx = ((px1 - px0) / res) / 2., phy = ((py1 - py0) / res) / 2.;
			const int ci = cell / res, cj = cell - ci * res;
			const double cx = cell_centre(px0, phx, ci), cy = cell_centre(py0, phy, cj);
			nx0 = cx - phx;
			nx1 = cx + phx;
			ny0 = cy - phy;
			ny1 = cy + phy;
			level = lv + 1;
			__builtin_amdgcn_wave_barrier();
		}
		if (lane == 0) nodes_out[(long)q.comp * g.DW * g.DH + q.pix] = nodes;
	}
}

// ------------------------------------------------------------------ convolution


// out(p) = sum_{ky,kx} in(p.y + oy + ky, p.x + ox + kx) * wt[ky*Kx + kx],
// in() is zero outside its image.  Block = 64 columns x 4 row groups; each
// thread produces R consecutive output rows of one column with a sliding
// register window along ky and PSF weights read through the scalar cache.
// The kernel rows are processed in chunks of at most KC rows held in LDS.
template <int R, int MODE>
__global__ __launch_bounds__(256) void k_conv(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                              const int *__restrict__ tile_off, int ninst, int kc_max,
                                              const int *__restrict__ lk_base, double *partial)
{
	extern __shared__ float lds[];
	const int b = blockIdx.x;
	const int inst = owner_of(tile_off, ninst, b);
	const InstDev I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	const int2 t = (MODE == CONV_LIKE) ? g.ctiles[b - tile_off[inst]] : g.atiles[b - tile_off[inst]];
	const int Kx = g.Kx, Ky = g.Ky;
	const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	const int x0 = t.x * kTileW, y0 = t.y * 4 * R;

	const float *__restrict__ in;
	const float *__restrict__ wt;
	int IW, IH, ox, oy, OW, OH;
	if (MODE == CONV_ADJ) {
		in = I.w;
		IW = g.W;
		IH = g.H;
		OW = g.DW;
		OH = g.DH;
		ox = -g.padx + Kx / 2 - Kx + 1;
		oy = -g.pady + Ky / 2 - Ky + 1;
		wt = g.psf_adj;
	}
	else {
		in = I.canvas;
		IW = g.DW;
		IH = g.DH;
		if (MODE == CONV_LIKE) {
			OW = g.W;
			OH = g.H;
			ox = g.padx - Kx / 2;
			oy = g.pady - Ky / 2;
		}
		else {
			OW = g.DW;
			OH = g.DH;
			ox = -(Kx / 2);
			oy = -(Ky / 2);
		}
		wt = g.psf_fwd;
	}
	const int LW = kTileW + Kx - 1;

	// per-column (kx) partial sums in FP32, accumulated across columns in FP64
	double acc[R];
#pragma unroll
	for (int r = 0; r < R; ++r) acc[r] = 0.0;

	for (int ky0 = 0; ky0 < Ky; ky0 += kc_max) {
		const int kc = min(kc_max, Ky - ky0);
		const int LH = 4 * R + kc - 1;
		__syncthreads();
		const int gy0 = y0 + oy + ky0, gx0 = x0 + ox;
		for (int idx = threadIdx.x; idx < LH * LW; idx += 256) {
			const int rr = idx / LW, cc = idx - rr * LW;
			const int gy = gy0 + rr, gx = gx0 + cc;
			float v = 0.f;
			if (gy >= 0 && gy < IH && gx >= 0 && gx < IW) v = in[(long)gy * IW + gx];
			lds[idx] = v;
		}
		__syncthreads();
		const float *colbase = lds + ty * R * LW + tx;
		for (int kx = 0; kx < Kx; ++kx) {
			const float *col = colbase + kx;
			const float *wcol = wt + (long)ky0 * Kx + kx;
			float part[R];
#pragma unroll
			for (int r = 0; r < R; ++r) part[r] = 0.f;
			float win[R];
#pragma unroll
			for (int r = 0; r < R - 1; ++r) win[r] = col[r * LW];
			int ky = 0;
			for (; ky + R <= kc; ky += R) {
#pragma unroll
				for (int u = 0; u < R; ++u) {
					win[(u + R - 1) % R] = col[(ky + u + R - 1) * LW];
					const float w = wcol[(ky + u) * Kx];
#pragma unroll
					for (int r = 0; r < R; ++r) part[r] = fmaf(win[(u + r) % R], w, part[r]);
				}
			}
			for (; ky < kc; ++ky) {
				const float w = wcol[ky * Kx];
#pragma unroll
				for (int r = 0; r < R; ++r) part[r] = fmaf(col[(ky + r) * LW], w, part[r]);
			}
#pragma unroll
			for (int r = 0; r < R; ++r) acc[r] += (double)part[r];
		}
	}

	const int oxx = x0 + tx;
	if (MODE == CONV_LIKE) {
		double part = 0;
		const bool pois = g.likefunc == LIKE_POIS, tl = g.likefunc == LIKE_T;
		const double s = g.pois_scale;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const int oyy = y0 + ty * R + r;
			if (oxx < OW && oyy < OH) {
				const long o = (long)oyy * OW + oxx;
				float wv = 0.f;
				if (g.region[o]) {
					// the model is zero outside the calc region (libprofit masks the
					// convolution output there); such pixels do not feed the adjoint
					const long q = (long)(oyy + g.pady) * g.DW + oxx + g.padx;
					const bool in_calc = !g.has_calc || g.calc[q];
					// an FP64 canvas meets k_conv only through the unit kernel of a
					// galaxy without a PSF (its PSF galaxies take the FFT)
					const double conv = g.canvas64 ? I.canvas64[q] : acc[r];
					const double m = in_calc ? conv + (I.has_aux ? (double)I.aux[q] : 0.0) : 0.0;
					const double d = g.data[o];
					if (pois) {
						const double mm = s * m, dd = s * d;
						part += mm - dd * log(mm);
						wv = (float)(-s * (1.0 - d / m));
					}
					else {
						const double diff = d - m;
						const double iv = g.ivar[o];
						part += diff * diff * iv;
						wv = (float)(diff * iv);
						// t: the standardised residual, for k_like_t
						if (tl) wv = (float)(diff * sqrt(iv));
					}
					if (!in_calc && !tl) wv = 0.f;
				}
				I.w[o] = wv;
			}
		}
		__shared__ double red[4];
		part = wave_sum(part);
		if (tx == 0) red[ty] = part;
		__syncthreads();
		if (threadIdx.x == 0) partial[lk_base[inst] + (b - tile_off[inst])] = (red[0] + red[1]) + (red[2] + red[3]);
	}
	else {
		float *out = I.G;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const int oyy = y0 + ty * R + r;
			if (oxx < OW && oyy < OH) {
				const long o = (long)oyy * OW + oxx;
				float v = (float)acc[r];
				if (MODE == CONV_MODEL && g.canvas64) v = (float)I.canvas64[o];
				if (MODE == CONV_MODEL && g.has_calc && !g.calc[o]) v = 0.f;
				if (MODE == CONV_MODEL && I.has_aux) v += I.aux[o];
				out[o] = v;
			}
		}
	}
}

// ------------------------------------------------------------------ likelihood finalisation

// The likelihood partials of instance i sit at [lk_off[i], lk_off[i+1]),
// written by whichever convolution kernel its galaxy uses.
__global__ void k_like_finalize(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                const int *__restrict__ lk_off, int ninst, const double *__restrict__ partial,
                                double *ll, double *gscale)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= ninst) return;
	const GalaxyDev &g = gals[insts[i].galaxy];
	// Neumaier-compensated sum of the block partials
	double s = 0, comp = 0;
	for (int b = lk_off[i]; b < lk_off[i + 1]; ++b) {
		const double v = partial[b];
		const double tsum = s + v;
		if (fabs(s) >= fabs(v)) comp += (s - tsum) + v;
		else comp += (v - tsum) + s;
		s = tsum;
	}
	const double X = s + comp;
	const double N = g.n_region;
	double L = 0, gsc = 1;
	switch (g.likefunc) {
	case LIKE_CHISQ: {
		// dchisq(X, N, log=TRUE) = dgamma(X, N/2, scale=2, log=TRUE)
		const double k = 0.5 * N;
		if (k >= 1) L = (k - 1) * log(X / 2) - X / 2 - lgamma(k) - log(2.0);
		else L = k * log(X / 2) - X / 2 - lgamma(k) - log(X);
		gsc = -2 * ((k - 1) / X - 0.5);
		break;
	}
	case LIKE_NORM:
		L = -N * 0.918938533204672741780329736406 - 0.5 * X;
		gsc = 1;
		break;
	case LIKE_POIS:
		L = -X;
		gsc = 1;
		break;
	default: L = NAN; break;
	}
	ll[i] = L;
	gscale[i] = gsc;
}

// Adjoint weight of canvas pixel (x, y) for a component drawn into the
// convolved canvas (G) or into the aux canvas, whose pixels enter the model
// image unconvolved (w at the image pixel, zero on the padding).
__device__ __forceinline__ double adjoint_weight(const GalaxyDev &g, const InstDev &I, bool via_aux, int x, int y)
{
	if (!via_aux) return (double)I.G[(long)y * g.DW + x];
	const int xi = x - g.padx, yi = y - g.pady;
	if (xi < 0 || yi < 0 || xi >= g.W || yi >= g.H) return 0.0;
	return (double)I.w[(long)yi * g.W + xi];
}

// Sum over the block (blockDim.x a multiple of 64, at most 1024) in a fixed
// order; every thread receives the total.
__device__ double block_sum(double v, double *sh)
{
	v = wave_sum(v);
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
	__syncthreads();
	if (lane == 0) sh[wv] = v;
	__syncthreads();
	double t = 0;
	for (int k = 0; k < nw; ++k) t += sh[k];
	return t;
}

// Student-t likelihood (R/profitLikeModel.R:281-286), one block per instance.
// The convolution epilogue has left the standardised residuals
// z = (data - model) / sigma of the likelihood pixels in w.  The degrees of
// freedom follow from their sample variance v, dof = max(1, 2v / (v - 1))
// (infinite, i.e. the normal density, when v = 1 or fewer than two pixels);
// logL = sum log dt(z, dof).  With `grad`, w becomes d logL / d model,
// including the dependence of dof on every residual:
//   d logL / dz = -(dof + 1) z / (dof + z^2) + dL/d dof * d dof/dv * 2 (z - mean) / (N - 1).
__global__ __launch_bounds__(1024) void k_like_t(const GalaxyDev *__restrict__ gals, const InstDev *__restrict__ insts,
                                                 int grad, double *ll, double *gscale)
{
	__shared__ double sh[16];
	const int inst = blockIdx.x;
	const InstDev I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	if (g.likefunc != LIKE_T) return;
	const long n = (long)g.W * g.H;
	const double N = g.n_region;
	double a = 0;
	for (long o = threadIdx.x; o < n; o += blockDim.x)
		if (g.region[o]) a += (double)I.w[o];
	const double mean = block_sum(a, sh) / N;
	a = 0;
	for (long o = threadIdx.x; o < n; o += blockDim.x)
		if (g.region[o]) {
			const double d = (double)I.w[o] - mean;
			a += d * d;
		}
	const double var = block_sum(a, sh) / (N - 1);
	double dof = 2 * var / (var - 1);
	bool clamped = false;
	if (N < 2 || dof != dof) dof = HUGE_VAL;
	else if (!(dof > 1)) {
		dof = 1;
		clamped = true;
	}
	const bool normal = sf::is_inf(dof);
	double sl = 0, sc = 0;
	for (long o = threadIdx.x; o < n; o += blockDim.x)
		if (g.region[o]) {
			const double z = (double)I.w[o];
			if (normal) sl += z * z;
			else {
				sl += log1p(z * z / dof);
				sc += z * z / (dof * (dof + z * z));
			}
		}
	sl = block_sum(sl, sh);
	sc = block_sum(sc, sh);
	double L;
	if (normal) L = -N * 0.918938533204672741780329736406 - 0.5 * sl;
	else L = N * (lgamma((dof + 1) / 2) - lgamma(dof / 2) - 0.5 * log(dof * M_PI)) - 0.5 * (dof + 1) * sl;
	if (threadIdx.x == 0) {
		ll[inst] = L;
		gscale[inst] = 1;
	}
	if (!grad) return;
	double kdof = 0;
	if (!normal && !clamped) {
		const double dLdnu = N * (0.5 * sf::digamma((dof + 1) / 2) - 0.5 * sf::digamma(dof / 2) - 0.5 / dof) - 0.5 * sl +
		                     0.5 * (dof + 1) * sc;
		kdof = dLdnu * (-2 / ((var - 1) * (var - 1))) * 2 / (N - 1);
	}
	for (long o = threadIdx.x; o < n; o += blockDim.x) {
		if (!g.region[o]) continue;
		const long y = o / g.W, x = o - y * g.W;
		const bool in_calc = !g.has_calc || g.calc[(y + g.pady) * g.DW + x + g.padx];
		const double z = (double)I.w[o];
		const double dz = normal ? -z : -(dof + 1) * z / (dof + z * z) + kdof * (z - mean);
		I.w[o] = in_calc ? (float)(-dz * sqrt((double)g.ivar[o])) : 0.f;
	}
}

// Gradient of the point sources and the sky (one block per instance): the
// adjoint weight of canvas pixel q is G(q) for a convolved component and
// w(q - padding) on the image for one drawn without convolution.
//   point source: sum_q W(q) fs d val(q) / d(xcen, ycen) and fs val(q) d ln fs / d mag
//   sky:          sum over calc pixels of W(q) / fs^2 into the bg slot
__global__ __launch_bounds__(256) void k_adj_nonradial(const GalaxyDev *__restrict__ gals,
                                                       const InstDev *__restrict__ insts,
                                                       const Comp *__restrict__ comps, double *grad_acc)
{
	__shared__ double sh[4];
	const int inst = blockIdx.x;
	const InstDev I = insts[inst];
	const GalaxyDev &g = gals[I.galaxy];
	for (int k = 0; k < I.ncomp; ++k) {
		const Comp &c = comps[I.comp0 + k];
		if (c.type != SKY && c.type != POINTSRC) continue;
		const bool via_aux = !c.convolve && I.has_aux;
		double a0 = 0, a1 = 0, a2 = 0;
		if (c.type == SKY) {
			const long n = (long)g.DW * g.DH;
			for (long o = threadIdx.x; o < n; o += blockDim.x) {
				if (g.has_calc && !g.calc[o]) continue;
				const int y = (int)(o / g.DW), x = (int)(o - (long)y * g.DW);
				a0 += adjoint_weight(g, I, via_aux, x, y);
			}
		}
		else {
			const PointFoot f = point_footprint(c, g);
			const int fw = f.x1 - f.x0 + 1, fh = f.y1 - f.y0 + 1;
			for (int t = threadIdx.x; t < fw * fh; t += blockDim.x) {
				const int y = f.y0 + t / fw, x = f.x0 + t % fw;
				const double wq = adjoint_weight(g, I, via_aux, x, y);
				if (wq == 0) continue;
				double dvx, dvy;
				const double v = point_value(g, f, x, y, &dvx, &dvy);
				a0 += wq * dvx;
				a1 += wq * dvy;
				a2 += wq * v;
			}
		}
		a0 = block_sum(a0, sh);
		a1 = block_sum(a1, sh);
		a2 = block_sum(a2, sh);
		if (threadIdx.x == 0) {
			double *gr = grad_acc + (long)(I.comp0 + k) * NGRAD;
			if (c.type == SKY) gr[G_S0] += a0 * c.s[2];
			else {
				gr[G_XCEN] += c.flux_scale * a0;
				gr[G_YCEN] += c.flux_scale * a1;
				gr[G_MAG] += c.flux_scale * a2 * c.dlnfs[G_MAG];
			}
		}
		__syncthreads();
	}
}

__global__ void k_grad_finalize(const Comp *__restrict__ comps, const double *__restrict__ gscale,
                                const double *__restrict__ grad_acc, double *grad_out, int ncomp, unsigned slot_mask)
{
	const int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= ncomp * NGRAD) return;
	const int c = idx / NGRAD, s = idx - c * NGRAD;
	grad_out[idx] = ((slot_mask >> s) & 1u) ? gscale[comps[c].instance] * grad_acc[idx] : 0.0;
}

// ------------------------------------------------------------------ masks

// Mask::expand_by: true where any input pixel within +-pad is true.  Done as a
// row pass followed by a column pass (box dilation is separable).
__global__ void k_dilate_rows(const uint8_t *__restrict__ in, uint8_t *out, int w, int h, int pad)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int j = blockIdx.y;
	if (i >= w || j >= h) return;
	const int lo = max(0, i - pad), hi = min(w - 1, i + pad);
	uint8_t v = 0;
	for (int k = lo; k <= hi && !v; ++k) v = in[(long)j * w + k] ? 1 : 0;
	out[(long)j * w + i] = v;
}

__global__ void k_dilate_cols(const uint8_t *__restrict__ in, uint8_t *out, int w, int h, int pad)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int j = blockIdx.y;
	if (i >= w || j >= h) return;
	const int lo = max(0, j - pad), hi = min(h - 1, j + pad);
	uint8_t v = 0;
	for (int k = lo; k <= hi && !v; ++k) v = in[(long)k * w + i] ? 1 : 0;
	out[(long)j * w + i] = v;
}

// ------------------------------------------------------------------ launch wrappers

#define PA_CHECK(x)                                  \
	do {                                             \
		hipError_t e_ = (x);                         \
		if (e_ != hipSuccess) return (int)e_;        \
	} while (0)

int launch_setup(const ProfileParams *params, const int *order, const int *comp_inst, const InstDev *insts,
                 const GalaxyDev *gals, Comp *comps, int *errors, int ncomp, hipStream_t st)
{
	if (ncomp <= 0) return 0;
	k_setup<<<(unsigned)(((long)ncomp * 64 + 255) / 256), 256, 0, st>>>(params, order, comp_inst, insts, gals, comps,
	                                                                   errors, ncomp);
	return (int)hipGetLastError();
}

int launch_setup_grad(const int *order, Comp *comps, const int *errors, int ncomp, hipStream_t st)
{
	if (ncomp <= 0) return 0;
	k_setup_grad<<<(unsigned)(((long)ncomp * 5 * 64 + 255) / 256), 256, 0, st>>>(order, comps, errors, ncomp);
	return (int)hipGetLastError();
}

int launch_zero_d(double *p, long n, hipStream_t st)
{
	if (n <= 0) return 0;
	k_fill_zero_d<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n);
	return (int)hipGetLastError();
}

int launch_zero_i(int *p, long n, hipStream_t st)
{
	if (n <= 0) return 0;
	k_fill_zero_i<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n);
	return (int)hipGetLastError();
}

int launch_render(const GalaxyDev *gals, const InstDev *insts, const int *tile_off, int ninst, int nblocks,
                  const Comp *comps, WorkQueue wq, int filter, int target, int *diag_cls, hipStream_t st, int type,
                  bool fp64)
{
	if (nblocks <= 0) return 0;
	// type >= 0: every component of the batch is of that radial family (the
	// kernel then holds that family's code alone); -1 any mix.  King and
	// Ferrer always take the kernel with the FP64 value path.
	if (type == KING || type == FERRER) fp64 = true;
#define PA_RENDER(T, F) k_render<T, F><<<nblocks, 256, 0, st>>>(gals, insts, tile_off, ninst, comps, wq, filter, target, diag_cls)
	if (fp64) {
		switch (type) {
		case SERSIC: PA_RENDER(SERSIC, true); break;
		case MOFFAT: PA_RENDER(MOFFAT, true); break;
		case FERRER: PA_RENDER(FERRER, true); break;
		case KING: PA_RENDER(KING, true); break;
		case CORESERSIC: PA_RENDER(CORESERSIC, true); break;
		case BROKENEXP: PA_RENDER(BROKENEXP, true); break;
		default: PA_RENDER(-1, true); break;
		}
	}
	else {
		switch (type) {
		case SERSIC: PA_RENDER(SERSIC, false); break;
		case MOFFAT: PA_RENDER(MOFFAT, false); break;
		case CORESERSIC: PA_RENDER(CORESERSIC, false); break;
		case BROKENEXP: PA_RENDER(BROKENEXP, false); break;
		default: PA_RENDER(-1, false); break;
		}
	}
#undef PA_RENDER
	return (int)hipGetLastError();
}

int launch_subsample_nodes(int nblocks, const GalaxyDev *gals, const InstDev *insts, const Comp *comps,
                           WorkQueue wq, int *nodes, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	k_subsample_nodes<<<nblocks, 256, 0, st>>>(gals, insts, comps, wq, nodes);
	return (int)hipGetLastError();
}

template <int R, int MODE>
static int conv_launch(int nblocks, size_t lds, const GalaxyDev *gals, const InstDev *insts, const int *tile_off,
                       int ninst, int kc, const int *lk_base, double *partial, hipStream_t st)
{
	if (lds > 64 * 1024) {
		PA_CHECK(hipFuncSetAttribute((const void *)k_conv<R, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                             (int)lds));
	}
	k_conv<R, MODE><<<nblocks, 256, lds, st>>>(gals, insts, tile_off, ninst, kc, lk_base, partial);
	return (int)hipGetLastError();
}

int launch_conv(int mode, int rows, int nblocks, int kx_max, int kc, const GalaxyDev *gals, const InstDev *insts,
                const int *tile_off, int ninst, const int *lk_base, double *partial, hipStream_t st)
{
	if (nblocks <= 0) return 0;
	const size_t lds = sizeof(float) * (size_t)(4 * rows + kc - 1) * (size_t)(kTileW + kx_max - 1);
	if (lds > 160 * 1024) return (int)hipErrorInvalidValue;
#define PA_CONV_CASE(RR)                                                                                  \
	case RR:                                                                                              \
		if (mode == CONV_LIKE) return conv_launch<RR, CONV_LIKE>(nblocks, lds, gals, insts, tile_off, ninst, kc, lk_base, partial, st); \
		if (mode == CONV_MODEL) return conv_launch<RR, CONV_MODEL>(nblocks, lds, gals, insts, tile_off, ninst, kc, lk_base, partial, st); \
		return conv_launch<RR, CONV_ADJ>(nblocks, lds, gals, insts, tile_off, ninst, kc, lk_base, partial, st);
	switch (rows) {
		PA_CONV_CASE(2)
		PA_CONV_CASE(4)
		PA_CONV_CASE(8)
		PA_CONV_CASE(16)
	default: return (int)hipErrorInvalidValue;
	}
#undef PA_CONV_CASE
}

int launch_like_finalize(const GalaxyDev *gals, const InstDev *insts, const int *lk_off, int ninst,
                         const double *partial, double *ll, double *gscale, hipStream_t st)
{
	if (ninst <= 0) return 0;
	k_like_finalize<<<(ninst + 63) / 64, 64, 0, st>>>(gals, insts, lk_off, ninst, partial, ll, gscale);
	return (int)hipGetLastError();
}

int launch_adj_nonradial(const GalaxyDev *gals, const InstDev *insts, int ninst, const Comp *comps, double *grad_acc,
                         hipStream_t st)
{
	if (ninst <= 0) return 0;
	k_adj_nonradial<<<ninst, 256, 0, st>>>(gals, insts, comps, grad_acc);
	return (int)hipGetLastError();
}

int launch_like_t(const GalaxyDev *gals, const InstDev *insts, int ninst, int grad, double *ll, double *gscale,
                  hipStream_t st)
{
	if (ninst <= 0) return 0;
	k_like_t<<<ninst, 1024, 0, st>>>(gals, insts, grad, ll, gscale);
	return (int)hipGetLastError();
}

int launch_grad_finalize(const Comp *comps, const double *gscale, const double *grad_acc, double *grad_out,
                         int ncomp, unsigned slot_mask, hipStream_t st)
{
	if (ncomp <= 0) return 0;
	const int n = ncomp * NGRAD;
	k_grad_finalize<<<(n + 255) / 256, 256, 0, st>>>(comps, gscale, grad_acc, grad_out, ncomp, slot_mask);
	return (int)hipGetLastError();
}

int launch_dilate(const uint8_t *in, uint8_t *tmp, uint8_t *out, int w, int h, int padx, int pady, hipStream_t st)
{
	dim3 grid((w + 255) / 256, h);
	k_dilate_rows<<<grid, 256, 0, st>>>(in, tmp, w, h, padx);
	k_dilate_cols<<<grid, 256, 0, st>>>(tmp, out, w, h, pady);
	return (int)hipGetLastError();
}

}  // namespace profit_amd
