"""ProFit's fitting interface on the batched engine: the parameter handling of
profitSetupData / profitRemakeModellist / profitLikeModel
(R/profitSetupData.R:151-345, R/profitRemakeModelList.R:1-166,
R/profitLikeModel.R:225-342) and profitMakePriors (R/profitUtility.R:124-176),
with the gradient of the log posterior with respect to the fitted parameters
(the engine's adjoint gradient through the same transforms).

A model list is a dict of components, each a dict of parameter vectors, as
in R: {"sersic": {"xcen": [x1, x2], "mag": [...], ...}, "pointsource":
{"xcen": [...], "ycen": [...], "mag": [...]}, "sky": {"bg": [b]}}.  tofit
(True / False / None = tied to the previous entry, R's NA), tolog and
intervals ({"sersic": {"re": [(lo, hi), (lo, hi)]}}) follow the same
structure.  Flattening follows R's unlist order (components, then parameters,
then vector elements) and names ("sersic.xcen1", or "sky.bg" for a length-1
vector).

Gradient of LP(parm) = LL(theta(parm)) + priors(theta(parm)): theta = 10^parm
for logged parameters (d theta / d parm = theta ln 10), offsets shift or scale
theta, tied entries take their parent's value (their gradient flows to the
parent), and an interval-clamped value does not move with parm (zero
derivative), exactly the steps profitRemakeModellist applies.  A constraints
function has no derivative here: like_grad refuses it.  Priors built by
make_priors have analytic gradients; any other priors callable is
differentiated by central differences in parm (host arithmetic only).
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from . import GRAD_SLOTS, PROFILE_TYPES, SHAPE_NAMES

RADIAL = ("sersic", "moffat", "ferrer", "ferrers", "coresersic", "king", "brokenexp")
LN10 = math.log(10.0)
SLOT = {n: i for i, n in enumerate(GRAD_SLOTS)}


def _vec(v):
    return [float(x) for x in np.atleast_1d(np.asarray(v, dtype=float))]


def unlist(tree, leaf=float):
    """R's unlist of a two-level list: (names, values)."""
    names, vals = [], []
    for comp, params in tree.items():
        for p, v in params.items():
            v = list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v]
            if len(v) == 1:
                names.append(f"{comp}.{p}")
            else:
                names += [f"{comp}.{p}{k + 1}" for k in range(len(v))]
            vals += [leaf(x) if x is not None else None for x in v]
    return names, vals


def relist(values, skeleton):
    """R's relist: values in unlist order into the skeleton's structure."""
    out, k = {}, 0
    for comp, params in skeleton.items():
        out[comp] = {}
        for p, v in params.items():
            n = len(v) if isinstance(v, (list, tuple, np.ndarray)) else 1
            out[comp][p] = [float(x) for x in values[k:k + n]]
            k += n
    return out


def _merge(default, given):
    """profitSetupData's partial tofit / tolog (R/profitSetupData.R:177-222):
    entries the user gives replace the default ones."""
    out = {c: {p: list(v) for p, v in ps.items()} for c, ps in default.items()}
    for c, ps in (given or {}).items():
        for p, v in ps.items():
            out[c][p] = list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v]
    return out


def _like_tree(modellist, value):
    return {c: {p: [value] * len(_vec(v)) for p, v in ps.items()} for c, ps in modellist.items()}


class NormalPriors:
    """profitMakePriors (R/profitUtility.R:124-176): a normal log density per
    parameter (of log10 of it where tolog), means from the initial model list
    unless given, infinite sigmas flat when allowflat; with its gradient in
    linear parameters."""

    def __init__(self, modellist, sigmas, tolog, means=None, tofit=None, allowflat=False):
        _, self.sigmas = unlist(sigmas)
        _, self.tolog = unlist(tolog, leaf=bool)
        self.means = None if means is None else unlist(means)[1]
        self.tofit = None if tofit is None else [bool(x) for x in unlist(tofit, leaf=lambda x: x)[1]]
        self.allowflat = allowflat
        n = len(unlist(modellist)[1])
        if not all(s > 0 for s in self.sigmas):
            raise ValueError("prior sigmas must be positive")
        if not allowflat and not all(math.isfinite(s) for s in self.sigmas):
            raise ValueError("infinite prior sigmas need allowflat")
        for v in (self.sigmas, self.tolog) + ((self.means,) if self.means else ()):
            if len(v) != n:
                raise ValueError("prior lists must match the model list")

    def _terms(self, new, init):
        parms = unlist(new)[1]
        means = self.means if self.means is not None else unlist(init)[1]
        ps = [p for p in range(len(parms)) if self.tofit is None or self.tofit[p]]
        for p in ps:
            if self.allowflat and math.isinf(self.sigmas[p]):
                continue
            yield p, parms[p], means[p], self.sigmas[p]

    def __call__(self, new, init):
        ll = 0.0
        for p, x, m, s in self._terms(new, init):
            if self.tolog[p]:
                x, m = math.log10(x), math.log10(m)
            ll += -0.5 * ((x - m) / s) ** 2 - math.log(s) - 0.5 * math.log(2 * math.pi)
        return ll

    def grad_flat(self, new, init):
        g = np.zeros(len(unlist(new)[1]))
        for p, x, m, s in self._terms(new, init):
            if self.tolog[p]:
                g[p] = -(math.log10(x) - math.log10(m)) / s ** 2 / (x * LN10)
            else:
                g[p] = -(x - m) / s ** 2
        return g


def make_priors(modellist, sigmas, tolog, means=None, tofit=None, allowflat=False):
    return NormalPriors(modellist, sigmas, tolog, means, tofit, allowflat)


def region_from_segim(segim, mask=None):
    """profitSetupData's default region (R/profitSetupData.R:229-232): the
    segment under the image centre, outside the mask.  Arrays are [y, x]; R's
    segim[ceiling(nx/2), ceiling(ny/2)] is 1-based [x, y]."""
    ny, nx = segim.shape
    keep = segim[math.ceil(ny / 2) - 1, math.ceil(nx / 2) - 1]
    region = segim == keep
    if mask is not None:
        region &= np.asarray(mask) <= 0
    return region


class FitData:
    """profitSetupData's result for one image, evaluated by the evaluator that
    `evaluator(fit_data)` builds (EngineEvaluator by default: the GPU engine
    through the C-ABI).  An evaluator has like(comp_lists) and
    like_grad(comp_lists) over a batch of component lists."""

    def __init__(self, image, modellist, sigma=None, segim=None, mask=None, region=None, psf=None, tofit=None,
                 tolog=None, priors=None, intervals=None, constraints=None, offset=None, rough=False,
                 finesample=1, magzero=0.0, like_func="norm", evaluator=None, device=0, conv_method=None):
        image = np.asarray(image, dtype=np.float64)
        self.image = image
        self.sigma = np.sqrt(np.abs(image)) if sigma is None else np.asarray(sigma, dtype=np.float64)
        self.segim = np.ones_like(image) if segim is None else np.asarray(segim)
        self.mask = np.zeros_like(image) if mask is None else np.asarray(mask)
        self.region = region_from_segim(self.segim, self.mask) if region is None else (np.asarray(region) == True)  # noqa: E712
        self.modellist = {c: {p: _vec(v) for p, v in ps.items()} for c, ps in modellist.items()}
        self.tofit = _merge(_like_tree(self.modellist, True), tofit)
        self.tolog = _merge(_like_tree(self.modellist, False), tolog)
        self.priors = priors
        self.intervals = intervals or {}
        self.constraints = constraints
        self.offset = offset
        self.rough = bool(rough)
        self.finesample = int(finesample)
        if self.finesample != 1:
            raise NotImplementedError("finesampled fits are not handled by this interface")
        self.magzero = float(magzero)
        self.like_func = like_func
        self.psf = None
        if psf is not None:
            p = np.array(psf, dtype=np.float64)
            p[p < 0] = 0  # R/profitSetupData.R:241
            if p.shape[0] % 2 == 0 or p.shape[1] % 2 == 0:
                raise NotImplementedError("even PSF dimensions need profitSetupData's regridding; pass an odd PSF")
            self.psf = p / p.sum()
        self.names, flat = unlist(self.modellist)
        self.flat0 = np.array(flat)
        _, tf = unlist(self.tofit, leaf=lambda x: x)
        _, tl = unlist(self.tolog, leaf=bool)
        self.fit_ids = [i for i, v in enumerate(tf) if v is True or v == 1]
        self.inherit_ids = [i for i, v in enumerate(tf) if v is None]
        self.tolog_flat = np.array([bool(v) for v in tl])
        init = self.flat0.copy()
        init[self.tolog_flat] = np.log10(init[self.tolog_flat])
        self.init = init[self.fit_ids]
        self.parm_names = [self.names[i] for i in self.fit_ids]
        self.N = int(self.region.sum())
        self._build_components()
        # evaluator: a factory taking this FitData (default: the GPU engine)
        if evaluator is None:
            self.evaluator = EngineEvaluator(self, device=device, conv_method=conv_method)
        else:
            self.evaluator = evaluator(self)

    # ---------------------------------------------------------------- model list -> profiles
    def _build_components(self):
        """Components in profitMakeModel's order (R/profitMakeModel.R:38-140):
        the radial families, then point sources, then the sky; the map from
        each flat parameter to the (component, gradient slot) pairs it feeds
        (a length-1 vector feeds every component of its family)."""
        comps, fmap = [], [[] for _ in self.names]
        offsets, k = {}, 0
        for c, ps in self.modellist.items():
            for p, v in ps.items():
                offsets[(c, p)] = (k, len(v))
                k += len(v)
        order = [c for c in RADIAL if c in self.modellist] + \
                [c for c in ("pointsource", "sky") if c in self.modellist]
        unknown = [c for c in self.modellist if c not in order]
        if unknown:
            raise KeyError(f"unsupported model list components: {unknown}")
        haspsf = self.psf is not None
        for c in order:
            ps = self.modellist[c]
            n = max(len(v) for v in ps.values())
            kind = {"pointsource": "psf", "ferrers": "ferrer"}.get(c, c)
            shape = SHAPE_NAMES[PROFILE_TYPES[kind]]
            for j in range(n):
                ci = len(comps)
                comps.append((c, kind, j))
                for p, v in ps.items():
                    base, ln = offsets[(c, p)]
                    i = base + (j if ln > 1 else 0)
                    if p in SLOT and p in ("xcen", "ycen", "mag", "ang", "axrat", "box"):
                        slot = SLOT[p]
                    elif p in shape:
                        slot = SLOT["s%d" % shape.index(p)]
                    else:
                        raise KeyError(f"unknown parameter {p} of {c}")
                    fmap[i].append((ci, slot))
        self.components = comps
        self.flat_map = fmap
        self.convolve = [haspsf and kind not in ("psf", "sky") for _, kind, _ in comps]

    def component_params(self, flat):
        """The per-component parameter dicts (for the evaluator) of a flat
        linear parameter vector."""
        out = []
        for ci, (c, kind, j) in enumerate(self.components):
            d = {}
            for p, v in self.modellist[c].items():
                i = self.names.index(f"{c}.{p}") if len(v) == 1 else self.names.index(f"{c}.{p}{j + 1}")
                d[p] = float(flat[i])
            if kind not in ("psf", "sky"):
                # profitMakeModel's per-profile settings (R/profitMakeModel.R:131-137)
                d.update(rough=self.rough, acc=0.1, rscale_max=0.0)
            out.append((kind, self.convolve[ci], d))
        return out

    # ---------------------------------------------------------------- profitRemakeModellist
    def remake(self, parm):
        """R/profitRemakeModelList.R:26-165 on one parameter vector: returns
        (parm after the intervals, linear flat values, Jacobian d flat / d parm)."""
        parm = np.asarray(parm, dtype=np.float64)
        if parm.shape != (len(self.fit_ids),):
            raise ValueError("Length of parm (i.e. number of parameters being updated) mismatches tofit list!")
        flat = self.flat0.copy()
        nf = len(self.fit_ids)
        J = np.zeros((len(flat), nf))
        if nf == 0:
            return parm.copy(), flat, J
        flat[self.fit_ids] = parm
        J[self.fit_ids, np.arange(nf)] = 1.0
        unlog = [i for i in self.fit_ids if self.tolog_flat[i]]
        flat[unlog] = 10.0 ** flat[unlog]
        J[unlog] *= (flat[unlog] * LN10)[:, None]
        if self.offset is not None:
            off = list(self.offset) + [None] * (4 - len(self.offset))
            xsel = [i for i, n in enumerate(self.names) if "xcen" in n]
            ysel = [i for i, n in enumerate(self.names) if "ycen" in n]
            flat[xsel] += off[0]
            flat[ysel] += off[1]
            if off[2] is not None and not (isinstance(off[2], float) and math.isnan(off[2])):
                asel = [i for i, n in enumerate(self.names) if "ang" in n]
                flat[asel] += off[2]
            if off[3] is not None and not (isinstance(off[3], float) and math.isnan(off[3])):
                pat = re.compile(r"\.re|\.fwhm|\.rb|\.rout|\.rc|\.rt|xcen|ycen")
                ssel = [i for i, n in enumerate(self.names) if pat.search(n)]
                flat[ssel] *= off[3]
                J[ssel] *= off[3]
        for i in self.inherit_ids:
            flat[i] = flat[i - 1]
            J[i] = J[i - 1]
        if self.constraints is not None:
            flat = np.array(unlist(self.constraints(relist(flat, self.modellist)))[1])
        for c, ps in self.intervals.items():
            for p, lims in ps.items():
                key = [n for n in self.names if n == f"{c}.{p}" or re.fullmatch(re.escape(f"{c}.{p}") + r"\d+", n)]
                for k, n in enumerate(key):
                    if k >= len(lims) or lims[k] is None:
                        continue
                    lo, hi = lims[k]
                    i = self.names.index(n)
                    v = max(lo, min(hi, flat[i]))
                    if v != flat[i]:
                        J[i] = 0.0
                    flat[i] = v
        for i in self.inherit_ids:
            flat[i] = flat[i - 1]
            J[i] = J[i - 1]
        mod = flat.copy()
        if self.offset is not None:
            mod[xsel] -= off[0]
            mod[ysel] -= off[1]
            if off[2] is not None and not (isinstance(off[2], float) and math.isnan(off[2])):
                mod[asel] -= off[2]
            if off[3] is not None and not (isinstance(off[3], float) and math.isnan(off[3])):
                mod[ssel] /= off[3]
        mod[unlog] = np.log10(mod[unlog])
        return mod[self.fit_ids], flat, J

    def modellist_of(self, flat):
        return relist(flat, self.modellist)

    # ---------------------------------------------------------------- profitLikeModel
    def _prior_terms(self, parms, flats, grad):
        n = len(flats)
        lp = np.zeros(n)
        g = np.zeros((n, len(self.fit_ids))) if grad else None
        if not self.priors:
            return lp, g
        for b in range(n):
            new = relist(flats[b], self.modellist)
            lp[b] = self.priors(new, self.modellist)
            if not grad:
                continue
            if isinstance(self.priors, NormalPriors):
                _, _, J = self.remake(parms[b])
                g[b] = self.priors.grad_flat(new, self.modellist) @ J
            else:
                for f in range(len(self.fit_ids)):
                    h = 1e-6 * max(1.0, abs(parms[b][f]))
                    pp, pm = parms[b].copy(), parms[b].copy()
                    pp[f] += h
                    pm[f] -= h
                    lpp = self.priors(relist(self.remake(pp)[1], self.modellist), self.modellist)
                    lpm = self.priors(relist(self.remake(pm)[1], self.modellist), self.modellist)
                    g[b, f] = (lpp - lpm) / (2 * h)
        return lp, g

    def like(self, parm):
        """LP = LL + priors for one parameter vector or a batch (rows)."""
        parms = np.atleast_2d(np.asarray(parm, dtype=np.float64))
        rem = [self.remake(p) for p in parms]
        flats = [r[1] for r in rem]
        ll = np.asarray(self.evaluator.like([self.component_params(f) for f in flats]), dtype=np.float64)
        lp, _ = self._prior_terms(parms, flats, False)
        out = ll + lp
        return out if np.ndim(parm) == 2 else float(out[0])

    def prepare(self, parm):
        """profitRemakeModellist on each row of parm: (rows, remakes, the
        evaluator's component lists)."""
        if self.constraints is not None:
            raise NotImplementedError("a constraints function has no derivative")
        parms = np.atleast_2d(np.asarray(parm, dtype=np.float64))
        rem = [self.remake(p) for p in parms]
        return parms, rem, [self.component_params(r[1]) for r in rem]

    def finish_grad(self, parms, rem, ll, g):
        """LP = LL + priors and d LP / d parm from the evaluator's logL and
        per-component slot gradients of prepared rows."""
        ll = np.asarray(ll, dtype=np.float64)
        g = np.asarray(g, dtype=np.float64).reshape(len(parms), len(self.components), -1)
        lp, gp = self._prior_terms(parms, [r[1] for r in rem], True)
        dparm = np.zeros((len(parms), len(self.fit_ids)))
        for b, (_, _, J) in enumerate(rem):
            gflat = np.array([sum(g[b, ci, s] for ci, s in self.flat_map[i]) for i in range(len(self.names))])
            dparm[b] = gflat @ J
        return ll + lp, dparm + gp

    def like_grad(self, parm):
        """(LP, d LP / d parm) for one parameter vector or a batch (rows)."""
        parms, rem, comps = self.prepare(parm)
        ll, g = self.evaluator.like_grad(comps)
        out, dparm = self.finish_grad(parms, rem, ll, g)
        if np.ndim(parm) == 2:
            return out, dparm
        return float(out[0]), dparm[0]

    def optimise(self, parm=None, maxiter=200, method="L-BFGS-B"):
        """Maximises LP with scipy's quasi-Newton methods on the adjoint
        gradient (the role of optim(method='BFGS', fnscale=-1) in the
        reference's examples)."""
        from scipy.optimize import minimize
        x0 = self.init.copy() if parm is None else np.asarray(parm, dtype=np.float64)

        def f(x):
            lp, g = self.like_grad(x)
            if not np.isfinite(lp):
                return 1e300, np.zeros_like(x)
            return -lp, -g
        res = minimize(f, x0, jac=True, method=method, options={"maxiter": maxiter})
        return res


class MultiFit:
    """Several galaxies' FitData evaluated together: one engine holds every
    galaxy and one call evaluates a parameter vector of each as one
    heterogeneous batch (different image sizes, PSFs, likelihoods and
    component counts)."""

    def __init__(self, datas, device=0, conv_method=None, precision=0):
        from . import Engine, adjust_mask
        import profit_amd as P
        self._P = P
        self.datas = list(datas)
        self.eng = Engine(device)
        self.eng.set_render_fp64(precision)
        if conv_method:
            self.eng.set_conv_method(conv_method)
        ids = []
        for d in self.datas:
            h, w = d.image.shape
            calc = None
            if d.psf is not None:
                calc = adjust_mask(d.region.astype(np.uint8), (w, h), d.psf, d.finesample)
            ids.append(self.eng.add_galaxy(w, h, d.image, d.sigma, psf=d.psf, calcregion=calc,
                                           region=d.region.astype(np.uint8), likefunc=d.like_func,
                                           magzero=d.magzero))
        self.eng.set_batch(ids, [len(d.components) for d in self.datas])

    def like_grad(self, parms):
        """LP and d LP / d parm of each galaxy at its own parameter vector."""
        prep = [d.prepare(p) for d, p in zip(self.datas, parms)]
        profiles = [p for (_, _, comps) in prep for p in _profiles(self._P, comps[0])]
        ll, g = self.eng.like_grad(profiles)
        out, row = [], 0
        for i, (d, (rows, rem, _)) in enumerate(zip(self.datas, prep)):
            n = len(d.components)
            lp, dp = d.finish_grad(rows, rem, ll[i:i + 1], g[row:row + n])
            out.append((float(lp[0]), dp[0]))
            row += n
        return out


def read_example_init(path):
    """data/ExampleInit.tab: the example galaxies' initial bulge + disk
    parameters, a whitespace table with a header of quoted column names;
    returns {CATAID: {column: value}}."""
    with open(path) as f:
        header = [h.strip('"') for h in f.readline().split()]
        rows = {}
        for line in f:
            v = line.split()
            if not v:
                continue
            rec = dict(zip(header, (float(x) for x in v)))
            rows[int(rec["CATAID"])] = rec
    return rows


def example_fit(directory, source, cataid, init, evaluator=None, device=0, conv_method=None):
    """The reference's galaxy-fit example (vignettes/ProFit-Galaxy-Fit-Example.Rmd)
    for one example galaxy: bulge + disk from the ExampleInit row, centred
    on the image, sizes scaled by 0.2/0.339 for SDSS, tied centres, logged
    sizes, Sersic index and axial ratio, its intervals and normal priors, the t
    likelihood and magzero 0.  directory holds <source>/G<cataid>{fitim,sigma,
    segim,psfim}.fits."""
    from . import fits
    base = os.path.join(directory, source, f"G{cataid}")
    image = fits.read_image(base + "fitim.fits")
    sigma = fits.read_image(base + "sigma.fits")
    segim = fits.read_image(base + "segim.fits")
    psf = fits.read_image(base + "psfim.fits")
    row = init[int(cataid)]
    ny, nx = image.shape
    rs = 1.0 if source == "KiDS" else 0.2 / 0.339
    modellist = {"sersic": {"xcen": [nx / 2, nx / 2], "ycen": [ny / 2, ny / 2],
                            "mag": [row["sersic.mag1"], row["sersic.mag2"]],
                            "re": [row["sersic.re1"] * rs, row["sersic.re2"] * rs],
                            "nser": [row["sersic.nser1"], 1.0], "ang": [row["sersic.ang2"], row["sersic.ang2"]],
                            "axrat": [1.0, row["sersic.axrat2"]], "box": [0.0, 0.0]}}
    tofit = {"sersic": {"xcen": [True, None], "ycen": [True, None], "mag": [True, True], "re": [True, True],
                        "nser": [True, False], "ang": [False, True], "axrat": [False, True], "box": [False, False]}}
    tolog = {"sersic": {"xcen": [False, False], "ycen": [False, False], "mag": [False, False], "re": [True, True],
                        "nser": [True, True], "ang": [False, False], "axrat": [True, True], "box": [False, False]}}
    s = [2, 2, 5, 1, 1, 30, 0.3, math.inf]
    sigmas = {"sersic": {k: [v, v] for k, v in zip(("xcen", "ycen", "mag", "re", "nser", "ang", "axrat", "box"), s)}}
    priors = make_priors(modellist, sigmas, tolog, allowflat=True)
    intervals = {"sersic": {"xcen": [(0, 300)] * 2, "ycen": [(0, 300)] * 2, "mag": [(10, 30)] * 2,
                            "re": [(1, 100)] * 2, "nser": [(0.5, 20)] * 2, "ang": [(-180, 360)] * 2,
                            "axrat": [(0.1, 1)] * 2, "box": [(-1, 1)] * 2}}
    return FitData(image, modellist, sigma=sigma, segim=segim, psf=psf, tofit=tofit, tolog=tolog, priors=priors,
                   intervals=intervals, magzero=0.0, like_func="t", evaluator=evaluator, device=device,
                   conv_method=conv_method)


def _profiles(module, comp_params):
    return [module.make_profile(kind, convolve=conv, **d) for kind, conv, d in comp_params]


class EngineEvaluator:
    """logL and its gradient of batches of component lists on the GPU engine
    (one galaxy; the batch is resized when a call brings a different count)."""

    def __init__(self, data, device=0, conv_method=None):
        from . import Engine, adjust_mask
        import profit_amd as P
        self._P = P
        h, w = data.image.shape
        self.calcregion = None
        if data.psf is not None:
            self.calcregion = adjust_mask(data.region.astype(np.uint8), (w, h), data.psf, data.finesample)
        self.eng = Engine(device)
        if conv_method:
            self.eng.set_conv_method(conv_method)
        self.gid = self.eng.add_galaxy(w, h, data.image, data.sigma, psf=data.psf, calcregion=self.calcregion,
                                       region=data.region.astype(np.uint8), likefunc=data.like_func,
                                       magzero=data.magzero)
        self.ncomp = len(data.components)
        self.nbatch = 0

    def _batch(self, n):
        if n != self.nbatch:
            self.eng.set_batch([self.gid] * n, [self.ncomp] * n)
            self.nbatch = n

    def like(self, comp_lists):
        self._batch(len(comp_lists))
        return self.eng.like([p for cl in comp_lists for p in _profiles(self._P, cl)])

    def like_grad(self, comp_lists):
        self._batch(len(comp_lists))
        return self.eng.like_grad([p for cl in comp_lists for p in _profiles(self._P, cl)])
