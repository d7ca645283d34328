"""ProFit's fitting glue (profit_amd.fit) on the CPU with the oracle as the
evaluator: R's unlist naming and order, profitSetupData's init and default
region, profitRemakeModellist's ties, logs, offsets and intervals, the normal
priors of profitMakePriors, and the chain rule of the log-posterior gradient
against central differences of LP in the fitted parameters."""
import math

import numpy as np
import pytest
import scipy.stats as st

import oracle as O
from profit_amd import fit as F
from profit_amd import workloads as W
from tests.helpers import OracleEvaluator

MODEL = {"sersic": {"xcen": [20.3, 20.3], "ycen": [19.6, 19.6], "mag": [16.5, 16.0], "re": [2.5, 6.0],
                    "nser": [3.0, 1.0], "ang": [30.0, 30.0], "axrat": [1.0, 0.5], "box": [0.0, 0.0]},
         "pointsource": {"xcen": [27.4], "ycen": [12.3], "mag": [18.0]},
         "sky": {"bg": [2.0]}}
TOFIT = {"sersic": {"xcen": [True, None], "ycen": [True, None], "mag": [True, True], "re": [True, True],
                    "nser": [True, False], "ang": [False, True], "axrat": [False, True], "box": [False, False]},
         "pointsource": {"xcen": [True], "ycen": [True], "mag": [True]},
         "sky": {"bg": [True]}}
TOLOG = {"sersic": {"re": [True, True], "nser": [True, True], "axrat": [True, True]}}
INTERVALS = {"sersic": {"re": [(1, 100), (1, 100)], "axrat": [(0.1, 1), (0.1, 1)], "nser": [(0.5, 20), (0.5, 20)]}}


def _data(like_func="norm", rough=True, priors=None, intervals=INTERVALS, offset=None):
    size = 40
    psf = W.gaussian_psf(9, 2.0)
    spec = O.ModelSpec(size, size, psf=psf, magzero=30.0)
    # truth: the model list itself, drawn with the oracle
    d0 = F.FitData(np.zeros((size, size)), MODEL, sigma=np.ones((size, size)), psf=psf, tofit=TOFIT, tolog=TOLOG,
                   magzero=30.0, rough=rough, evaluator=lambda d: None)
    model = O.model_evaluate(spec, [O.make_profile(k, convolve=c, **p)
                                    for k, c, p in d0.component_params(d0.flat0)])["image"]
    sigma = np.sqrt(np.maximum(model, 0) + 25.0)
    image = model + W.normal_draws(17, size * size).reshape(size, size) * sigma * (1.5 if like_func == "t" else 1)
    segim = np.ones((size, size))
    segim[:3, :] = 2  # a neighbouring segment the default region excludes
    return F.FitData(image, MODEL, sigma=sigma, segim=segim, psf=psf, tofit=TOFIT, tolog=TOLOG, priors=priors,
                     intervals=intervals, offset=offset, magzero=30.0, like_func=like_func, rough=rough,
                     evaluator=OracleEvaluator)


def test_unlist_names_and_order():
    names, vals = F.unlist(MODEL)
    assert names[:4] == ["sersic.xcen1", "sersic.xcen2", "sersic.ycen1", "sersic.ycen2"]
    assert names[-4:] == ["pointsource.xcen", "pointsource.ycen", "pointsource.mag", "sky.bg"]
    assert vals[4:6] == [16.5, 16.0]
    assert F.relist(vals, MODEL) == {c: {p: [float(x) for x in v] for p, v in ps.items()} for c, ps in MODEL.items()}


def test_init_default_region_and_components():
    d = _data()
    # R: init = unlist(modellist); init[tolog] = log10(init[tolog]); init[which(tofit)]
    assert d.parm_names == ["sersic.xcen1", "sersic.ycen1", "sersic.mag1", "sersic.mag2", "sersic.re1", "sersic.re2",
                            "sersic.nser1", "sersic.ang2", "sersic.axrat2", "pointsource.xcen", "pointsource.ycen",
                            "pointsource.mag", "sky.bg"]
    assert np.allclose(d.init, [20.3, 19.6, 16.5, 16.0, math.log10(2.5), math.log10(6.0), math.log10(3.0), 30.0,
                                math.log10(0.5), 27.4, 12.3, 18.0, 2.0])
    assert d.region.sum() == 40 * 37 and not d.region[:3].any()
    assert [c[1] for c in d.components] == ["sersic", "sersic", "psf", "sky"]
    assert d.convolve == [True, True, False, False]


def test_remake_ties_logs_and_intervals():
    d = _data()
    p = d.init.copy()
    p[0] = 21.0                      # xcen1, tied to xcen2
    p[5] = math.log10(150.0)         # re2 beyond its interval
    p[8] = math.log10(0.3)           # axrat2
    pout, flat, J = d.remake(p)
    ml = d.modellist_of(flat)
    assert ml["sersic"]["xcen"] == [21.0, 21.0]
    assert ml["sersic"]["re"][1] == 100.0 and abs(ml["sersic"]["axrat"][1] - 0.3) < 1e-15
    assert abs(pout[5] - 2.0) < 1e-15          # parm comes back clamped (log10 100)
    # Jacobian: the tie feeds both xcen entries, the clamp cuts re2, logs scale by theta ln 10
    i1, i2 = d.names.index("sersic.xcen1"), d.names.index("sersic.xcen2")
    assert J[i1, 0] == 1 and J[i2, 0] == 1
    assert np.all(J[d.names.index("sersic.re2")] == 0)
    ia = d.names.index("sersic.axrat2")
    assert abs(J[ia, 8] - 0.3 * math.log(10)) < 1e-12


def test_offsets_follow_remake():
    d = _data(offset=[1.5, -2.0, 10.0, 2.0])
    _, flat, J = d.remake(d.init)
    ml = d.modellist_of(flat)
    # xcen: (x + 1.5) * 2 (sizes include xcen, ycen); re scaled; ang shifted
    assert np.allclose(ml["sersic"]["xcen"], [(20.3 + 1.5) * 2] * 2)
    assert np.allclose(ml["sersic"]["re"], [5.0, 12.0])
    assert np.allclose(ml["sersic"]["ang"], [40.0, 40.0])
    assert J[d.names.index("sersic.xcen1"), 0] == 2.0


def test_normal_priors_and_gradient():
    sig = {c: {p: [0.5 + 0.1 * k for k in range(len(v))] for p, v in ps.items()} for c, ps in MODEL.items()}
    tl = F._merge(F._like_tree(MODEL, False), TOLOG)
    pri = F.make_priors(MODEL, sig, tl)
    d = _data(priors=pri)
    new = d.modellist_of(d.remake(d.init + 0.01)[1])
    names, x = F.unlist(new)
    _, m = F.unlist(MODEL)
    _, s = F.unlist(sig)
    _, lg = F.unlist(tl, leaf=bool)
    ref = sum(st.norm.logpdf(math.log10(a) if l else a, math.log10(b) if l else b, c)
              for a, b, c, l in zip(x, m, s, lg))
    assert abs(pri(new, MODEL) - ref) < 1e-10
    g = pri.grad_flat(new, MODEL)
    for i in range(len(x)):
        h = 1e-6 * max(1.0, abs(x[i]))
        xp, xm = list(x), list(x)
        xp[i] += h
        xm[i] -= h
        fd = (pri(F.relist(xp, MODEL), MODEL) - pri(F.relist(xm, MODEL), MODEL)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), names[i]


@pytest.mark.parametrize("like_func", ["norm", "chisq", "t"])
def test_posterior_gradient_matches_central_differences(like_func):
    """rough = True (no sub-pixel recursion) makes LP smooth in every
    parameter, so central differences of LP through the whole chain (remake,
    oracle likelihood, priors) check the chain rule to ~1e-6."""
    sig = {c: {p: [2.0] * len(v) for p, v in ps.items()} for c, ps in MODEL.items()}
    pri = F.make_priors(MODEL, sig, F._merge(F._like_tree(MODEL, False), TOLOG))
    d = _data(like_func=like_func, priors=pri)
    p = d.init + np.linspace(-0.02, 0.03, len(d.init))
    lp, g = d.like_grad(p)
    assert abs(lp - d.like(p)) < 1e-9 * abs(lp)
    for f in range(len(p)):
        h = 1e-6 * max(1.0, abs(p[f]))
        pp, pm = p.copy(), p.copy()
        pp[f] += h
        pm[f] -= h
        fd = (d.like(pp) - d.like(pm)) / (2 * h)
        assert abs(g[f] - fd) <= 2e-5 * max(abs(fd), 1e-2 * np.abs(g).max()), (like_func, d.parm_names[f], g[f], fd)


def test_clamped_parameter_has_zero_gradient_and_batches_agree():
    d = _data()
    p = d.init.copy()
    p[5] = math.log10(150.0)   # re2 clamped at 100
    lp, g = d.like_grad(p)
    assert g[5] == 0.0
    batch = np.stack([p, d.init])
    lpb, gb = d.like_grad(batch)
    # the oracle's threaded reductions may round differently from call to call
    assert abs(lpb[0] - lp) <= 1e-12 * abs(lp) and np.allclose(gb[0], g, rtol=1e-10, atol=0)
    assert abs(lpb[1] - d.like(d.init)) <= 1e-12 * abs(lpb[1])


def test_black_box_priors_use_differences():
    def pri(new, init):
        return -0.5 * ((new["sky"]["bg"][0] - 2.0) / 0.1) ** 2 - (new["sersic"]["re"][0] - 2.5) ** 2
    d = _data(priors=pri)
    d0 = _data()
    p = d.init + 0.01
    _, g = d.like_grad(p)
    _, g0 = d0.like_grad(p)
    bg = d.parm_names.index("sky.bg")
    assert abs((g - g0)[bg] - (-(p[bg] - 2.0) / 0.01)) < 1e-4
    r1 = d.parm_names.index("sersic.re1")
    re1 = 10 ** p[r1]
    assert abs((g - g0)[r1] - (-2 * (re1 - 2.5) * re1 * math.log(10))) < 1e-4
