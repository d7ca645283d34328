"""A real galaxy end to end: the KiDS cut-out G267525 that ProFit ships
(tests/golden/kids, read with profit_amd.fits), set up as the reference's
galaxy-fit example does (vignettes/ProFit-Galaxy-Fit-Example.Rmd: bulge +
disk from data/ExampleInit.tab, tied centres, logged sizes, intervals,
profitMakePriors normal priors, the t likelihood, magzero 0).  The GPU
engine's log posterior and its gradient in the fitted parameters are checked
against the same glue over the FP64 oracle, then a quasi-Newton fit on the
adjoint gradient must raise LP."""
import math
import os

import numpy as np
import pytest

from profit_amd import fit as F
from profit_amd import fits
from tests.helpers import OracleEvaluator

pytestmark = pytest.mark.gpu

KIDS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kids")
# data/ExampleInit.tab, CATAID 267525
INIT = dict(mag1=18.89297498916, mag2=18.89297498916, re1=7.00115, re2=14.0023, nser1=2.0818, ang2=157.1785,
            axrat2=0.6997)


def _setup(evaluator=None):
    image = fits.read_image(os.path.join(KIDS, "G267525fitim.fits"))
    sigma = fits.read_image(os.path.join(KIDS, "G267525sigma.fits"))
    segim = fits.read_image(os.path.join(KIDS, "G267525segim.fits"))
    psf = fits.read_image(os.path.join(KIDS, "G267525psfim.fits"))
    ny, nx = image.shape
    modellist = {"sersic": {"xcen": [nx / 2, nx / 2], "ycen": [ny / 2, ny / 2],
                            "mag": [INIT["mag1"], INIT["mag2"]], "re": [INIT["re1"], INIT["re2"]],
                            "nser": [INIT["nser1"], 1.0], "ang": [INIT["ang2"], INIT["ang2"]],
                            "axrat": [1.0, INIT["axrat2"]], "box": [0.0, 0.0]}}
    tofit = {"sersic": {"xcen": [True, None], "ycen": [True, None], "mag": [True, True], "re": [True, True],
                        "nser": [True, False], "ang": [False, True], "axrat": [False, True], "box": [False, False]}}
    tolog = {"sersic": {"xcen": [False, False], "ycen": [False, False], "mag": [False, False], "re": [True, True],
                        "nser": [True, True], "ang": [False, False], "axrat": [True, True], "box": [False, False]}}
    s = [2, 2, 5, 1, 1, 30, 0.3, math.inf]
    sigmas = {"sersic": {k: [v, v] for k, v in zip(("xcen", "ycen", "mag", "re", "nser", "ang", "axrat", "box"), s)}}
    priors = F.make_priors(modellist, sigmas, tolog, allowflat=True)
    intervals = {"sersic": {"xcen": [(0, 300)] * 2, "ycen": [(0, 300)] * 2, "mag": [(10, 30)] * 2,
                            "re": [(1, 100)] * 2, "nser": [(0.5, 20)] * 2, "ang": [(-180, 360)] * 2,
                            "axrat": [(0.1, 1)] * 2, "box": [(-1, 1)] * 2}}
    return F.FitData(image, modellist, sigma=sigma, segim=segim, psf=psf, tofit=tofit, tolog=tolog, priors=priors,
                     intervals=intervals, magzero=0.0, like_func="t", evaluator=evaluator)


@pytest.fixture(scope="module")
def datasets():
    return _setup(), _setup(OracleEvaluator)


def test_posterior_and_gradient_match_the_oracle(datasets):
    gpu, ora = datasets
    assert gpu.N == ora.N > 4000
    rng = np.random.default_rng(3)
    for k in range(3):
        p = gpu.init + (0 if k == 0 else rng.normal(size=len(gpu.init)) * 0.02)
        lp, g = gpu.like_grad(p)
        olp, og = ora.like_grad(p)
        print(f"[kids] point {k}: LP gpu {lp:.6f} oracle {olp:.6f}", flush=True)
        print(f"[kids]   grad gpu    {np.array2string(g, precision=5)}", flush=True)
        print(f"[kids]   grad oracle {np.array2string(og, precision=5)}", flush=True)
        assert abs(lp - olp) <= 1e-3
        assert np.all(np.abs(g - og) <= 2e-3 * np.maximum(np.abs(og), 1e-2 * np.abs(og).max()))


def test_fit_raises_the_posterior(datasets):
    gpu, _ = datasets
    lp0 = gpu.like(gpu.init)
    res = gpu.optimise(maxiter=60)
    lp1 = gpu.like(res.x)
    print(f"[kids] LP {lp0:.3f} -> {lp1:.3f} after {res.nit} iterations ({res.nfev} evaluations): "
          f"{dict(zip(gpu.parm_names, np.round(res.x, 4)))}", flush=True)
    assert np.isfinite(lp1) and lp1 > lp0 + 10
