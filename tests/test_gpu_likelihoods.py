"""The four likelihoods of R/profitLikeModel.R:244-329 (norm, chisq, t, pois)
on the batched engine, through the C-ABI, against the FP64 oracle: logL
within 1e-3 and the adjoint gradient against the oracle's analytic gradient
(the same fixed recursion tree in FP64).  The t likelihood's degrees of
freedom follow the sample variance of the residuals; its data carry 1.5 sigma
noise so that dof = 2v / (v - 1) stays unclamped and its dependence on every
pixel enters the gradient."""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import gpu_profiles, oracle_profiles

pytestmark = pytest.mark.gpu

SIZE = 100
MAGZERO = 30.0
COMPS = [dict(type="sersic", xcen=50.3, ycen=49.7, mag=15.0, re=8.0, nser=2.5, ang=30.0, axrat=0.6, box=0.0),
         dict(type="moffat", xcen=49.6, ycen=50.8, mag=16.5, fwhm=5.0, con=2.5, ang=75.0, axrat=0.8, box=0.0)]


def _perturbed():
    out = [dict(c) for c in COMPS]
    out[0].update(xcen=50.45, mag=15.03, re=8.2, nser=2.4, axrat=0.62)
    out[1].update(ycen=50.6, mag=16.45, fwhm=5.1, con=2.6)
    return out


@pytest.fixture(scope="module")
def galaxies():
    psf = W.gaussian_psf(25, 3.0)
    spec = O.ModelSpec(SIZE, SIZE, psf=psf, magzero=MAGZERO, threads=8)
    model = O.model_evaluate(spec, oracle_profiles(COMPS))["image"]
    region = W.ellipse_mask(SIZE, SIZE, 50.0, 50.0, 48.0, 40.0)
    calc = O.adjust_mask(region, (SIZE, SIZE), psf)
    out = {}
    draws = W.normal_draws(11, SIZE * SIZE).reshape(SIZE, SIZE)
    for lf in ("norm", "chisq", "t", "pois"):
        if lf == "pois":
            sigma = np.sqrt(np.maximum(model, 1e-3))
            data = np.maximum(model + draws * sigma, 0.5)
        else:
            sigma = np.sqrt(np.maximum(model, 0) + 25.0)
            data = model + (1.5 if lf == "t" else 1.0) * draws * sigma
        cspec = O.ModelSpec(SIZE, SIZE, psf=psf, mask=calc, adjust_mask=False, magzero=MAGZERO, threads=8)
        out[lf] = dict(psf=psf, data=data, sigma=sigma, region=region, calc=calc,
                       fit=O.FitSpec(cspec, data, sigma, region, likefunc=lf))
    return out


@pytest.mark.parametrize("method", ["brute", "fft"])
@pytest.mark.parametrize("likefunc", ["norm", "chisq", "t", "pois"])
def test_likelihood_and_gradient(galaxies, likefunc, method):
    g = galaxies[likefunc]
    eng = P.Engine()
    eng.set_conv_method(method)
    gid = eng.add_galaxy(SIZE, SIZE, g["data"], g["sigma"], psf=g["psf"], calcregion=g["calc"], region=g["region"],
                         likefunc=likefunc, magzero=MAGZERO)
    sets = [COMPS, _perturbed()]
    eng.set_batch([gid] * len(sets), [len(COMPS)] * len(sets))
    params = [p for s in sets for p in gpu_profiles(s)]
    ll_only = eng.like(params)
    ll, grad = eng.like_grad(params)
    grad = grad.reshape(len(sets), len(COMPS), -1)
    for i, s in enumerate(sets):
        oll, ograd = O.loglike_grad(g["fit"], oracle_profiles(s))
        print(f"[like] {likefunc} {method} set {i}: logL gpu {ll[i]:.6f} (like() {ll_only[i]:.6f}) oracle {oll:.6f}",
              flush=True)
        assert abs(ll[i] - oll) <= 1e-3
        assert ll_only[i] == ll[i]
        for k, c in enumerate(s):
            nshape = 2
            slots = [0, 1, 2, 3, 4] + [6 + j for j in range(nshape)]
            o = ograd[k, slots]
            gv = grad[i, k, slots]
            big = np.abs(o).max()
            print(f"[like]   comp {k} {c['type']}: gpu {np.array2string(gv, precision=5)} "
                  f"oracle {np.array2string(o, precision=5)}", flush=True)
            assert np.all(np.abs(gv - o) <= 2e-3 * np.maximum(np.abs(o), 5e-2 * big)), (i, k)


def test_t_data_keep_dof_unclamped(galaxies):
    """The t case above runs where dof = 2v / (v - 1) lies between 3 and 6, so
    the term of the gradient that carries the dof's dependence on the
    residuals is active (at v < 1 dof is clamped at 1 and the term vanishes)."""
    g = galaxies["t"]
    z = (g["data"] - O.model_evaluate(g["fit"].spec, oracle_profiles(COMPS))["image"]) / g["sigma"]
    v = np.var(z[g["region"] == 1], ddof=1)
    assert 1.5 < v < 3.0  # dof = 2v / (v - 1) between 3 and 6: unclamped


def test_gradient_slot_mask(galaxies):
    """Slots left out of the mask come back as 0 and the others are unchanged
    (leaving out the box slot skips its per-sample logarithms), up to the
    grouping of the adjoint walker's FP32 partial sums, which follows the
    sub-pixel queue's order (appended by atomics): 1e-6 relative."""
    g = galaxies["chisq"]
    eng = P.Engine()
    gid = eng.add_galaxy(SIZE, SIZE, g["data"], g["sigma"], psf=g["psf"], calcregion=g["calc"], region=g["region"],
                         magzero=MAGZERO)
    eng.set_batch([gid], [len(COMPS)])
    params = gpu_profiles(_perturbed())
    ll, full = eng.like_grad(params)
    eng.set_grad_slots(["xcen", "ycen", "mag", "ang", "axrat", "s0", "s1"])
    ll2, part = eng.like_grad(params)
    assert abs(ll2[0] - ll[0]) <= 1e-9 * abs(ll[0])
    keep = [0, 1, 2, 3, 4, 6, 7]
    assert np.all(part[:, [5, 8, 9, 10]] == 0)
    assert np.allclose(part[:, keep], full[:, keep], rtol=1e-6, atol=0)
    eng.set_grad_slots(None)
    _, again = eng.like_grad(params)
    assert np.allclose(again, full, rtol=1e-6, atol=0)
