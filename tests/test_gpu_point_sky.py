"""Point sources ("psf", psf.cpp:58-145) and sky (sky.cpp:47-63) on the GPU,
through the C-ABI, against the FP64 oracle: model images (per-pixel relative
error <= 1e-5 above 1e-6 of the peak) through Model::evaluate, and logL
(<= 1e-3) and the adjoint gradient of the batched engine, where the point
source and the sky are drawn unconvolved next to a convolved Sersic."""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

MAGZERO = 30.0


def _pair(kind, convolve=False, **kw):
    return O.make_profile(kind, convolve=convolve, **kw), P.make_profile(kind, convolve=convolve, **kw)


def _scene(sky_convolved=False):
    return [_pair("sersic", True, xcen=30.4, ycen=29.7, mag=17.0, re=4.0, nser=2.0, ang=20.0, axrat=0.7),
            _pair("psf", xcen=37.3, ycen=24.6, mag=17.5),
            _pair("psf", xcen=20.55, ycen=38.2, mag=18.2),
            _pair("sky", sky_convolved, bg=3.5)]


@pytest.mark.parametrize("convolver", ["brute", "fft"])
@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("psf_scale", [1.0, 0.8])
def test_model_images_with_point_sources_and_sky(convolver, masked, psf_scale):
    size = 60
    psf = W.moffat_psf(15, 3.0, 2.5)
    mask = W.ellipse_mask(size, size, 30.0, 30.0, 26.0, 22.0) if masked else None
    pairs = _scene()
    spec = O.ModelSpec(size, size, psf=psf, mask=mask, magzero=MAGZERO, psf_scale=(psf_scale, psf_scale),
                       convolver="fft" if convolver == "fft" else "brute", threads=8)
    ref = O.model_evaluate(spec, [o for o, _ in pairs])["image"]
    out = P.make_model(size, size, [g for _, g in pairs], psf=psf, mask=mask, magzero=MAGZERO,
                       psf_scale=(psf_scale, psf_scale), convolver=convolver)["image"]
    err, n = rel_err(out, ref)
    print(f"[point] {convolver} masked={masked} psf_scale={psf_scale}: pixel err {err:.2e} over {n} px", flush=True)
    assert err <= 1e-5


def test_convolved_sky_and_unconvolved_model():
    size = 40
    psf = W.gaussian_psf(9, 2.0)
    pairs = [_pair("sky", True, bg=2.0), _pair("psf", xcen=19.3, ycen=21.8, mag=16.0)]
    spec = O.ModelSpec(size, size, psf=psf, magzero=MAGZERO)
    ref = O.model_evaluate(spec, [o for o, _ in pairs])["image"]
    out = P.make_model(size, size, [g for _, g in pairs], psf=psf, magzero=MAGZERO)["image"]
    err, _ = rel_err(out, ref)
    assert err <= 1e-5
    # only point sources: the PSF draws them but nothing is convolved, so no padding
    only = [_pair("psf", xcen=19.3, ycen=21.8, mag=16.0)]
    ref = O.model_evaluate(spec, [o for o, _ in only])
    out = P.make_model(size, size, [g for _, g in only], psf=psf, magzero=MAGZERO)
    assert ref["padding"] == (0, 0)
    assert rel_err(out["image"], ref["image"])[0] <= 1e-5


def test_point_source_without_psf_fails_like_libprofit():
    with pytest.raises(P.ProfitAmdError, match="No psf present in the model, cannot produce a psf profile"):
        P.make_model(20, 20, [P.make_profile("psf", xcen=10.0, ycen=10.0, mag=15.0)])


@pytest.mark.parametrize("method", ["brute", "fft"])
def test_engine_likelihood_and_gradient(method):
    size = 60
    psf = W.moffat_psf(15, 3.0, 2.5)
    truth = _scene()
    spec = O.ModelSpec(size, size, psf=psf, magzero=MAGZERO, threads=8)
    model = O.model_evaluate(spec, [o for o, _ in truth])["image"]
    sigma = np.sqrt(np.maximum(model, 0) + 25.0)
    data = model + W.normal_draws(21, size * size).reshape(size, size) * sigma
    region = W.ellipse_mask(size, size, 30.0, 30.0, 27.0, 24.0)
    calc = O.adjust_mask(region, (size, size), psf)
    cspec = O.ModelSpec(size, size, psf=psf, mask=calc, adjust_mask=False, magzero=MAGZERO, threads=8)
    fit = O.FitSpec(cspec, data, sigma, region)
    eng = P.Engine()
    eng.set_conv_method(method)
    gid = eng.add_galaxy(size, size, data, sigma, psf=psf, calcregion=calc, region=region, magzero=MAGZERO)
    pert = [_pair("sersic", True, xcen=30.5, ycen=29.6, mag=17.05, re=4.1, nser=1.95, ang=21.0, axrat=0.71),
            _pair("psf", xcen=37.45, ycen=24.5, mag=17.45),
            _pair("psf", xcen=20.4, ycen=38.35, mag=18.3),
            _pair("sky", bg=3.3)]
    sets = [truth, pert]
    eng.set_batch([gid] * 2, [len(truth)] * 2)
    ll, grad = eng.like_grad([g for s in sets for _, g in s])
    grad = grad.reshape(2, len(truth), -1)
    for i, s in enumerate(sets):
        oll, og = O.loglike_grad(fit, [o for o, _ in s])
        print(f"[point] {method} set {i}: logL gpu {ll[i]:.6f} oracle {oll:.6f}", flush=True)
        assert abs(ll[i] - oll) <= 1e-3
        checks = [(0, [0, 1, 2, 3, 4, 6, 7]), (1, [0, 1, 2]), (2, [0, 1, 2]), (3, [6])]
        big = max(np.abs(og[k, sl]).max() for k, sl in checks)
        for k, slots in checks:
            o, gv = og[k, slots], grad[i, k, slots]
            print(f"[point]   comp {k}: gpu {np.array2string(gv, precision=5)} "
                  f"oracle {np.array2string(o, precision=5)}", flush=True)
            assert np.all(np.abs(gv - o) <= 2e-3 * np.maximum(np.abs(o), 1e-3 * big)), (i, k)
