"""Branches of the model and likelihood path against the FP64 oracle:

* boxy profiles (box != 0: boxy_r, radial.h:93-102, and the boxy Sersic,
  sersic.cpp:164-198) rendered and differentiated, box slot included;
* the broken-exponential gradient (brokenexponential.cpp:38-149) against
  the oracle's fixed-tree central differences;
* the Sersic rescale_flux factor (sersic.cpp:303-311) and its gradient;
* the Poisson likelihood's rescale branch (R/profitLikeModel.R:324-328) on
  both sides of [0.1, 10];
* finesampling (model.cpp:326-329, image.cpp:283-296) through the model path;
* the null convolver (convolve.cpp:521-524) and the FFT convolver's
  uncropped frame (crop = false: model.cpp:307-322, convolve.cpp:81-99,
  262-311) through the model path;
* repeated make_model calls on one device set-up, and clear_cache.

Gates: pixels within 1e-5 relative above 1e-6 of the peak, |dlogL| <= 1e-3,
gradients within 2e-3 of the oracle's fixed-tree gradient (relative to the
slot, or to 5% of the component's largest slot)."""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import fp32_logl_floor, gpu_profiles, oracle_profiles, rel_err

pytestmark = pytest.mark.gpu

SIZE = 60
MAGZERO = 30.0
PSF = W.gaussian_psf(9, 2.0)
GRAD_GATE = 2e-3


def _report(name, err, n):
    print(f"[branches] {name}: max rel pixel err {err:.3e} over {n} px", flush=True)


def _fit(comps, likefunc="chisq", sigma_fn=None, data_fn=None, size=SIZE, psf=PSF, region=None):
    spec = O.ModelSpec(size, size, psf=psf, magzero=MAGZERO, threads=8)
    model = O.model_evaluate(spec, oracle_profiles(comps))["image"]
    draws = W.normal_draws(17, size * size).reshape(size, size)
    sigma = sigma_fn(model) if sigma_fn else np.sqrt(np.maximum(model, 0) + 25.0)
    data = data_fn(model, sigma, draws) if data_fn else model + draws * sigma
    if region is None:
        region = np.ones((size, size), dtype=np.uint8)
    return dict(spec=spec, model=model, data=data, sigma=sigma, region=region,
                fit=O.FitSpec(spec, data, sigma, region, likefunc=likefunc))


def _engine(f, likefunc="chisq", ncomp=1, slots=None, size=SIZE, psf=PSF, precision=0):
    eng = P.Engine()
    eng.set_render_fp64(precision)
    g = eng.add_galaxy(size, size, f["data"], f["sigma"], psf=psf, region=f["region"], likefunc=likefunc,
                       magzero=MAGZERO)
    eng.set_batch([g], [ncomp])
    eng.set_grad_slots(slots)
    return eng


def _check_grad(tag, gpu, ref, slots):
    for k in range(ref.shape[0]):
        o = ref[k, slots]
        gv = gpu[k, slots]
        big = np.abs(o).max()
        print(f"[branches] {tag} comp {k}: gpu {np.array2string(gv, precision=5)} "
              f"oracle {np.array2string(o, precision=5)}", flush=True)
        assert np.all(np.abs(gv - o) <= GRAD_GATE * np.maximum(np.abs(o), 5e-2 * big)), (tag, k)


BOXY = {
    "sersic": dict(re=6.0, nser=2.2),
    "moffat": dict(fwhm=5.0, con=2.5),
}


@pytest.mark.parametrize("box", [0.3, -0.3])
@pytest.mark.parametrize("kind", sorted(BOXY))
def test_boxy_model_image(kind, box):
    comps = [dict(type=kind, xcen=30.4, ycen=29.7, mag=16.0, ang=25.0, axrat=0.65, box=box, **BOXY[kind])]
    spec = O.ModelSpec(SIZE, SIZE, psf=PSF, magzero=MAGZERO, threads=8)
    ref = O.model_evaluate(spec, oracle_profiles(comps))["image"]
    got = P.make_model(SIZE, SIZE, gpu_profiles(comps), psf=PSF, magzero=MAGZERO)["image"]
    err, n = rel_err(got, ref)
    _report(f"{kind} box {box}", err, n)
    assert err <= 1e-5


@pytest.mark.parametrize("box", [0.3, -0.3])
@pytest.mark.parametrize("kind", sorted(BOXY))
def test_boxy_gradient_including_box_slot(kind, box):
    truth = [dict(type=kind, xcen=30.4, ycen=29.7, mag=16.0, ang=25.0, axrat=0.65, box=box, **BOXY[kind])]
    f = _fit(truth)
    comps = [dict(truth[0], xcen=30.55, mag=16.04, ang=27.0, axrat=0.62, box=box * 1.1)]
    comps[0][list(BOXY[kind])[0]] *= 1.03
    eng = _engine(f, slots=None)
    gll, ggrad = eng.like_grad(gpu_profiles(comps))
    oll, ograd = O.loglike_grad(f["fit"], oracle_profiles(comps))
    print(f"[branches] {kind} box {box}: logL gpu {gll[0]:.6f} oracle {oll:.6f}", flush=True)
    assert abs(gll[0] - oll) <= 1e-3
    slots = [0, 1, 2, 3, 4, 5, 6, 7]
    _check_grad(f"{kind} box {box}", ggrad, ograd, slots)
    # the box slot also against fixed-tree central differences of the oracle
    mask = np.zeros((1, O.NGRAD), dtype=np.uint8)
    mask[0, 5] = 1
    _, fd = O.loglike_grad_fd(f["fit"], oracle_profiles(comps), rel_step=1e-5, slots=mask)
    print(f"[branches] {kind} box {box}: d/dbox gpu {ggrad[0, 5]:.6e} fd {fd[0, 5]:.6e}", flush=True)
    assert abs(ggrad[0, 5] - fd[0, 5]) <= 5e-3 * abs(fd[0, 5])


def test_brokenexp_gradient_against_fixed_tree_differences():
    truth = [dict(type="brokenexp", xcen=30.3, ycen=29.6, mag=15.5, ang=35.0, axrat=0.7, box=0.0,
                  h1=5.0, h2=2.5, rb=8.0, a=3.0)]
    f = _fit(truth)
    comps = [dict(truth[0], xcen=30.45, ycen=29.5, mag=15.55, ang=37.0, axrat=0.68, h1=5.2, h2=2.4, rb=8.3, a=2.8)]
    eng = _engine(f)
    gll, ggrad = eng.like_grad(gpu_profiles(comps))
    steps = np.full((1, O.NGRAD), 1e-6)
    steps[:, :2] = 1e-8
    oll, fd = O.loglike_grad_fd(f["fit"], oracle_profiles(comps), rel_step=steps)
    print(f"[branches] brokenexp: logL gpu {gll[0]:.6f} oracle {oll:.6f}", flush=True)
    assert abs(gll[0] - oll) <= 1e-3
    _check_grad("brokenexp", ggrad, fd, [0, 1, 2, 3, 4, 6, 7, 8, 9])


@pytest.mark.parametrize("nser", [1.7, 4.5])
def test_sersic_rescale_flux_gradient(nser):
    # a profile truncated at rscale_max = 3 re, whose flux the factor
    # 1 / P(b_n (rscale_max / re)^(1/n), 2n) restores (a sizeable correction
    # here, unlike at the adjusted 99.99% radius).  The source is bright (peak
    # ~5e4 counts) and its FP32-canvas floor (4e-3 in logL) exceeds the gate:
    # the model runs in the FP64 mode (FP64 values, canvas and FFT).
    truth = [dict(type="sersic", xcen=30.3, ycen=29.6, mag=15.5, ang=35.0, axrat=0.7, box=0.0, re=4.0, nser=nser,
                  rescale_flux=1, rscale_max=3.0)]
    f = _fit(truth)
    comps = [dict(truth[0], xcen=30.4, mag=15.53, re=4.15, nser=nser * 0.97, axrat=0.68)]
    eng = _engine(f, precision=2)
    gll, ggrad = eng.like_grad(gpu_profiles(comps))
    oll = O.loglike(f["fit"], oracle_profiles(comps))
    assert abs(gll[0] - oll) <= 1e-3
    steps = np.full((1, O.NGRAD), 1e-6)
    steps[:, :2] = 1e-8
    _, fd = O.loglike_grad_fd(f["fit"], oracle_profiles(comps), rel_step=steps)
    print(f"[branches] rescale_flux n {nser}: logL gpu {gll[0]:.6f} oracle {oll:.6f}", flush=True)
    _check_grad(f"rescale_flux n {nser}", ggrad, fd, [0, 1, 2, 3, 4, 6, 7])
    # the rescale factor moves d/dre and d/dnser: the gradient without it differs
    _, fd0 = O.loglike_grad_fd(f["fit"], oracle_profiles([dict(comps[0], rescale_flux=0)]), rel_step=steps)
    assert abs(fd0[0, 7] - fd[0, 7]) > 10 * GRAD_GATE * abs(fd[0, 7])


# Poisson data whose scale max |d| / sigma^2 lies above 10 (about 16) and
# below 0.1 (about 0.04): R/profitLikeModel.R:324-328 then multiplies model
# and data by it.  The scaled sums stay near 1e4, where the FP32 model image
# keeps logL within 1e-3.
POIS = {
    "above_10": dict(mag=22.6, sigma=lambda m: np.full_like(m, 1.0)),
    "below_0.1": dict(mag=19.0, sigma=lambda m: np.full_like(m, 100.0)),
}


@pytest.mark.parametrize("case", sorted(POIS))
def test_poisson_rescale_branch(case):
    c = POIS[case]
    truth = [dict(type="sersic", xcen=30.3, ycen=29.6, mag=c["mag"], ang=35.0, axrat=0.7, box=0.0, re=6.0,
                  nser=1.5)]
    # the region stays inside the profile's support (log m is finite there)
    f = _fit(truth, likefunc="pois", sigma_fn=c["sigma"],
             data_fn=lambda m, s, z: np.maximum(m * (1 + 0.05 * z), 1e-3 * m.max()),
             region=W.ellipse_mask(SIZE, SIZE, 30.0, 30.0, 22.0, 20.0))
    scale = np.max(np.abs(f["data"]) / f["sigma"] ** 2)
    assert scale > 10 or scale < 0.1, scale
    comps = [dict(truth[0], xcen=30.4, mag=c["mag"] + 0.03, re=6.2, nser=1.45, axrat=0.68)]
    eng = _engine(f, likefunc="pois")
    gll, ggrad = eng.like_grad(gpu_profiles(comps))
    oll, ograd = O.loglike_grad(f["fit"], oracle_profiles(comps))
    _, om = O.loglike(f["fit"], oracle_profiles(comps), want_model=True)
    floor = fp32_logl_floor("pois", f["data"], f["sigma"], om, f["region"])
    print(f"[branches] poisson {case} (scale {scale:.3g}): logL gpu {gll[0]:.6f} oracle {oll:.6f} "
          f"(FP32 floor {floor:.2e})", flush=True)
    assert abs(gll[0] - oll) <= 1e-3
    _check_grad(f"poisson {case}", ggrad, ograd, [0, 1, 2, 3, 4, 6, 7])


@pytest.mark.parametrize("return_finesampled", [False, True])
def test_finesampling_through_the_model(return_finesampled):
    fs = 3
    comps = [dict(type="sersic", xcen=20.3, ycen=19.6, mag=16.0, ang=35.0, axrat=0.7, box=0.0, re=3.0, nser=2.5),
             dict(type="moffat", xcen=22.1, ycen=18.4, mag=17.0, ang=80.0, axrat=0.9, box=0.0, fwhm=2.5, con=3.0)]
    psf = W.gaussian_psf(15, 4.0)  # on the fine grid
    spec = O.ModelSpec(40, 40, psf=psf, magzero=MAGZERO, finesampling=fs, return_finesampled=return_finesampled,
                       threads=8)
    ref = O.model_evaluate(spec, oracle_profiles(comps))
    got = P.make_model(40, 40, gpu_profiles(comps), psf=psf, magzero=MAGZERO, finesampling=fs,
                       return_finesampled=return_finesampled)
    assert got["image"].shape == ref["image"].shape == ((120, 120) if return_finesampled else (40, 40))
    assert got["offset"] == ref["offset"]
    err, n = rel_err(got["image"], ref["image"])
    _report(f"finesampling {fs} return_finesampled {return_finesampled}", err, n)
    assert err <= 1e-5


COMPS_MODEL = [dict(type="sersic", xcen=30.4, ycen=27.7, mag=16.0, ang=25.0, axrat=0.65, box=0.0, re=5.0, nser=2.0),
               dict(type="moffat", xcen=25.1, ycen=31.4, mag=17.0, ang=80.0, axrat=0.9, box=0.0, fwhm=3.0, con=2.5,
                    convolve=False)]


def _comps_with_flags(comps):
    """Oracle and product profiles with each component's own convolve flag."""
    o, g = [], []
    for c in comps:
        c = dict(c)
        conv = c.pop("convolve", True)
        o += oracle_profiles([c], convolve=conv)
        g += gpu_profiles([c], convolve=conv)
    return o, g


def test_null_convolver_returns_the_unconvolved_canvas():
    o, g = _comps_with_flags(COMPS_MODEL)
    ref = O.model_evaluate(O.ModelSpec(SIZE, SIZE, psf=PSF, magzero=MAGZERO, convolver="null", threads=8), o)
    got = P.make_model(SIZE, SIZE, g, psf=PSF, magzero=MAGZERO, convolver="null")
    conv = P.make_model(SIZE, SIZE, g, psf=PSF, magzero=MAGZERO, convolver="brute")
    err, n = rel_err(got["image"], ref["image"])
    _report("null convolver", err, n)
    assert got["image"].shape == ref["image"].shape and got["offset"] == ref["offset"]
    assert err <= 1e-5
    # the brute-force convolver does convolve
    assert rel_err(conv["image"], ref["image"])[0] > 1e-2


@pytest.mark.parametrize("with_mask", [False, True])
@pytest.mark.parametrize("dims", [(60, 60), (64, 52)])
def test_fft_uncropped_frame(dims, with_mask):
    w, h = dims
    o, g = _comps_with_flags(COMPS_MODEL)
    mask = None
    if with_mask:
        mask = W.ellipse_mask(w, h, w / 2, h / 2, 0.45 * w, 0.35 * h)
    spec = O.ModelSpec(w, h, psf=PSF, mask=mask, magzero=MAGZERO, convolver="fft", crop=False, threads=8)
    ref = O.model_evaluate(spec, o)
    got = P.make_model(w, h, g, psf=PSF, mask=mask, magzero=MAGZERO, convolver="fft", crop=False)
    print(f"[branches] fft crop=false {dims} mask {with_mask}: shape {got['image'].shape} vs {ref['image'].shape}, "
          f"offset {got['offset']} vs {ref['offset']}", flush=True)
    assert got["image"].shape == ref["image"].shape
    assert got["offset"] == ref["offset"]
    err, n = rel_err(got["image"], ref["image"])
    _report(f"fft crop=false {dims} mask {with_mask}", err, n)
    assert err <= 1e-5
    # crop = false with the brute-force convolver returns the cropped image
    cropped = P.make_model(w, h, g, psf=PSF, mask=mask, magzero=MAGZERO, convolver="brute", crop=False)
    assert cropped["image"].shape == (h, w) and cropped["offset"] == (0, 0)


def test_repeated_make_model_reuses_the_set_up():
    spec = O.ModelSpec(SIZE, SIZE, psf=PSF, magzero=MAGZERO, threads=8)
    for step, dx in enumerate([0.0, 0.37, -0.21]):
        comps = [dict(c, xcen=c["xcen"] + dx, mag=c["mag"] + 0.1 * step) for c in COMPS_MODEL]
        o, g = _comps_with_flags(comps)
        ref = O.model_evaluate(spec, o)["image"]
        got = P.make_model(SIZE, SIZE, g, psf=PSF, magzero=MAGZERO)["image"]
        err, n = rel_err(got, ref)
        _report(f"repeated make_model call {step}", err, n)
        assert err <= 1e-5
        if step == 1:
            P.clear_cache()
