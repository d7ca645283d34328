"""GPU path vs the FP64 oracle on the same inputs, through the C-ABI.

Gates: pixel classes and recursion trees bit-exact; per-pixel relative error
<= 1e-5 above 1e-6 of the peak; |logL difference| <= 1e-3.
"""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import gpu_profiles, oracle_model_fn, oracle_profiles, rel_err

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _report(name, err, n):
    print(f"[parity] {name}: max rel err {err:.3e} over {n} pixels", flush=True)


def test_config1_model_image():
    wl = W.config1()
    oimg = O.model_evaluate(O.ModelSpec(100, 100, psf=wl["psf"], threads=8),
                            oracle_profiles(wl["components"]))["image"]
    gimg = P.make_model(100, 100, gpu_profiles(wl["components"]), psf=wl["psf"])["image"]
    assert gimg.shape == oimg.shape == (100, 100)
    err, n = rel_err(gimg, oimg)
    _report("config1 model", err, n)
    assert err <= REL_TOL


def test_config1_classes_and_recursion_trees_bit_exact():
    wl = W.config1(oracle_model_fn())
    ref = O.model_evaluate(O.ModelSpec(100, 100, psf=wl["psf"], threads=8),
                           oracle_profiles(wl["components"]), diag=True)
    eng = P.Engine()
    g = eng.add_galaxy(100, 100, wl["data"], wl["sigma"], psf=wl["psf"])
    eng.set_batch([g], [1])
    cls, nodes = eng.diagnostics(gpu_profiles(wl["components"]))
    assert cls.shape == ref["cls"].shape
    assert np.array_equal(cls, ref["cls"].astype(np.int32))
    assert np.array_equal(nodes, ref["nodes"].astype(np.int32))
    print(f"[parity] config1 classes: {np.bincount(cls.ravel(), minlength=4)}, nodes {nodes.sum()}", flush=True)


def test_config1_likelihood():
    wl = W.config1(oracle_model_fn())
    fit = O.FitSpec(O.ModelSpec(100, 100, psf=wl["psf"], threads=8), wl["data"], wl["sigma"], wl["region"])
    oll = O.loglike(fit, oracle_profiles(wl["components"]))
    eng = P.Engine()
    g = eng.add_galaxy(100, 100, wl["data"], wl["sigma"], psf=wl["psf"], region=wl["region"])
    eng.set_batch([g], [1])
    gll = eng.like(gpu_profiles(wl["components"]))[0]
    print(f"[parity] config1 logL gpu {gll:.9f} oracle {oll:.9f} diff {gll - oll:.3e}", flush=True)
    assert abs(gll - oll) <= 1e-3


def _config2_fit(magzero=30.0):
    wl = W.config2(oracle_model_fn(), magzero=magzero)
    calc = O.adjust_mask(wl["region"], (400, 400), wl["psf"])
    spec = O.ModelSpec(400, 400, psf=wl["psf"], mask=calc, adjust_mask=False, magzero=magzero, threads=8)
    fit = O.FitSpec(spec, wl["data"], wl["sigma"], wl["region"])
    return wl, calc, fit


def test_config2_model_and_likelihood():
    wl, calc, fit = _config2_fit()
    assert calc.shape == (450, 450)
    oll, omodel = O.loglike(fit, oracle_profiles(wl["components"]), want_model=True)
    eng = P.Engine()
    g = eng.add_galaxy(400, 400, wl["data"], wl["sigma"], psf=wl["psf"], calcregion=calc, region=wl["region"],
                       magzero=30.0)
    assert eng.geometry(g) == (450, 450, 25, 25)
    eng.set_batch([g], [2])
    params = gpu_profiles(wl["components"])
    gmodel = eng.model_images(params)[0]
    err, n = rel_err(gmodel, omodel)
    _report("config2 model", err, n)
    assert err <= REL_TOL
    gll = eng.like(params)[0]
    print(f"[parity] config2 logL gpu {gll:.6f} oracle {oll:.6f} diff {gll - oll:.3e}", flush=True)
    assert abs(gll - oll) <= 1e-3


def _walker(comps):
    out = [dict(c) for c in comps]
    out[0].update(xcen=out[0]["xcen"] + 0.15, mag=out[0]["mag"] + 0.05, re=out[0]["re"] * 1.03,
                  nser=out[0]["nser"] * 0.98, axrat=0.88)
    out[1].update(ycen=out[1]["ycen"] - 0.1, ang=out[1]["ang"] + 2.0, re=out[1]["re"] * 0.98, nser=1.05)
    return out


def test_config2_gradient_against_oracle():
    """Adjoint gradient vs the FP64 oracle's fixed-tree analytic gradient away
    from the optimum (at the optimum the gradient is a small residual of
    cancelling sums, compared below on the Fisher scale)."""
    wl, calc, fit = _config2_fit()
    eng = P.Engine()
    g = eng.add_galaxy(400, 400, wl["data"], wl["sigma"], psf=wl["psf"], calcregion=calc, region=wl["region"],
                       magzero=30.0)
    eng.set_batch([g], [2])
    slots = [0, 1, 2, 3, 4, 6, 7]  # xcen ycen mag ang axrat re nser
    comps = _walker(wl["components"])
    oll, ograd = O.loglike_grad(fit, oracle_profiles(comps))
    gll, ggrad = eng.like_grad(gpu_profiles(comps))
    print(f"[parity] walker logL gpu {gll[0]:.6f} oracle {oll:.6f}", flush=True)
    assert abs(gll[0] - oll) <= 1e-3
    for k in range(2):
        big = np.abs(ograd[k, slots]).max()
        for s in slots:
            o, gv = ograd[k, s], ggrad[k, s]
            print(f"[parity] grad comp {k} slot {O.GRAD_SLOTS[s]}: gpu {gv:.6e} oracle {o:.6e}", flush=True)
            assert abs(gv - o) <= 1e-3 * max(abs(o), 5e-2 * big), (k, s)
    # at the truth: errors on the scale of the Fisher information sqrt(sum (dm/dtheta)^2 / sigma^2)
    oll0, ograd0 = O.loglike_grad(fit, oracle_profiles(wl["components"]))
    gll0, ggrad0 = eng.like_grad(gpu_profiles(wl["components"]))
    fisher_mag = np.sqrt(np.sum((0.4 * np.log(10) * wl["model"]) ** 2 / wl["sigma"] ** 2 * wl["region"]))
    for k in range(2):
        for s in slots:
            assert abs(ggrad0[k, s] - ograd0[k, s]) <= 1e-4 * fisher_mag, (k, s)


def test_adjust_mask_bit_exact():
    wl = W.config2()
    ref = O.adjust_mask(wl["region"], (400, 400), wl["psf"])
    got = P.adjust_mask(wl["region"], (400, 400), wl["psf"])
    assert got.shape == ref.shape and np.array_equal(got, ref)


@pytest.mark.parametrize("ksize", [5, 25, 51])
def test_brute_convolution(ksize):
    rng = np.random.default_rng(ksize)
    src = rng.random((90, 70))
    krn = rng.random((ksize, ksize))
    mask = (rng.random((90, 70)) > 0.3).astype(np.uint8)
    ref = O.convolve(src, krn, mask)
    got = P.convolve(src, krn, mask)
    err, n = rel_err(got, ref)
    _report(f"brute conv k={ksize}", err, n)
    assert err <= REL_TOL
    assert np.all(got[mask == 0] == 0)


def test_fp32_values_meet_the_gates():
    """The FP32 value option (direct pixels and recursion leaves in FP32,
    decisions still FP64-exact) against the same gates: config 1 at magzero 30
    (the logL most sensitive to the model) and config 2."""
    wl = W.config1(oracle_model_fn(), magzero=30.0)
    fit = O.FitSpec(O.ModelSpec(100, 100, psf=wl["psf"], magzero=30.0, threads=8), wl["data"], wl["sigma"],
                    wl["region"])
    oll, omodel = O.loglike(fit, oracle_profiles(wl["components"]), want_model=True)
    eng = P.Engine()
    eng.set_render_fp64(False)
    g = eng.add_galaxy(100, 100, wl["data"], wl["sigma"], psf=wl["psf"], region=wl["region"], magzero=30.0)
    eng.set_batch([g], [1])
    params = gpu_profiles(wl["components"])
    err, n = rel_err(eng.model_images(params)[0], omodel)
    gll = eng.like(params)[0]
    print(f"[parity] fp32 values config1 magzero 30: pixel err {err:.3e}, logL diff {gll - oll:.3e}", flush=True)
    assert err <= REL_TOL
    assert abs(gll - oll) <= 1e-3
    wl2, calc, fit2 = _config2_fit()
    oll2, omodel2 = O.loglike(fit2, oracle_profiles(wl2["components"]), want_model=True)
    eng2 = P.Engine()
    eng2.set_render_fp64(False)
    g2 = eng2.add_galaxy(400, 400, wl2["data"], wl2["sigma"], psf=wl2["psf"], calcregion=calc, region=wl2["region"],
                         magzero=30.0)
    eng2.set_batch([g2], [2])
    params2 = gpu_profiles(wl2["components"])
    err2, _ = rel_err(eng2.model_images(params2)[0], omodel2)
    gll2 = eng2.like(params2)[0]
    print(f"[parity] fp32 values config2: pixel err {err2:.3e}, logL diff {gll2 - oll2:.3e}", flush=True)
    assert err2 <= REL_TOL
    assert abs(gll2 - oll2) <= 1e-3
