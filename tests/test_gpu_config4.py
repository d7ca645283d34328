"""Config 4 at its stated size: 4096 galaxies of uneven image size (sides
32-256 px), component count (1-3, Sersic / Moffat / Ferrer / King /
core-Sersic) and PSF (Gaussian or Moffat, 15-41 px), evaluated as one batch
with the autotuner's per-galaxy convolution kernels.

* every instance: finite logL and gradient;
* batch order: the same batch in a permuted order gives the same values, to
  the FP32 canvas's rounding (its sums follow the order in which waves add
  sub-pixel values: each galaxy's logL within four times its FP32 floor,
  below);
* a seeded sample of 64 galaxies against the FP64 oracle, in the default FP32
  mode and in the FP64 model mode (FP64 values, canvas and FFT): model image
  per pixel <= 1e-5 relative above 1e-6 of the peak; the adjoint gradient
  within 2e-3 of the oracle's fixed-tree gradient (its analytic gradient, or
  fixed-tree central differences for core-Sersic); |dlogL| <= 1e-3 in the
  FP64 mode, and in the FP32 mode <= max(1e-3, 3 x the galaxy's FP32 floor),
  the floor being the logL change that rounding each model pixel to FP32 can
  cause (helpers.fp32_logl_floor): bright compact sources at magzero 30 reach
  floors above 1e-3, the stated limit of the FP32 canvas.

The truth models that make the synthetic data come from the GPU model path
(one call for the whole batch); the oracle then checks the sample."""
import os
import time

import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import fp32_logl_floor, gpu_profiles, oracle_profiles, rel_err

pytestmark = pytest.mark.gpu

NGAL = 4096
NCHECK = 64
SHAPES = {"sersic": 2, "moffat": 2, "ferrer": 3, "king": 3, "coresersic": 5}
THREADS = min(16, os.cpu_count() or 1)


def _walker(comps, k):
    out = [dict(c) for c in comps]
    for c in out:
        c["xcen"] += 0.13 * (k % 3 - 1)
        c["mag"] += 0.04
        c["axrat"] = min(0.98, c["axrat"] * 1.02)
    return out


def _add(eng, g):
    return eng.add_galaxy(g["width"], g["height"], g["data"], g["sigma"], psf=g["psf"], calcregion=g["calc"],
                          region=g["region"], magzero=30.0)


@pytest.fixture(scope="module")
def c4():
    t0 = time.perf_counter()
    gals = W.config4_galaxies(NGAL, seed=4)
    for g in gals:
        g["calc"] = P.adjust_mask(g["region"], (g["width"], g["height"]), g["psf"])
        g["data"] = np.zeros((g["height"], g["width"]))
        g["sigma"] = np.ones((g["height"], g["width"]))
    # noiseless truth models from the GPU model path, one batch
    meng = P.Engine()
    ids = [_add(meng, g) for g in gals]
    meng.set_batch(ids, [len(g["components"]) for g in gals])
    models = meng.model_images([p for g in gals for p in gpu_profiles(g["components"])])
    meng.close()
    W.config4_attach_data(gals, [m.astype(np.float64) for m in models])
    eng = P.Engine()
    ids = [_add(eng, g) for g in gals]
    eng.set_batch(ids, [len(g["components"]) for g in gals])
    sets = [_walker(g["components"], k) for k, g in enumerate(gals)]
    params = P.profile_array([p for s in sets for p in gpu_profiles(s)])
    tune = eng.autotune(params, reps=1)
    kinds = tune["kinds"]
    print(f"[config4] {NGAL} galaxies set up in {time.perf_counter() - t0:.1f} s; autotune "
          f"{ {k: kinds.count(k) for k in set(kinds)} } ms {tune['ms']}", flush=True)
    ll, grad = eng.like_grad(params)
    return dict(gals=gals, eng=eng, ids=ids, sets=sets, params=params, ll=ll, grad=grad, kinds=kinds)


def test_every_instance_is_finite(c4):
    ll, grad = c4["ll"], c4["grad"]
    assert ll.shape == (NGAL,)
    assert np.all(np.isfinite(ll)) and np.all(np.isfinite(grad))
    ncomp = sum(len(g["components"]) for g in c4["gals"])
    assert grad.shape == (ncomp, O.NGRAD)


def _floors(c4, models):
    return np.array([fp32_logl_floor("chisq", g["data"], g["sigma"], m.astype(np.float64), g["region"])
                     for g, m in zip(c4["gals"], models)])


def test_batch_order_invariance(c4):
    gals, eng, ids, sets = c4["gals"], c4["eng"], c4["ids"], c4["sets"]
    floors = _floors(c4, eng.model_images(c4["params"]))
    perm = np.random.default_rng(44).permutation(NGAL)
    eng.set_batch([ids[i] for i in perm], [len(gals[i]["components"]) for i in perm])
    ll_p, grad_p = eng.like_grad(P.profile_array([p for i in perm for p in gpu_profiles(sets[i])]))
    rows = np.cumsum([0] + [len(g["components"]) for g in gals])
    ref_rows = np.concatenate([np.arange(rows[i], rows[i + 1]) for i in perm])
    # restore the original order for the later tests
    eng.set_batch(ids, [len(g["components"]) for g in gals])
    d_ll = np.abs(ll_p - c4["ll"][perm])
    ref = c4["grad"][ref_rows]
    d_g = np.abs(grad_p - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max(axis=1, keepdims=True) + 1e-300)
    print(f"[config4] permuted batch: max |dlogL| {d_ll.max():.3e} (max FP32 floor {floors.max():.3e}), "
          f"max rel grad diff {d_g.max():.3e}", flush=True)
    # each run's canvas pixels are within a few FP32 ulps (one per wave that
    # adds to the pixel): within 4 floors of each other
    assert np.all(d_ll <= 1e-9 + 4 * floors[perm])
    assert d_g.max() <= 2e-3


def _sample_check(c4, eng, tag, fp64):
    gals = c4["gals"]
    pick = sorted(np.random.default_rng(4096).choice(NGAL, NCHECK, replace=False).tolist())
    sub_ids = [c4["ids"][i] for i in pick] if not fp64 else list(range(NCHECK))
    eng.set_batch(sub_ids, [len(gals[i]["components"]) for i in pick])
    sub = P.profile_array([p for i in pick for p in gpu_profiles(c4["sets"][i])])
    models = eng.model_images(sub)
    ll, grad = eng.like_grad(sub)
    rows = np.cumsum([0] + [len(gals[i]["components"]) for i in pick])
    worst = dict(pix=0.0, ll=0.0, grad=0.0, floor=0.0)
    for n, i in enumerate(pick):
        g = gals[i]
        s = c4["sets"][i]
        spec = O.ModelSpec(g["width"], g["height"], psf=g["psf"], mask=g["calc"], adjust_mask=False, magzero=30.0,
                           threads=THREADS)
        fit = O.FitSpec(spec, g["data"], g["sigma"], g["region"])
        oll, omodel = O.loglike(fit, oracle_profiles(s), want_model=True)
        err, _ = rel_err(models[n], omodel)
        dll = abs(ll[n] - oll)
        floor = fp32_logl_floor("chisq", g["data"], g["sigma"], omodel, g["region"])
        gate = 1e-3 if fp64 else max(1e-3, 3 * floor)
        if any(c["type"] == "coresersic" for c in s):
            steps = np.full((len(s), O.NGRAD), 1e-6)
            steps[:, :2] = 1e-8
            _, ograd = O.loglike_grad_fd(fit, oracle_profiles(s), rel_step=steps)
        else:
            _, ograd = O.loglike_grad(fit, oracle_profiles(s))
        ggrad = grad[rows[n]:rows[n + 1]]
        gerr = 0.0
        for k, c in enumerate(s):
            slots = [0, 1, 2, 3, 4] + [6 + j for j in range(SHAPES[c["type"]])]
            o, gv = ograd[k, slots], ggrad[k, slots]
            scale = np.maximum(np.abs(o), 5e-2 * np.abs(o).max())
            gerr = max(gerr, float(np.max(np.abs(gv - o) / scale)))
        kinds = ",".join(c["type"] for c in s)
        print(f"[config4] {tag} galaxy {i} {g['width']}x{g['height']} psf {g['psf'].shape[0]} [{kinds}]: "
              f"pixel err {err:.2e}, |dlogL| {dll:.2e} (logL {oll:.2f}, FP32 floor {floor:.1e}), "
              f"grad err {gerr:.2e} of the gate's scale", flush=True)
        for k, v in (("pix", err), ("ll", dll), ("grad", gerr), ("floor", floor)):
            worst[k] = max(worst[k], v)
        assert err <= 1e-5, i
        assert dll <= gate, i
        assert gerr <= 2e-3, i
    print(f"[config4] {tag} sample of {NCHECK}: worst pixel err {worst['pix']:.2e}, |dlogL| {worst['ll']:.2e} "
          f"(largest FP32 floor {worst['floor']:.2e}), grad {worst['grad']:.2e}", flush=True)


def test_seeded_sample_against_the_oracle(c4):
    _sample_check(c4, c4["eng"], "fp32", fp64=False)
    c4["eng"].set_batch(c4["ids"], [len(g["components"]) for g in c4["gals"]])


def test_seeded_sample_in_the_fp64_model_mode(c4):
    gals = c4["gals"]
    pick = sorted(np.random.default_rng(4096).choice(NGAL, NCHECK, replace=False).tolist())
    eng = P.Engine()
    eng.set_render_fp64(2)
    for i in pick:
        _add(eng, gals[i])
    _sample_check(c4, eng, "fp64", fp64=True)
    eng.close()
