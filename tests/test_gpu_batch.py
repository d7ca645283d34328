"""Heterogeneous batch (config 4 draws): galaxies of different sizes, PSFs,
component counts and profile types evaluated in one launch sequence, each
instance checked against the FP64 oracle: model image (per-pixel relative
error <= 1e-5 above 1e-6 of the peak), logL and the adjoint gradient against
the oracle's central differences on the frozen recursion tree (h = 1e-6
|theta|, 1e-8 |theta| for the centroid).

logL gate: 1e-3 absolute, or twice the FP32 floor of the model image when
that is larger (fp32_logl_floor: the logL change that rounding every model
pixel to FP32 alone can cause).  At magzero 30 a compact, mis-fitted source
(galaxy 8 of these draws at the walker positions: a 1 px Moffat peaking at
2.3e4 counts, logL -561) has a floor of 1.2e-3; the GPU canvas is FP32."""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import fp32_logl_floor, gpu_profiles, oracle_profiles, rel_err

pytestmark = pytest.mark.gpu

NGAL = 10
SHAPES = {"sersic": 2, "moffat": 2, "ferrer": 3, "king": 3, "coresersic": 5}


@pytest.fixture(scope="module")
def batch():
    gals = W.config4_galaxies(NGAL, seed=4)
    fits, models = [], []
    for g in gals:
        w, h = g["width"], g["height"]
        calc = O.adjust_mask(g["region"], (w, h), g["psf"])
        spec = O.ModelSpec(w, h, psf=g["psf"], mask=calc, adjust_mask=False, magzero=30.0, threads=8)
        models.append(O.model_evaluate(spec, oracle_profiles(g["components"]))["image"])
        g["calc"] = calc
        g["spec"] = spec
    W.config4_attach_data(gals, models)
    for g in gals:
        fits.append(O.FitSpec(g["spec"], g["data"], g["sigma"], g["region"]))
    return gals, fits


def _engine(gals, method):
    eng = P.Engine()
    eng.set_conv_method(method)
    ids = [eng.add_galaxy(g["width"], g["height"], g["data"], g["sigma"], psf=g["psf"], calcregion=g["calc"],
                          region=g["region"], magzero=30.0) for g in gals]
    eng.set_batch(ids, [len(g["components"]) for g in gals])
    return eng


def _walker(comps, k):
    out = [dict(c) for c in comps]
    for c in out:
        c["xcen"] += 0.13 * (k % 3 - 1)
        c["mag"] += 0.04
        c["axrat"] = min(0.98, c["axrat"] * 1.02)
    return out


@pytest.mark.parametrize("method", ["brute", "fft", "brute-bf16x3"])
def test_heterogeneous_batch_models_and_likelihoods(batch, method):
    gals, fits = batch
    eng = _engine(gals, method)
    params = [p for g in gals for p in gpu_profiles(g["components"])]
    models = eng.model_images(params)
    ll = eng.like(params)
    for i, (g, f) in enumerate(zip(gals, fits)):
        oll, omodel = O.loglike(f, oracle_profiles(g["components"]), want_model=True)
        err, n = rel_err(models[i], omodel)
        kinds = ",".join(c["type"] for c in g["components"])
        print(f"[batch] {method} galaxy {i} {g['width']}x{g['height']} psf {g['psf'].shape[0]} [{kinds}]: "
              f"pixel err {err:.2e}, logL {ll[i]:.4f} vs {oll:.4f}", flush=True)
        assert err <= 1e-5, i
        assert abs(ll[i] - oll) <= 1e-3, i


def test_heterogeneous_batch_gradients(batch):
    gals, fits = batch
    eng = _engine(gals, "brute")
    sets = [_walker(g["components"], k) for k, g in enumerate(gals)]
    params = [p for s in sets for p in gpu_profiles(s)]
    ll, grad = eng.like_grad(params)
    row = 0
    for i, (g, f, s) in enumerate(zip(gals, fits, sets)):
        # cusps (core-Sersic inner slopes, Sersic n > 1) make the centroid
        # derivative converge only for steps far below the deepest sub-pixel
        # cells (n = 7.4 needs 1e-8 |theta|); 1e-6 |theta| elsewhere
        steps = np.full((len(s), O.NGRAD), 1e-6)
        steps[:, :2] = 1e-8
        oll, ograd = O.loglike_grad_fd(f, oracle_profiles(s), rel_step=steps)
        _, om = O.loglike(f, oracle_profiles(s), want_model=True)
        gate = max(1e-3, 2 * fp32_logl_floor("chisq", g["data"], g["sigma"], om, g["region"]))
        print(f"[batch] galaxy {i}: logL {ll[i]:.6f} vs {oll:.6f} (gate {gate:.2e})", flush=True)
        assert abs(ll[i] - oll) <= gate, i
        for k, c in enumerate(s):
            slots = [0, 1, 2, 3, 4] + [6 + j for j in range(SHAPES[c["type"]])]
            o = ograd[k, slots]
            gv = grad[row + k, slots]
            big = np.abs(o).max()
            print(f"[batch] galaxy {i} comp {k} {c['type']}: gpu {np.array2string(gv, precision=4)} "
                  f"fd {np.array2string(o, precision=4)}", flush=True)
            assert np.all(np.abs(gv - o) <= 2e-3 * np.maximum(np.abs(o), 5e-2 * big)), (i, k)
        row += len(s)


def test_autotuned_batch_keeps_parity(batch):
    """The autotuner's per-galaxy kernel mix evaluates the same logL (FP32 brute
    force and FFT are both inside the gates)."""
    gals, fits = batch
    eng = _engine(gals, "brute")
    params = [p for g in gals for p in gpu_profiles(g["components"])]
    rep = eng.autotune(params, reps=2)
    print(f"[batch] autotune kinds {rep['kinds']} ms {rep['ms']}", flush=True)
    assert len(rep["kinds"]) == len(gals)
    assert set(rep["kinds"]) <= {"brute", "fft"}
    assert rep["ms"]["chosen"] <= min(rep["ms"]["brute_rows4"], rep["ms"]["brute_rows8"], rep["ms"]["brute_rows16"],
                                      rep["ms"]["fft"]) * 1.0001
    ll = eng.like(params)
    for i, (g, f) in enumerate(zip(gals, fits)):
        oll = O.loglike(f, oracle_profiles(g["components"]))
        assert abs(ll[i] - oll) <= 1e-3, i
