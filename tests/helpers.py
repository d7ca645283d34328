"""Shared helpers: the same component description for the oracle and the GPU path."""
import numpy as np

import oracle as O
import profit_amd as P


def oracle_profiles(comps, convolve=True, **extra):
    out = []
    for c in comps:
        c = dict(c)
        t = c.pop("type")
        c.update(extra)
        out.append(O.make_profile(t, convolve=convolve, **c))
    return out


def gpu_profiles(comps, convolve=True, **extra):
    out = []
    for c in comps:
        c = dict(c)
        t = c.pop("type")
        c.update(extra)
        out.append(P.make_profile(t, convolve=convolve, **c))
    return out


def oracle_model_fn(threads=8):
    def fn(w, h, psf, comps, magzero):
        spec = O.ModelSpec(w, h, psf=psf, magzero=magzero, threads=threads)
        return O.model_evaluate(spec, oracle_profiles(comps))["image"]
    return fn


def rel_err(test, ref, floor_frac=1e-6):
    """Max relative error over pixels above floor_frac of the reference peak."""
    ref = np.asarray(ref, dtype=np.float64)
    test = np.asarray(test, dtype=np.float64)
    peak = np.abs(ref).max()
    sel = np.abs(ref) > floor_frac * peak
    return float(np.max(np.abs(test[sel] - ref[sel]) / np.abs(ref[sel]))), int(sel.sum())
