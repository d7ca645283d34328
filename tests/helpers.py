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


class OracleEvaluator:
    """The FP64 oracle behind profit_amd.fit.FitData's evaluator interface
    (tests only): logL and its fixed-tree analytic gradient per batch member,
    with the calc region Model::adjust makes."""

    def __init__(self, data):
        h, w = data.image.shape
        self.calcregion = None
        if data.psf is not None:
            self.calcregion = O.adjust_mask(data.region.astype(np.uint8), (w, h), data.psf, data.finesample)
        spec = O.ModelSpec(w, h, psf=data.psf, mask=self.calcregion, adjust_mask=False, magzero=data.magzero,
                           threads=8)
        self.fit = O.FitSpec(spec, data.image, data.sigma, data.region.astype(np.uint8), likefunc=data.like_func)

    @staticmethod
    def _profiles(cl):
        return [O.make_profile(kind, convolve=conv, **d) for kind, conv, d in cl]

    def like(self, comp_lists):
        return np.array([O.loglike(self.fit, self._profiles(cl)) for cl in comp_lists])

    def like_grad(self, comp_lists):
        out = [O.loglike_grad(self.fit, self._profiles(cl)) for cl in comp_lists]
        return np.array([o[0] for o in out]), np.stack([o[1] for o in out])


def fp32_logl_floor(likefunc, data, sigma, model, region=None):
    """The largest logL change that rounding each pixel of the model image to
    FP32 can cause (sum over the likelihood pixels of |dlogL/dm| |m| 2^-24):
    the precision floor of a model held in FP32, as the GPU canvas is."""
    _, dll = O.loglike_image(likefunc, data, sigma, model, region, want_grad=True)
    m = np.asarray(model, dtype=np.float64).ravel()
    return float(np.sum(np.abs(dll * m)) * 2.0 ** -24)
