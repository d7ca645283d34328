"""THE ORACLE -- test infrastructure only.

ctypes binding of ``liboracle.so``, the plain-C FP64 restatement of
libprofit's CPU model-image path (see ``oracle.h``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
package; the product (``profit_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SERSIC, MOFFAT, FERRER, KING, CORESERSIC, BROKENEXP, POINT, SKY = range(8)
PROFILE_TYPES = {"sersic": SERSIC, "moffat": MOFFAT, "ferrer": FERRER, "ferrers": FERRER,
                 "king": KING, "coresersic": CORESERSIC, "brokenexp": BROKENEXP, "psf": POINT, "sky": SKY}
SHAPE_NAMES = {
    SERSIC: ("re", "nser"),
    MOFFAT: ("fwhm", "con"),
    FERRER: ("rout", "a", "b"),
    KING: ("rc", "rt", "a"),
    CORESERSIC: ("re", "rb", "nser", "a", "b"),
    BROKENEXP: ("h1", "h2", "rb", "a"),
    POINT: (),
    SKY: ("bg",),
}
CONV = {"brute": 0, "brute-old": 1, "fft": 2, "null": 3}
LIKE = {"norm": 0, "chisq": 1, "t": 2, "pois": 3}
NGRAD = 11
GRAD_SLOTS = ("xcen", "ycen", "mag", "ang", "axrat", "box", "s0", "s1", "s2", "s3", "s4")


class Profile(C.Structure):
    _fields_ = [("type", C.c_int), ("convolve", C.c_int), ("rough", C.c_int),
                ("adjust", C.c_int), ("rescale_flux", C.c_int),
                ("xcen", C.c_double), ("ycen", C.c_double), ("mag", C.c_double),
                ("ang", C.c_double), ("axrat", C.c_double), ("box", C.c_double),
                ("acc", C.c_double), ("rscale_switch", C.c_double), ("rscale_max", C.c_double),
                ("resolution", C.c_uint), ("max_recursions", C.c_uint),
                ("shape", C.c_double * 5)]


class Model(C.Structure):
    _fields_ = [("width", C.c_uint), ("height", C.c_uint), ("scale_x", C.c_double),
                ("scale_y", C.c_double), ("magzero", C.c_double), ("finesampling", C.c_uint),
                ("return_finesampled", C.c_int), ("crop", C.c_int), ("adjust_mask", C.c_int),
                ("dry_run", C.c_int), ("psf", C.POINTER(C.c_double)), ("psf_w", C.c_uint),
                ("psf_h", C.c_uint), ("mask", C.POINTER(C.c_ubyte)), ("mask_w", C.c_uint),
                ("mask_h", C.c_uint), ("convolver", C.c_int), ("threads", C.c_int),
                ("psf_scale_x", C.c_double), ("psf_scale_y", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("cls", C.POINTER(C.c_ubyte)), ("nodes", C.POINTER(C.c_int)), ("cap", C.c_uint)]


class CompInfo(C.Structure):
    _fields_ = [("rscale", C.c_double), ("ie", C.c_double), ("bn", C.c_double),
                ("lumtot", C.c_double), ("rbox", C.c_double), ("flux_scale", C.c_double),
                ("acc", C.c_double), ("rscale_switch", C.c_double), ("rscale_max", C.c_double),
                ("resolution", C.c_uint), ("max_recursions", C.c_uint),
                ("cos_ang", C.c_double), ("sin_ang", C.c_double)]


class Fit(C.Structure):
    _fields_ = [("model", Model), ("data", C.POINTER(C.c_double)),
                ("sigma", C.POINTER(C.c_double)), ("region", C.POINTER(C.c_ubyte)),
                ("likefunc", C.c_int)]


class QuadInfo(C.Structure):
    _fields_ = [("abserr", C.c_double), ("ier", C.c_int), ("last", C.c_int), ("neval", C.c_int)]


INTEGRAND = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, u, i, p = C.c_double, C.c_uint, C.c_int, C.c_void_p
        for name, res, args in [
            ("or_gammafn", d, [d]), ("or_lgammafn", d, [d]), ("or_beta", d, [d, d]),
            ("or_digamma", d, [d]), ("or_pgamma", d, [d, d]), ("or_qgamma", d, [d, d]),
            ("or_dpgamma_da", d, [d, d]),
            ("or_integrate_qags_tol", d, [INTEGRAND, p, d, d, d, d, i, C.POINTER(QuadInfo)]),
            ("or_integrate_qagi_tol", d, [INTEGRAND, p, d, d, d, i, C.POINTER(QuadInfo)]),
            ("or_quadpack_nodes", C.POINTER(d), [i]),
            ("or_profile_defaults", None, [i, C.POINTER(Profile)]),
            ("or_model_defaults", None, [C.POINTER(Model)]),
            ("or_last_error", C.c_char_p, []),
            ("or_model_evaluate", i, [C.POINTER(Model), C.POINTER(Profile), i, C.POINTER(d), u,
                                      C.POINTER(u), C.POINTER(u), C.POINTER(i), C.POINTER(i),
                                      C.POINTER(Diag)]),
            ("or_model_drawing_dims", i, [C.POINTER(Model), C.POINTER(Profile), i, C.POINTER(u),
                                          C.POINTER(u), C.POINTER(u), C.POINTER(u)]),
            ("or_adjust_mask", i, [C.POINTER(C.c_ubyte), u, u, u, u, C.POINTER(d), u, u, u,
                                   C.POINTER(C.c_ubyte), u, C.POINTER(u), C.POINTER(u)]),
            ("or_convolve", i, [i, C.POINTER(d), u, u, C.POINTER(d), u, u, C.POINTER(C.c_ubyte),
                                C.POINTER(d), i]),
            ("or_mask_expand_by", None, [C.POINTER(C.c_ubyte), u, u, u, u, C.POINTER(C.c_ubyte)]),
            ("or_profile_info", i, [C.POINTER(Profile), d, d, d, C.POINTER(CompInfo)]),
            ("or_profile_eval_at", d, [C.POINTER(Profile), d, d, d]),
            ("or_loglike_image", i, [i, C.POINTER(d), C.POINTER(d), C.POINTER(C.c_ubyte),
                                     C.POINTER(d), u, C.POINTER(d), C.POINTER(d)]),
            ("or_loglike", i, [C.POINTER(Fit), C.POINTER(Profile), i, C.POINTER(d), C.POINTER(d)]),
            ("or_loglike_grad", i, [C.POINTER(Fit), C.POINTER(Profile), i, C.POINTER(d),
                                    C.POINTER(d)]),
            ("or_loglike_grad_fd", i, [C.POINTER(Fit), C.POINTER(Profile), i, C.POINTER(d),
                                       C.POINTER(C.c_ubyte), i, C.POINTER(d), C.POINTER(d)]),
            ("or_model_image_frozen", i, [C.POINTER(Fit), C.POINTER(Profile), C.POINTER(Profile),
                                          i, C.POINTER(d)]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().or_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _bptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_ubyte))


# ---------------------------------------------------------------- scalars
def qgamma(p, shape):
    return lib().or_qgamma(p, shape)


def pgamma(x, shape):
    return lib().or_pgamma(x, shape)


def gammafn(x):
    return lib().or_gammafn(x)


def beta(a, b):
    return lib().or_beta(a, b)


def digamma(x):
    return lib().or_digamma(x)


def dpgamma_da(x, a):
    return lib().or_dpgamma_da(x, a)


def quadpack_nodes(which):
    n = (11, 11, 5, 8, 8, 8)[which]
    ptr = lib().or_quadpack_nodes(which)
    return np.array([ptr[i] for i in range(n)])


def integrate(f, a, b=None, epsabs=1e-4, epsrel=1e-4, limit=100):
    """QUADPACK dqags over [a,b] or dqagi over [a,inf) when b is None."""
    cb = INTEGRAND(lambda x, _ctx: float(f(x)))
    info = QuadInfo()
    if b is None:
        r = lib().or_integrate_qagi_tol(cb, None, a, epsabs, epsrel, limit, C.byref(info))
    else:
        r = lib().or_integrate_qags_tol(cb, None, a, b, epsabs, epsrel, limit, C.byref(info))
    return r, {"abserr": info.abserr, "ier": info.ier, "last": info.last, "neval": info.neval}


# ---------------------------------------------------------------- profiles
def make_profile(name, **params):
    """A libprofit profile with its constructor defaults, then ``params``."""
    t = PROFILE_TYPES[name]
    p = Profile()
    lib().or_profile_defaults(t, C.byref(p))
    shape = SHAPE_NAMES[t]
    for k, v in params.items():
        if k in shape:
            p.shape[shape.index(k)] = float(v)
        elif k in ("convolve", "rough", "adjust", "rescale_flux"):
            setattr(p, k, int(bool(v)))
        elif k in ("resolution", "max_recursions"):
            setattr(p, k, int(v))
        elif k in ("xcen", "ycen", "mag", "ang", "axrat", "box", "acc", "rscale_switch",
                   "rscale_max"):
            setattr(p, k, float(v))
        else:
            raise KeyError(f"unknown parameter {k} for profile {name}")
    return p


def profile_array(profiles):
    arr = (Profile * max(1, len(profiles)))()
    for i, p in enumerate(profiles):
        arr[i] = p
    return arr


def profile_info(p, magzero=0.0, scale=(1.0, 1.0)):
    info = CompInfo()
    _check(lib().or_profile_info(C.byref(p), magzero, scale[0], scale[1], C.byref(info)))
    return {f[0]: getattr(info, f[0]) for f in CompInfo._fields_}


def eval_at(p, x, y, magzero=0.0):
    return lib().or_profile_eval_at(C.byref(p), magzero, x, y)


# ---------------------------------------------------------------- model
class ModelSpec:
    """Holds numpy buffers alive for an ``or_model``."""

    def __init__(self, width, height, psf=None, mask=None, magzero=0.0, scale=(1.0, 1.0),
                 finesampling=1, return_finesampled=True, crop=True, adjust_mask=True,
                 convolver="brute", threads=1, dry_run=False, psf_scale=(1.0, 1.0)):
        m = Model()
        lib().or_model_defaults(C.byref(m))
        m.width, m.height = int(width), int(height)
        m.scale_x, m.scale_y = scale
        m.magzero = magzero
        m.finesampling = finesampling
        m.return_finesampled = int(return_finesampled)
        m.crop = int(crop)
        m.adjust_mask = int(adjust_mask)
        m.dry_run = int(dry_run)
        m.convolver = CONV[convolver]
        m.threads = int(threads)
        m.psf_scale_x, m.psf_scale_y = psf_scale
        self._psf = None
        self._mask = None
        if psf is not None:
            self._psf = np.ascontiguousarray(psf, dtype=np.float64)
            m.psf = _dptr(self._psf)
            m.psf_h, m.psf_w = self._psf.shape
        if mask is not None:
            self._mask = np.ascontiguousarray(mask, dtype=np.uint8)
            m.mask = _bptr(self._mask)
            m.mask_h, m.mask_w = self._mask.shape
        self.m = m


def model_evaluate(spec, profiles, diag=False):
    """Model::evaluate.  Images are numpy (rows=y, cols=x), i.e. [j, i]."""
    L = lib()
    arr = profile_array(profiles)
    dw, dh, px, py = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
    _check(L.or_model_drawing_dims(C.byref(spec.m), arr, len(profiles), C.byref(dw), C.byref(dh),
                                   C.byref(px), C.byref(py)))
    cap = max(dw.value * dh.value, 1) * 4 + 64
    out = np.zeros(cap)
    ow, oh, ox, oy = C.c_uint(), C.c_uint(), C.c_int(), C.c_int()
    d = None
    cls = nodes = None
    if diag:
        n = dw.value * dh.value
        cls = np.zeros(n * max(1, len(profiles)), dtype=np.uint8)
        nodes = np.zeros(n * max(1, len(profiles)), dtype=np.int32)
        d = Diag(_bptr(cls), nodes.ctypes.data_as(C.POINTER(C.c_int)), n)
    _check(L.or_model_evaluate(C.byref(spec.m), arr, len(profiles), _dptr(out), cap, C.byref(ow),
                               C.byref(oh), C.byref(ox), C.byref(oy),
                               C.byref(d) if d is not None else None))
    img = out[: ow.value * oh.value].reshape(oh.value, ow.value).copy()
    res = {"image": img, "offset": (ox.value, oy.value),
           "drawing": (dw.value, dh.value), "padding": (px.value, py.value)}
    if diag:
        shp = (max(1, len(profiles)), dh.value, dw.value)
        res["cls"] = cls.reshape(shp)
        res["nodes"] = nodes.reshape(shp)
    return res


def adjust_mask(mask, img_dims, psf, finesampling=1):
    """Model::adjust(mask, dims, psf, fs) -> adjusted mask (numpy [j, i])."""
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    psf = np.ascontiguousarray(psf, dtype=np.float64) if psf is not None else np.zeros((0, 0))
    mh, mw = mask.shape
    ph, pw = psf.shape if psf.size else (0, 0)
    cap = (mw * finesampling + pw) * (mh * finesampling + ph) + 16
    out = np.zeros(cap, dtype=np.uint8)
    ow, oh = C.c_uint(), C.c_uint()
    _check(lib().or_adjust_mask(_bptr(mask), mw, mh, img_dims[0], img_dims[1],
                                _dptr(psf) if psf.size else None, pw, ph, finesampling,
                                _bptr(out), cap, C.byref(ow), C.byref(oh)))
    return out[: ow.value * oh.value].reshape(oh.value, ow.value).copy()


def convolve(src, krn, mask=None, convolver="brute", threads=1):
    src = np.ascontiguousarray(src, dtype=np.float64)
    krn = np.ascontiguousarray(krn, dtype=np.float64)
    out = np.zeros_like(src)
    m = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        m = _bptr(mask)
    sh, sw = src.shape
    kh, kw = krn.shape
    _check(lib().or_convolve(CONV[convolver], _dptr(src), sw, sh, _dptr(krn), kw, kh, m,
                             _dptr(out), threads))
    return out


def expand_by(mask, pad):
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.zeros_like(mask)
    h, w = mask.shape
    lib().or_mask_expand_by(_bptr(mask), w, h, pad[0], pad[1], _bptr(out))
    return out


# ---------------------------------------------------------------- likelihood
def loglike_image(likefunc, data, sigma, model, region=None, want_grad=False):
    data = np.ascontiguousarray(data, dtype=np.float64).ravel()
    sigma = np.ascontiguousarray(sigma, dtype=np.float64).ravel()
    model = np.ascontiguousarray(model, dtype=np.float64).ravel()
    reg = None
    if region is not None:
        region = np.ascontiguousarray(region, dtype=np.uint8).ravel()
        reg = _bptr(region)
    ll = C.c_double()
    g = np.zeros_like(model) if want_grad else None
    _check(lib().or_loglike_image(LIKE[likefunc], _dptr(data), _dptr(sigma), reg, _dptr(model),
                                  data.size, C.byref(ll), _dptr(g) if want_grad else None))
    return (ll.value, g) if want_grad else ll.value


class FitSpec:
    """An ``or_fit``: a model whose mask is the pre-adjusted calcregion (as
    profitSetupData prepares it) plus data, sigma and the likelihood region."""

    def __init__(self, spec, data, sigma, region, likefunc="chisq"):
        self.spec = spec
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        self.region = None if region is None else np.ascontiguousarray(region, dtype=np.uint8)
        f = Fit()
        f.model = spec.m
        f.data = _dptr(self.data)
        f.sigma = _dptr(self.sigma)
        f.region = _bptr(self.region) if self.region is not None else None
        f.likefunc = LIKE[likefunc]
        self.f = f


def loglike(fit, profiles, want_model=False):
    arr = profile_array(profiles)
    ll = C.c_double()
    n = fit.spec.m.width * fit.spec.m.height
    mod = np.zeros(n) if want_model else None
    _check(lib().or_loglike(C.byref(fit.f), arr, len(profiles), C.byref(ll),
                            _dptr(mod) if want_model else None))
    if want_model:
        return ll.value, mod.reshape(fit.spec.m.height, fit.spec.m.width)
    return ll.value


def loglike_grad(fit, profiles):
    arr = profile_array(profiles)
    ll = C.c_double()
    g = np.zeros(NGRAD * len(profiles))
    _check(lib().or_loglike_grad(C.byref(fit.f), arr, len(profiles), C.byref(ll), _dptr(g)))
    return ll.value, g.reshape(len(profiles), NGRAD)


def loglike_grad_fd(fit, profiles, rel_step=1e-4, slots=None, frozen=True):
    n = len(profiles)
    arr = profile_array(profiles)
    steps = np.full(n * NGRAD, rel_step, dtype=np.float64) if np.isscalar(rel_step) else \
        np.ascontiguousarray(rel_step, dtype=np.float64).ravel()
    mask = None
    if slots is not None:
        mask = np.ascontiguousarray(slots, dtype=np.uint8).ravel()
    ll = C.c_double()
    g = np.zeros(NGRAD * n)
    _check(lib().or_loglike_grad_fd(C.byref(fit.f), arr, n, _dptr(steps),
                                    _bptr(mask) if mask is not None else None, int(frozen),
                                    C.byref(ll), _dptr(g)))
    return ll.value, g.reshape(n, NGRAD)


def model_image_frozen(fit, base, pert):
    n = fit.spec.m.width * fit.spec.m.height
    out = np.zeros(n)
    _check(lib().or_model_image_frozen(C.byref(fit.f), profile_array(base), profile_array(pert),
                                       len(base), _dptr(out)))
    return out.reshape(fit.spec.m.height, fit.spec.m.width)
