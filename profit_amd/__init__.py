"""profit_amd: MI355X-native model-image path of libprofit (ProFit's C++ core).

Python binding of the C-ABI in ``include/profit_amd.h`` (``libprofit_amd.so``,
built in-tree by ``profit_amd/hipbuild.py``).  Every compute call runs the HIP
kernels; there is no CPU fallback: without the built library, or without a
GPU, calls raise.

Parameter names follow libprofit (src/libprofit/src/profit/radial.cpp:724-752
and each profile's constructor); images are numpy arrays indexed [y, x], which
is libprofit's row-major layout with x fastest.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprofit_amd.so")

SERSIC, MOFFAT, FERRER, KING, CORESERSIC, BROKENEXP = range(6)
SKY, NULL, PSF = 100, 101, 102
PROFILE_TYPES = {"sersic": SERSIC, "moffat": MOFFAT, "ferrer": FERRER, "ferrers": FERRER,
                 "king": KING, "coresersic": CORESERSIC, "brokenexp": BROKENEXP,
                 "sky": SKY, "null": NULL, "psf": PSF}
SHAPE_NAMES = {
    SERSIC: ("re", "nser"),
    MOFFAT: ("fwhm", "con"),
    FERRER: ("rout", "a", "b"),
    KING: ("rc", "rt", "a"),
    CORESERSIC: ("re", "rb", "nser", "a", "b"),
    BROKENEXP: ("h1", "h2", "rb", "a"),
    SKY: ("bg",),
    NULL: (),
    PSF: (),
}
CONVOLVERS = {"brute-old": 0, "brute": 1, "fft": 3, "null": 4, "brute-bf16x3": 16}
LIKE = {"norm": 0, "chisq": 1, "t": 2, "pois": 3}
NGRAD = 11
GRAD_SLOTS = ("xcen", "ycen", "mag", "ang", "axrat", "box", "s0", "s1", "s2", "s3", "s4")


class ProfitAmdError(RuntimeError):
    pass


class Profile(C.Structure):
    _fields_ = [("type", C.c_int32), ("convolve", C.c_int32), ("rough", C.c_int32),
                ("adjust", C.c_int32), ("rescale_flux", C.c_int32), ("reserved", C.c_int32),
                ("xcen", C.c_double), ("ycen", C.c_double), ("mag", C.c_double),
                ("ang", C.c_double), ("axrat", C.c_double), ("box", C.c_double),
                ("acc", C.c_double), ("rscale_switch", C.c_double), ("rscale_max", C.c_double),
                ("resolution", C.c_uint32), ("max_recursions", C.c_uint32),
                ("shape", C.c_double * 5)]


class ModelDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("scale_x", C.c_double),
                ("scale_y", C.c_double), ("magzero", C.c_double), ("finesampling", C.c_uint32),
                ("return_finesampled", C.c_int32), ("crop", C.c_int32), ("adjust_mask", C.c_int32),
                ("dry_run", C.c_int32), ("psf", C.POINTER(C.c_double)), ("psf_w", C.c_uint32),
                ("psf_h", C.c_uint32), ("mask", C.POINTER(C.c_uint8)), ("mask_w", C.c_uint32),
                ("mask_h", C.c_uint32), ("convolver", C.c_int32), ("device", C.c_int32),
                ("psf_scale_x", C.c_double), ("psf_scale_y", C.c_double)]


class GalaxyDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("scale_x", C.c_double),
                ("scale_y", C.c_double), ("magzero", C.c_double), ("psf", C.POINTER(C.c_double)),
                ("psf_w", C.c_uint32), ("psf_h", C.c_uint32),
                ("calcregion", C.POINTER(C.c_uint8)), ("calc_w", C.c_uint32),
                ("calc_h", C.c_uint32), ("region", C.POINTER(C.c_uint8)),
                ("data", C.POINTER(C.c_double)), ("sigma", C.POINTER(C.c_double)),
                ("likefunc", C.c_int32), ("psf_scale_x", C.c_double), ("psf_scale_y", C.c_double)]


# every symbol include/profit_amd.h declares: (name, restype, argtypes)
_P = C.POINTER
SYMBOLS = [
    ("profit_amd_last_error", C.c_char_p, []),
    ("profit_amd_version", C.c_char_p, []),
    ("profit_amd_profile_defaults", C.c_int, [C.c_int32, _P(Profile)]),
    ("profit_amd_make_model", C.c_int, [_P(ModelDesc), _P(Profile), C.c_int32, _P(C.c_double), C.c_size_t,
                                        _P(C.c_uint32), _P(C.c_uint32), _P(C.c_int32), _P(C.c_int32)]),
    ("profit_amd_drawing_dims", C.c_int, [_P(ModelDesc), _P(Profile), C.c_int32, _P(C.c_uint32),
                                          _P(C.c_uint32)]),
    ("profit_amd_output_dims", C.c_int, [_P(ModelDesc), _P(Profile), C.c_int32, _P(C.c_uint32),
                                         _P(C.c_uint32)]),
    ("profit_amd_clear_cache", C.c_int, []),
    ("profit_amd_adjust_mask", C.c_int, [_P(C.c_uint8), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                         _P(C.c_double), C.c_uint32, C.c_uint32, C.c_uint32,
                                         _P(C.c_uint8), C.c_size_t, _P(C.c_uint32), _P(C.c_uint32)]),
    ("profit_amd_convolve", C.c_int, [C.c_int32, _P(C.c_double), C.c_uint32, C.c_uint32, _P(C.c_double),
                                      C.c_uint32, C.c_uint32, _P(C.c_uint8), _P(C.c_double), C.c_int32]),
    ("profit_amd_convolve_noalloc", C.c_int, [C.c_int32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                              C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int32,
                                              C.c_void_p]),
    ("profit_amd_convolvers", C.c_char_p, []),
    ("profit_amd_downsample", C.c_int, [_P(C.c_double), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32,
                                        _P(C.c_double), C.c_size_t, _P(C.c_uint32), _P(C.c_uint32)]),
    ("profit_amd_upsample", C.c_int, [_P(C.c_double), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32,
                                      _P(C.c_double), C.c_size_t, _P(C.c_uint32), _P(C.c_uint32)]),
    ("profit_amd_engine_create", C.c_int, [C.c_int32, _P(C.c_void_p)]),
    ("profit_amd_engine_destroy", None, [C.c_void_p]),
    ("profit_amd_engine_add_galaxy", C.c_int, [C.c_void_p, _P(GalaxyDesc), _P(C.c_int32)]),
    ("profit_amd_engine_set_batch", C.c_int, [C.c_void_p, _P(C.c_int32), _P(C.c_int32), C.c_int32]),
    ("profit_amd_engine_set_tile_rows", C.c_int, [C.c_void_p, C.c_int32]),
    ("profit_amd_engine_set_render_fp64", C.c_int, [C.c_void_p, C.c_int32]),
    ("profit_amd_engine_set_grad_slots", C.c_int, [C.c_void_p, C.c_uint32]),
    ("profit_amd_engine_set_conv_method", C.c_int, [C.c_void_p, C.c_int32]),
    ("profit_amd_engine_autotune", C.c_int, [C.c_void_p, _P(Profile), C.c_int32, C.c_int32, _P(C.c_int32),
                                             _P(C.c_double)]),
    ("profit_amd_engine_like", C.c_int, [C.c_void_p, _P(Profile), _P(C.c_double)]),
    ("profit_amd_engine_like_grad", C.c_int, [C.c_void_p, _P(Profile), _P(C.c_double), _P(C.c_double)]),
    ("profit_amd_engine_model_images", C.c_int, [C.c_void_p, _P(Profile), _P(C.c_float)]),
    ("profit_amd_engine_diagnostics", C.c_int, [C.c_void_p, _P(Profile), _P(C.c_int32), _P(C.c_int32)]),
    ("profit_amd_engine_galaxy_geometry", C.c_int, [C.c_void_p, C.c_int32, _P(C.c_uint32), _P(C.c_uint32),
                                                     _P(C.c_uint32), _P(C.c_uint32)]),
    ("profit_amd_engine_set_timing", C.c_int, [C.c_void_p, C.c_int32]),
    ("profit_amd_engine_stats", C.c_int, [C.c_void_p, _P(C.c_double), _P(C.c_int64), _P(C.c_int64),
                                          _P(C.c_int32), _P(C.c_double), _P(C.c_double)]),
    ("profit_amd_engine_adjusted_settings", C.c_int, [C.c_void_p, _P(Profile), _P(C.c_int32), _P(C.c_int32),
                                                      _P(C.c_double), _P(C.c_double), _P(C.c_double)]),
    ("profit_amd_engine_fft_stats", C.c_int, [C.c_void_p, _P(C.c_double), _P(C.c_int64)]),
    ("profit_amd_engine_reset_stats", C.c_int, [C.c_void_p]),
]
STAGES = ("setup", "render", "subsample", "conv_like", "like_finalize", "adjoint_conv", "adjoint_render",
          "adjoint_subsample", "grad_finalize")

_lib = None


def build():
    from . import hipbuild
    return hipbuild.build()


def lib():
    """The loaded C-ABI library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ProfitAmdError(f"{LIB_PATH} is missing: run profit_amd.build() (hipcc, gfx950) first")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SYMBOLS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise ProfitAmdError(lib().profit_amd_last_error().decode())


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def make_profile(name, **params):
    """A profile with libprofit's constructor defaults, then ``params``."""
    t = PROFILE_TYPES[name]
    p = Profile()
    _check(lib().profit_amd_profile_defaults(t, C.byref(p)))
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


def _model_desc(width, height, psf, mask, magzero, scale, finesampling, return_finesampled, crop,
                adjust_mask, convolver, device, dry_run, keep, psf_scale=(1.0, 1.0)):
    d = ModelDesc()
    d.psf_scale_x, d.psf_scale_y = psf_scale
    d.width, d.height = int(width), int(height)
    d.scale_x, d.scale_y = scale
    d.magzero = magzero
    d.finesampling = int(finesampling)
    d.return_finesampled = int(return_finesampled)
    d.crop = int(crop)
    d.adjust_mask = int(adjust_mask)
    d.dry_run = int(dry_run)
    d.convolver = CONVOLVERS[convolver]
    d.device = int(device)
    if psf is not None:
        psf = np.ascontiguousarray(psf, dtype=np.float64)
        keep.append(psf)
        d.psf = _d(psf)
        d.psf_h, d.psf_w = psf.shape
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        keep.append(mask)
        d.mask = _u8(mask)
        d.mask_h, d.mask_w = mask.shape
    return d


def make_model(width, height, profiles, psf=None, mask=None, magzero=0.0, scale=(1.0, 1.0),
               finesampling=1, return_finesampled=True, crop=True, adjust_mask=True,
               convolver="brute", device=0, dry_run=False, psf_scale=(1.0, 1.0)):
    """Model::evaluate on the GPU.  Returns dict(image=[y, x] array, offset=(x, y))."""
    keep = []
    d = _model_desc(width, height, psf, mask, magzero, scale, finesampling, return_finesampled,
                    crop, adjust_mask, convolver, device, dry_run, keep, psf_scale)
    arr = profile_array(profiles)
    L = lib()
    dw, dh = C.c_uint32(), C.c_uint32()
    _check(L.profit_amd_output_dims(C.byref(d), arr, len(profiles), C.byref(dw), C.byref(dh)))
    cap = max(1, dw.value * dh.value)
    out = np.zeros(cap)
    ow, oh, ox, oy = C.c_uint32(), C.c_uint32(), C.c_int32(), C.c_int32()
    _check(L.profit_amd_make_model(C.byref(d), arr, len(profiles), _d(out), cap, C.byref(ow),
                                   C.byref(oh), C.byref(ox), C.byref(oy)))
    img = out[: ow.value * oh.value].reshape(oh.value, ow.value).copy()
    return {"image": img, "offset": (ox.value, oy.value)}


def clear_cache():
    """Releases the device set-ups kept for repeated make_model calls."""
    _check(lib().profit_amd_clear_cache())


def adjust_mask(mask, img_dims, psf, finesampling=1):
    """Model::adjust(mask, dims, psf, finesampling); img_dims = (width, height)."""
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    psf = np.ascontiguousarray(psf, dtype=np.float64) if psf is not None else np.zeros((0, 0))
    mh, mw = mask.shape
    ph, pw = psf.shape if psf.size else (0, 0)
    cap = (mw * finesampling + pw) * (mh * finesampling + ph) + 16
    out = np.zeros(cap, dtype=np.uint8)
    ow, oh = C.c_uint32(), C.c_uint32()
    _check(lib().profit_amd_adjust_mask(_u8(mask), mw, mh, img_dims[0], img_dims[1],
                                        _d(psf) if psf.size else None, pw, ph, finesampling,
                                        _u8(out), cap, C.byref(ow), C.byref(oh)))
    return out[: ow.value * oh.value].reshape(oh.value, ow.value).copy()


def convolve(src, krn, mask=None, convolver="brute", device=0):
    """Convolver::convolve(src, krn, mask) on the GPU."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    krn = np.ascontiguousarray(krn, dtype=np.float64)
    out = np.zeros_like(src)
    m = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        m = _u8(mask)
    sh, sw = src.shape
    kh, kw = krn.shape
    _check(lib().profit_amd_convolve(CONVOLVERS[convolver], _d(src), sw, sh, _d(krn), kw, kh, m,
                                     _d(out), device))
    return out


def downsample(img, factor, mode="sum"):
    img = np.ascontiguousarray(img, dtype=np.float64)
    h, w = img.shape
    cap = ((w + factor - 1) // factor) * ((h + factor - 1) // factor)
    out = np.zeros(max(cap, 1))
    ow, oh = C.c_uint32(), C.c_uint32()
    _check(lib().profit_amd_downsample(_d(img), w, h, factor, {"sample": 0, "sum": 1, "average": 2}[mode],
                                       _d(out), cap, C.byref(ow), C.byref(oh)))
    return out[: ow.value * oh.value].reshape(oh.value, ow.value)


def upsample(img, factor, mode="copy"):
    img = np.ascontiguousarray(img, dtype=np.float64)
    h, w = img.shape
    cap = w * h * factor * factor
    out = np.zeros(max(cap, 1))
    ow, oh = C.c_uint32(), C.c_uint32()
    _check(lib().profit_amd_upsample(_d(img), w, h, factor, {"copy": 0, "scale": 1}[mode], _d(out), cap,
                                     C.byref(ow), C.byref(oh)))
    return out[: ow.value * oh.value].reshape(oh.value, ow.value)


class Engine:
    """Batched likelihood (+ adjoint gradient) engine on one GPU."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        _check(lib().profit_amd_engine_create(device, C.byref(self._h)))
        self._keep = []
        self.galaxies = []
        self.inst_ncomp = []

    def close(self):
        if self._h:
            lib().profit_amd_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_galaxy(self, width, height, data, sigma, psf=None, calcregion=None, region=None,
                   likefunc="chisq", magzero=0.0, scale=(1.0, 1.0), psf_scale=(1.0, 1.0)):
        g = GalaxyDesc()
        g.width, g.height = int(width), int(height)
        g.scale_x, g.scale_y = scale
        g.psf_scale_x, g.psf_scale_y = psf_scale
        g.magzero = magzero
        keep = []
        if psf is not None:
            psf = np.ascontiguousarray(psf, dtype=np.float64)
            keep.append(psf)
            g.psf = _d(psf)
            g.psf_h, g.psf_w = psf.shape
        if calcregion is not None:
            calcregion = np.ascontiguousarray(calcregion, dtype=np.uint8)
            keep.append(calcregion)
            g.calcregion = _u8(calcregion)
            g.calc_h, g.calc_w = calcregion.shape
        if region is not None:
            region = np.ascontiguousarray(region, dtype=np.uint8)
            keep.append(region)
            g.region = _u8(region)
        data = np.ascontiguousarray(data, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        keep += [data, sigma]
        g.data = _d(data)
        g.sigma = _d(sigma)
        g.likefunc = LIKE[likefunc]
        gid = C.c_int32()
        _check(lib().profit_amd_engine_add_galaxy(self._h, C.byref(g), C.byref(gid)))
        self._keep.append(keep)
        self.galaxies.append((int(width), int(height)))
        return gid.value

    def geometry(self, galaxy_id):
        dw, dh, px, py = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(lib().profit_amd_engine_galaxy_geometry(self._h, galaxy_id, C.byref(dw), C.byref(dh),
                                                       C.byref(px), C.byref(py)))
        return dw.value, dh.value, px.value, py.value

    def set_batch(self, inst_galaxy, inst_ncomp):
        g = np.ascontiguousarray(inst_galaxy, dtype=np.int32)
        n = np.ascontiguousarray(inst_ncomp, dtype=np.int32)
        _check(lib().profit_amd_engine_set_batch(self._h, g.ctypes.data_as(C.POINTER(C.c_int32)),
                                                 n.ctypes.data_as(C.POINTER(C.c_int32)), len(g)))
        self.inst_galaxy = list(g)
        self.inst_ncomp = list(n)

    def set_tile_rows(self, rows):
        _check(lib().profit_amd_engine_set_tile_rows(self._h, rows))

    def set_render_fp64(self, on=True):
        """Model precision: False / 0 FP32 profile values (default), True / 1
        FP64 values, 2 FP64 values, FP64 canvas and FP64 FFT convolution."""
        _check(lib().profit_amd_engine_set_render_fp64(self._h, int(on)))

    def set_grad_slots(self, slots=None):
        """Gradient slots to compute (names from GRAD_SLOTS or indices; None =
        all); the others come back as 0."""
        if slots is None:
            mask = (1 << NGRAD) - 1
        else:
            mask = 0
            for s in slots:
                mask |= 1 << (GRAD_SLOTS.index(s) if isinstance(s, str) else int(s))
        _check(lib().profit_amd_engine_set_grad_slots(self._h, mask))

    def set_conv_method(self, method="brute"):
        """'brute' (tiled FP32 brute force), 'fft' (FP64 FFT) or 'brute-bf16x3' (matrix
        cores, split-BF16x3) for the galaxies with a PSF."""
        _check(lib().profit_amd_engine_set_conv_method(self._h, CONVOLVERS[method]))

    def autotune(self, params, reps=2, allow_bf16x3=False):
        """Per-galaxy convolution kernels for the current batch; returns the
        chosen convolver name per galaxy and the timings (ms per evaluation).
        The split-BF16x3 kernel is a candidate only with allow_bf16x3."""
        arr = self._params(params)
        ngal = len(self.galaxies)
        kinds = (C.c_int32 * max(1, ngal))()
        ms = (C.c_double * 6)()
        _check(lib().profit_amd_engine_autotune(self._h, arr, int(reps), int(bool(allow_bf16x3)), kinds, ms))
        names = {v: k for k, v in CONVOLVERS.items()}
        return dict(kinds=[names[kinds[i]] for i in range(ngal)],
                    ms=dict(zip(("brute_rows4", "brute_rows8", "brute_rows16", "fft", "brute_bf16x3", "chosen"),
                                list(ms))))

    def _params(self, params):
        if isinstance(params, C.Array):
            return params
        return profile_array(list(params))

    def like(self, params):
        arr = self._params(params)
        ll = np.zeros(len(self.inst_ncomp))
        _check(lib().profit_amd_engine_like(self._h, arr, _d(ll)))
        return ll

    def like_grad(self, params):
        arr = self._params(params)
        ll = np.zeros(len(self.inst_ncomp))
        g = np.zeros(int(sum(self.inst_ncomp)) * NGRAD)
        _check(lib().profit_amd_engine_like_grad(self._h, arr, _d(ll), _d(g)))
        return ll, g.reshape(-1, NGRAD)

    def model_images(self, params):
        arr = self._params(params)
        sizes = [self.galaxies[g][0] * self.galaxies[g][1] for g in self.inst_galaxy]
        out = np.zeros(sum(sizes), dtype=np.float32)
        _check(lib().profit_amd_engine_model_images(self._h, arr,
                                                    out.ctypes.data_as(C.POINTER(C.c_float))))
        imgs, off = [], 0
        for g, n in zip(self.inst_galaxy, sizes):
            w, h = self.galaxies[g]
            imgs.append(out[off:off + n].reshape(h, w))
            off += n
        return imgs

    def set_timing(self, on=True):
        _check(lib().profit_amd_engine_set_timing(self._h, int(on)))

    def reset_stats(self):
        _check(lib().profit_amd_engine_reset_stats(self._h))

    def stats(self):
        ms = np.zeros(len(STAGES))
        calls = np.zeros(len(STAGES), dtype=np.int64)
        ev, q = C.c_int64(), C.c_int32()
        cf, af = C.c_double(), C.c_double()
        _check(lib().profit_amd_engine_stats(self._h, _d(ms), calls.ctypes.data_as(C.POINTER(C.c_int64)),
                                             C.byref(ev), C.byref(q), C.byref(cf), C.byref(af)))
        return {"stage_ms": dict(zip(STAGES, ms.tolist())), "stage_calls": dict(zip(STAGES, calls.tolist())),
                "evaluations": ev.value, "queue_length": q.value, "conv_flops": cf.value,
                "adjoint_flops": af.value}

    def adjusted_settings(self, params):
        """Per component: resolution, max_recursions, acc, rscale_switch and
        rscale_max as adjusted by the set-up."""
        arr = self._params(params)
        nc = int(sum(self.inst_ncomp))
        res = np.zeros(max(1, nc), dtype=np.int32)
        mr = np.zeros(max(1, nc), dtype=np.int32)
        acc, rsw, rmax = np.zeros(max(1, nc)), np.zeros(max(1, nc)), np.zeros(max(1, nc))
        _check(lib().profit_amd_engine_adjusted_settings(self._h, arr, res.ctypes.data_as(C.POINTER(C.c_int32)),
                                                         mr.ctypes.data_as(C.POINTER(C.c_int32)), _d(acc), _d(rsw),
                                                         _d(rmax)))
        return dict(resolution=res[:nc], max_recursions=mr[:nc], acc=acc[:nc], rscale_switch=rsw[:nc],
                    rscale_max=rmax[:nc])

    def fft_stats(self):
        """Milliseconds per timed evaluation of each FFT kernel: forward rows,
        columns, inverse rows of the likelihood stage and of the adjoint stage."""
        ms = np.zeros(6)
        calls = np.zeros(2, dtype=np.int64)
        _check(lib().profit_amd_engine_fft_stats(self._h, _d(ms), calls.ctypes.data_as(C.POINTER(C.c_int64))))
        names = ("rows_fwd", "cols", "rows_inv")
        out = {}
        for f, stage in enumerate(("like", "adj")):
            for k, n in enumerate(names):
                out[f"{stage}_{n}"] = ms[3 * f + k] / calls[f] if calls[f] else 0.0
        return out

    def diagnostics(self, params):
        arr = self._params(params)
        dw, dh, _, _ = self.geometry(self.inst_galaxy[0])
        nc = int(sum(self.inst_ncomp))
        cls = np.zeros(nc * dw * dh, dtype=np.int32)
        nodes = np.zeros(nc * dw * dh, dtype=np.int32)
        _check(lib().profit_amd_engine_diagnostics(self._h, arr,
                                                   cls.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   nodes.ctypes.data_as(C.POINTER(C.c_int32))))
        return cls.reshape(nc, dh, dw), nodes.reshape(nc, dh, dw)
