"""Synthetic workloads of the model-image path.

Each builder returns the inputs ProFit's profitSetupData would hold: image
size, PSF, likelihood region, component parameters and (given a function that
renders the noiseless model) data and sigma maps drawn from a seeded
counter-based generator, so any host regenerates identical inputs.

  config1()  100x100, one Sersic (n=2.5, re=8), 25x25 Gaussian PSF (FWHM 3),
             sigma = sqrt(model + 10^2), all-true mask, chi^2.
  config2()  400x400 bulge (n=4, re=3) + disk (n=1, re=15), 51x51 Moffat PSF
             (FWHM 4, con 3), elliptical region of ~70% of the pixels,
             sigma = sqrt(max(model, 0)/gain + 5^2), gain 1.
"""
from __future__ import annotations

import numpy as np

GAUSS_FWHM_TO_SIGMA = 1.0 / (2.0 * np.sqrt(2.0 * np.log(2.0)))


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uint64 draws of SplitMix64 seeded with ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def normal_draws(seed: int, n: int) -> np.ndarray:
    """Standard normal deviates by Box-Muller over SplitMix64 uniforms."""
    m = (n + 1) // 2
    u = splitmix64(seed, 2 * m)
    u1 = ((u[0::2] >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0
    u2 = ((u[1::2] >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def gaussian_psf(size: int, fwhm: float) -> np.ndarray:
    s = fwhm * GAUSS_FWHM_TO_SIGMA
    k = np.arange(size) - size // 2
    p = np.exp(-(k[None, :] ** 2 + k[:, None] ** 2) / (2 * s * s))
    return p / p.sum()


def moffat_psf(size: int, fwhm: float, con: float) -> np.ndarray:
    rs = fwhm / (2 * np.sqrt(2 ** (1 / con) - 1))
    k = np.arange(size) - size // 2
    r2 = (k[None, :] ** 2 + k[:, None] ** 2) / (rs * rs)
    p = (1 + r2) ** (-con)
    return p / p.sum()


def ellipse_mask(width: int, height: int, cx: float, cy: float, a: float, b: float) -> np.ndarray:
    x = np.arange(width) + 0.5
    y = np.arange(height) + 0.5
    return ((((x[None, :] - cx) / a) ** 2 + ((y[:, None] - cy) / b) ** 2) <= 1).astype(np.uint8)


def config1_components():
    return [dict(type="sersic", xcen=50.3, ycen=49.7, mag=15.0, re=8.0, nser=2.5, ang=30.0, axrat=0.6,
                 box=0.0)]


def config2_components():
    return [dict(type="sersic", xcen=200.4, ycen=199.8, mag=17.0, re=3.0, nser=4.0, ang=0.0, axrat=0.9, box=0.0),
            dict(type="sersic", xcen=200.4, ycen=199.8, mag=16.0, re=15.0, nser=1.0, ang=60.0, axrat=0.4,
                 box=0.0)]


def config1(model_fn=None, magzero=0.0, seed=0):
    """Config 1.  model_fn(width, height, psf, components, magzero) -> noiseless model."""
    w = h = 100
    psf = gaussian_psf(25, 3.0)
    comps = config1_components()
    region = np.ones((h, w), dtype=np.uint8)
    out = dict(width=w, height=h, psf=psf, components=comps, region=region, magzero=magzero,
               likefunc="chisq")
    if model_fn is not None:
        model = model_fn(w, h, psf, comps, magzero)
        sigma = np.sqrt(model + 10.0 ** 2)
        data = model + normal_draws(seed, w * h).reshape(h, w) * sigma
        out.update(model=model, data=data, sigma=sigma)
    return out


def config2(model_fn=None, magzero=30.0, seed=2, gain=1.0, read_noise=5.0):
    """Config 2 (magzero 30 by default so that the galaxy rises above the noise)."""
    w = h = 400
    psf = moffat_psf(51, 4.0, 3.0)
    comps = config2_components()
    region = ellipse_mask(w, h, 200.0, 200.0, 199.5, 178.7)
    out = dict(width=w, height=h, psf=psf, components=comps, region=region, magzero=magzero,
               likefunc="chisq")
    if model_fn is not None:
        model = model_fn(w, h, psf, comps, magzero)
        sigma = np.sqrt(np.maximum(model, 0) / gain + read_noise ** 2)
        data = model + normal_draws(seed, w * h).reshape(h, w) * sigma
        out.update(model=model, data=data, sigma=sigma)
    return out


GRAD_FREE = ("xcen", "ycen", "mag", "re", "nser", "ang", "axrat")


def uniforms(seed: int, n: int) -> np.ndarray:
    """n uniform deviates in (0, 1) from SplitMix64."""
    return ((splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0


CONFIG4_TYPES = ("sersic", "moffat", "ferrer", "king", "coresersic")


def _config4_component(kind, u, w, h):
    """One component from 8 uniforms: centre within +-2 px of the image centre,
    mag U(14, 20), ang U(0, 180), axrat U(0.2, 1), re U(1, 30), nser
    log-uniform in (0.5, 8).  Non-Sersic profiles take their scale radius from
    re and their shape from nser."""
    re = 1.0 + 29.0 * u[4]
    nser = float(np.exp(np.log(0.5) + (np.log(8.0) - np.log(0.5)) * u[5]))
    c = dict(type=kind, xcen=w / 2 + 4 * u[0] - 2, ycen=h / 2 + 4 * u[1] - 2, mag=14.0 + 6.0 * u[2],
             ang=180.0 * u[3], axrat=0.2 + 0.8 * u[6], box=0.0)
    if kind == "sersic":
        c.update(re=re, nser=nser)
    elif kind == "moffat":
        c.update(fwhm=re, con=1.5 + 0.5 * nser)
    elif kind == "ferrer":
        c.update(rout=2.0 * re, a=0.5 + nser / 4, b=0.5)
    elif kind == "king":
        c.update(rc=re / 4, rt=4.0 * re, a=2.0)
    else:
        c.update(re=re, rb=re / 4, nser=nser, a=2.0, b=0.5)
    return c


def config4_galaxies(n=4096, seed=4, types=CONFIG4_TYPES):
    """Config 4 geometry and parameters: n galaxies with sides uniform in
    32..256 px (width and height independent), 1-3 components of the given
    types, each galaxy with its own Gaussian or Moffat PSF of odd side 15..41
    px (FWHM U(2, 5), Moffat con U(2, 5)) and an elliptical likelihood region
    with semi-axes 0.48 of the sides.  No data: see config4_attach_data."""
    out = []
    for i in range(n):
        u = uniforms(seed * 1000003 + i, 64)
        w = 32 + int(u[0] * 225)
        h = 32 + int(u[1] * 225)
        ncomp = 1 + int(u[2] * 3)
        k = 15 + 2 * int(u[3] * 14)
        fwhm = 2.0 + 3.0 * u[4]
        psf = gaussian_psf(k, fwhm) if u[5] < 0.5 else moffat_psf(k, fwhm, 2.0 + 3.0 * u[6])
        comps = []
        for j in range(ncomp):
            v = u[8 + 8 * j: 16 + 8 * j]
            kind = types[int(v[7] * len(types)) % len(types)]
            comps.append(_config4_component(kind, v, w, h))
        region = ellipse_mask(w, h, w / 2, h / 2, 0.48 * w, 0.48 * h)
        out.append(dict(width=w, height=h, psf=psf, components=comps, region=region, magzero=30.0,
                        likefunc="chisq"))
    return out


def config4_attach_data(gals, models, seed=40, read_noise=5.0):
    """Adds sigma = sqrt(max(model, 0) + read_noise^2) and seeded noisy data."""
    for i, (g, m) in enumerate(zip(gals, models)):
        sigma = np.sqrt(np.maximum(m, 0) + read_noise ** 2)
        g.update(model=m, sigma=sigma,
                 data=m + normal_draws(seed * 1000003 + i, m.size).reshape(m.shape) * sigma)
    return gals


def config5(model_fn=None, seed=5, nsrc=500, magzero=30.0, read_noise=5.0):
    """Config 5, the precision stress field: 2048 x 2048 with 500 Sersic
    sources (nser log-uniform 0.5-8, re U(0.5, 20) px, mag U(16, 24),
    positions uniform, ang U(0, 180), axrat U(0.2, 1)), sub-pixel recursion
    with max_recursions = 4, a 101 x 101 Moffat PSF (FWHM 5, con 2.5), all
    pixels in the likelihood, sigma = sqrt(max(model, 0) + read_noise^2)."""
    w = h = 2048
    psf = moffat_psf(101, 5.0, 2.5)
    u = uniforms(seed * 7919 + 1, 8 * nsrc).reshape(nsrc, 8)
    comps = []
    for v in u:
        comps.append(dict(type="sersic", xcen=w * v[0], ycen=h * v[1], mag=16.0 + 8.0 * v[2],
                          re=0.5 + 19.5 * v[3], nser=float(np.exp(np.log(0.5) + np.log(16.0) * v[4])),
                          ang=180.0 * v[5], axrat=0.2 + 0.8 * v[6], box=0.0, max_recursions=4))
    region = np.ones((h, w), dtype=np.uint8)
    out = dict(width=w, height=h, psf=psf, components=comps, region=region, magzero=magzero, likefunc="chisq")
    if model_fn is not None:
        model = model_fn(w, h, psf, comps, magzero)
        sigma = np.sqrt(np.maximum(model, 0) + read_noise ** 2)
        data = model + normal_draws(seed, w * h).reshape(h, w) * sigma
        out.update(model=model, data=data, sigma=sigma)
    return out
