"""Pins of the FP64 oracle against independent references (CPU only).

The reference (R + libprofit) ships no tests or golden outputs and cannot be
built here (its sources need the configure-generated config.h and R's nmath),
so the oracle is pinned by: scipy (gamma family; QUADPACK through
scipy.integrate.quad, the same dqagse/dqagie routines), closed forms, brute
force (numpy/scipy.signal/scipy.ndimage) and invariants.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.integrate as si
import scipy.ndimage as ndi
import scipy.signal as ssig
import scipy.special as sp
import scipy.stats as st

import oracle as O
from profit_amd import workloads as W

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "special_golden.json")))


# ---------------------------------------------------------------- special functions

def test_qgamma_matches_scipy_golden():
    for p, a, ref in GOLDEN["qgamma"]:
        got = O.qgamma(p, a)
        assert abs(got - ref) <= 1e-13 * ref, (p, a, got, ref)


def test_bn_closed_forms():
    # b_{1/2} = ln 2 ; b_1 solves (1 + x) e^-x = 1/2
    assert abs(O.qgamma(0.5, 1.0) - math.log(2)) <= 1e-15
    b1 = O.qgamma(0.5, 2.0)
    assert abs((1 + b1) * math.exp(-b1) - 0.5) <= 1e-15
    assert abs(b1 - 1.678346990016661) <= 1e-14


def test_pgamma_gamma_beta_golden():
    for x, a, ref in GOLDEN["pgamma"]:
        assert abs(O.pgamma(x, a) - ref) <= 1e-14 + 1e-13 * ref
    for x, ref in GOLDEN["gamma"]:
        assert abs(O.gammafn(x) - ref) <= 1e-14 * ref
    for a, b, ref in GOLDEN["beta"]:
        assert abs(O.beta(a, b) - ref) <= 1e-13 * ref
    # B(1/2, 1/2) = pi, so r_box(box=0) = 1
    assert abs(O.beta(0.5, 0.5) - math.pi) <= 1e-14


def test_digamma_and_shape_derivative():
    for x in [0.3, 1.0, 2.5, 9.7, 31.0]:
        assert abs(O.digamma(x) - sp.digamma(x)) <= 1e-13 * max(1, abs(sp.digamma(x)))
    for a in [1.0, 5.0, 8.0]:
        for x in [0.5, 3.0, 7.669]:
            h = 1e-5
            fd = (sp.gammainc(a + h, x) - sp.gammainc(a - h, x)) / (2 * h)
            assert abs(O.dpgamma_da(x, a) - fd) <= 1e-8


def test_quadpack_nodes_are_gauss_kronrod():
    xk21, wk21, wg10 = (O.quadpack_nodes(i) for i in range(3))
    gx, gw = np.polynomial.legendre.leggauss(10)
    assert np.allclose(np.sort(np.concatenate([xk21[1:10:2], -xk21[1:10:2]])), np.sort(gx), atol=1e-15)
    assert np.allclose(np.sort(np.concatenate([wg10, wg10])), np.sort(gw), atol=1e-15)
    # the 21-point Kronrod rule integrates polynomials exactly up to degree 31
    nodes = np.concatenate([-xk21[:10], xk21[:10], [0.0]])
    weights = np.concatenate([wk21[:10], wk21[:10], [wk21[10]]])
    for k in range(0, 32):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(weights, nodes ** k) - exact) <= 1e-14, k
    xk15, wk15, wg7 = (O.quadpack_nodes(i) for i in range(3, 6))
    gx, gw = np.polynomial.legendre.leggauss(7)
    gnodes = np.concatenate([xk15[1:7:2], -xk15[1:7:2], [0.0]])
    assert np.allclose(np.sort(gnodes), np.sort(gx), atol=1e-15)
    nodes = np.concatenate([-xk15[:7], xk15[:7], [0.0]])
    weights = np.concatenate([wk15[:7], wk15[:7], [wk15[7]]])
    for k in range(0, 24):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(weights, nodes ** k) - exact) <= 1e-14, k


QUAD_CASES = [
    (lambda x: math.exp(-x) * x, 0.0, 5.0),
    (lambda x: 1 / math.sqrt(x) if x > 0 else 0.0, 0.0, 1.0),            # end-point singularity: extrapolation
    (lambda x: math.log(x) if x > 0 else 0.0, 0.0, 1.0),
    (lambda x: math.sin(30 * x) * math.exp(-x), 0.0, 3.0),
    (lambda x: x * (1 / (1 + (x / 1.5) ** 2) ** 0.5 - 1 / (1 + (10 / 1.5) ** 2) ** 0.5) ** 2, 0.0, 10.0),
]
QUAD_INF_CASES = [
    lambda r: r * math.exp(-r),
    lambda r: r / (1 + r * r) ** 3,
    lambda r: 1 / (1 + r) ** 1.3,                                          # slow decay
    lambda r: r * math.exp(-r / 1.3 + (1 / 1.3 - 1 / 3.0) * (math.log1p(math.exp(2 * (r - 4))) / 2
                                                             if 2 * (r - 4) < 40 else (r - 4))),
]


@pytest.mark.parametrize("case", range(len(QUAD_CASES)))
def test_qags_bitwise_equal_to_scipy_quadpack(case):
    f, a, b = QUAD_CASES[case]
    got, info = O.integrate(f, a, b)
    ref = si.quad(f, a, b, epsabs=1e-4, epsrel=1e-4, limit=100, full_output=1)
    assert got == ref[0]
    assert info["neval"] == ref[2]["neval"] and info["last"] == ref[2]["last"]


@pytest.mark.parametrize("case", range(len(QUAD_INF_CASES)))
def test_qagi_bitwise_equal_to_scipy_quadpack(case):
    f = QUAD_INF_CASES[case]
    got, info = O.integrate(f, 0.0)
    ref = si.quad(f, 0, np.inf, epsabs=1e-4, epsrel=1e-4, limit=100, full_output=1)
    assert got == ref[0]
    assert info["neval"] == ref[2]["neval"] and info["last"] == ref[2]["last"]


def test_numerical_luminosities_match_scipy_golden():
    for g in GOLDEN["quad_finite"]:
        rc, rt, a = g["args"]
        info = O.profile_info(O.make_profile("king", rc=rc, rt=rt, a=a))
        assert abs(info["lumtot"] / (2 * math.pi) - g["value"]) <= 1e-12 * g["value"]
    for g in GOLDEN["quad_infinite"]:
        if g["kind"] == "coresersic":
            re, rb, nser, a, b = g["args"]
            p = O.make_profile("coresersic", re=re, rb=rb, nser=nser, a=a, b=b)
        else:
            h1, h2, rb, a = g["args"]
            p = O.make_profile("brokenexp", h1=h1, h2=h2, rb=rb, a=a)
        info = O.profile_info(p)
        assert abs(info["lumtot"] / (2 * math.pi) - g["value"]) <= 1e-10 * g["value"], g


def test_closed_form_luminosities():
    re, n = 8.0, 2.5
    bn = sp.gammaincinv(2 * n, 0.5)
    lum = re ** 2 * 2 * math.pi * n * sp.gamma(2 * n) * math.exp(bn) / bn ** (2 * n)
    assert abs(O.profile_info(O.make_profile("sersic", re=re, nser=n))["lumtot"] - lum) <= 1e-12 * lum
    fwhm, con = 4.0, 3.0
    rs = fwhm / (2 * math.sqrt(2 ** (1 / con) - 1))
    info = O.profile_info(O.make_profile("moffat", fwhm=fwhm, con=con))
    assert abs(info["rscale"] - rs) <= 1e-14 * rs
    assert abs(info["lumtot"] - math.pi * rs ** 2 / (con - 1)) <= 1e-12 * info["lumtot"]
    rout, a, b = 10.0, 1.5, 0.7
    lf = rout ** 2 * math.pi * a * sp.beta(a, 1 + 2 / (2 - b))
    assert abs(O.profile_info(O.make_profile("ferrer", rout=rout, a=a, b=b))["lumtot"] - lf) <= 1e-12 * lf
    # King with a = 2 has an elementary r f(r) antiderivative check through quad
    rc, rt = 2.0, 20.0
    exact = si.quad(lambda r: r * (1 / math.sqrt(1 + (r / rc) ** 2) - 1 / math.sqrt(1 + (rt / rc) ** 2)) ** 2,
                    0, rt, epsabs=1e-13, epsrel=1e-13)[0]
    lk = O.profile_info(O.make_profile("king", rc=rc, rt=rt, a=2.0))["lumtot"] / (2 * math.pi)
    assert abs(lk - exact) <= 1e-6 * exact


def test_adjusted_settings_of_the_benchmark_components():
    """libprofit adjust=true rewrites resolution, rscale_switch, rscale_max, acc."""
    info = O.profile_info(O.make_profile("sersic", re=8, nser=2.5, axrat=0.6))
    assert abs(info["bn"] - 4.670908882795985) <= 1e-14
    assert info["rscale_switch"] * 8 == 20.0 and info["resolution"] == 8
    assert info["rscale_max"] == 29.0 and abs(info["acc"] - 0.1 / 0.6) <= 1e-15
    bulge = O.profile_info(O.make_profile("sersic", re=3, nser=4, axrat=0.9))
    disk = O.profile_info(O.make_profile("sersic", re=15, nser=1, axrat=0.4))
    assert abs(bulge["bn"] - 7.669249442500804) <= 1e-14
    assert bulge["rscale_max"] * 3 == 243 and disk["rscale_max"] * 15 == 120
    assert abs(bulge["acc"] - 0.1 / 0.9) <= 1e-15 and abs(disk["acc"] - 0.25) <= 1e-15
    # res = clamp(even(ceil(160 / (rsw * rscale))), 4, 16)
    m = O.profile_info(O.make_profile("moffat", fwhm=1.0, con=2.0))
    assert m["rscale_switch"] * m["rscale"] == 20.0 and m["resolution"] == 8
    f = O.profile_info(O.make_profile("ferrer", rout=10.0))
    assert f["rscale_switch"] == 0.5 and f["resolution"] == 16


# ---------------------------------------------------------------- rendering

def _sersic_numpy(x, y, xcen, ycen, re, n, ang, axrat, bn):
    angrad = math.fmod(ang + 90, 360.0) * math.pi / 180
    c, s = math.cos(angrad), math.sin(angrad)
    xr, yr = x - xcen, y - ycen
    xp = xr * c + yr * s
    yp = (-xr * s + yr * c) / axrat
    return np.exp(-bn * (np.sqrt((xp * xp + yp * yp) / (re * re)) ** (1 / n) - 1))


def test_flux_conservation():
    """Sum of an uncut profile = 10^(-0.4 (mag - magzero)) (within the
    99.99% rscale_max truncation and pixelisation)."""
    cases = [("sersic", dict(re=4.0, nser=1.0), 5e-4), ("sersic", dict(re=3.0, nser=2.5), 1e-3),
             ("moffat", dict(fwhm=3.0, con=3.0), 5e-4), ("ferrer", dict(rout=12.0, a=1.5, b=0.5), 3e-3),
             ("king", dict(rc=2.0, rt=20.0, a=2.0), 1e-3)]
    for name, shape, tol in cases:
        p = O.make_profile(name, xcen=100.3, ycen=99.6, mag=10.0, ang=20.0, axrat=0.7, acc=0.1, **shape)
        img = O.model_evaluate(O.ModelSpec(200, 200, magzero=0.0, threads=8), [p])["image"]
        total = 10 ** (-0.4 * 10.0)
        assert abs(img.sum() / total - 1) <= tol, (name, img.sum() / total)


def test_midpoint_rule_without_recursion():
    """max_recursions = 0: a pixel is the mean of its res x res cell centres."""
    res = 6
    p = O.make_profile("sersic", xcen=20.3, ycen=19.8, mag=12, re=4.0, nser=1.7, ang=33, axrat=0.8,
                       adjust=False, resolution=res, max_recursions=0, rscale_switch=100.0)
    img = O.model_evaluate(O.ModelSpec(40, 40), [p])["image"]
    info = O.profile_info(p)
    for (i, j) in [(20, 19), (18, 22), (5, 7), (33, 30)]:
        h = 1.0 / res
        xs = i + h * (np.arange(res) + 0.5)
        ys = j + h * (np.arange(res) + 0.5)
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        val = _sersic_numpy(X, Y, 20.3, 19.8, 4.0, 1.7, 33, 0.8, info["bn"]).mean()
        assert abs(img[j, i] / info["flux_scale"] - val) <= 1e-12 * val


def test_recursion_converges_to_pixel_integral():
    """With a tight accuracy the adaptive recursion approaches the exact pixel integral."""
    p = O.make_profile("sersic", xcen=10.5, ycen=10.5, mag=12, re=2.0, nser=1.0, ang=0, axrat=1.0,
                       adjust=False, resolution=4, max_recursions=5, rscale_switch=100.0, acc=1e-3,
                       rscale_max=1.6)
    img = O.model_evaluate(O.ModelSpec(21, 21), [p])["image"]
    info = O.profile_info(p)
    for (i, j) in [(10, 10), (11, 10), (12, 12)]:
        f = lambda y, x: float(_sersic_numpy(np.array(x), np.array(y), 10.5, 10.5, 2.0, 1.0, 0, 1.0, info["bn"]))
        exact = si.dblquad(f, i, i + 1, j, j + 1, epsabs=1e-12, epsrel=1e-10)[0]
        assert abs(img[j, i] / info["flux_scale"] - exact) <= 2e-4 * exact


def test_rotation_and_mirror_invariance():
    base = dict(xcen=30.0, ycen=30.0, mag=12, re=5.0, nser=2.0, axrat=0.5)
    a = O.model_evaluate(O.ModelSpec(60, 60), [O.make_profile("sersic", ang=25.0, **base)])["image"]
    b = O.model_evaluate(O.ModelSpec(60, 60), [O.make_profile("sersic", ang=205.0, **base)])["image"]
    assert np.allclose(a, b, rtol=1e-10, atol=0)
    c = O.model_evaluate(O.ModelSpec(60, 60), [O.make_profile("sersic", ang=0.0, **base)])["image"]
    assert np.allclose(c, c[::-1, :], rtol=1e-10, atol=0)
    assert np.allclose(c, c[:, ::-1], rtol=1e-10, atol=0)


def test_gaussian_convolved_with_gaussian():
    """Sersic n=1/2 is a Gaussian (b = ln 2); convolved with a Gaussian PSF it is
    a Gaussian with the variances added."""
    re = 3.0
    s_src = re / math.sqrt(2 * math.log(2))
    s_psf = 1.5
    psf = W.gaussian_psf(31, s_psf / W.GAUSS_FWHM_TO_SIGMA)
    p = O.make_profile("sersic", xcen=40.0, ycen=40.0, mag=10.0, re=re, nser=0.5, convolve=True)
    img = O.model_evaluate(O.ModelSpec(80, 80, psf=psf, threads=8), [p])["image"]
    s2 = s_src ** 2 + s_psf ** 2
    # pixel integrals of the analytic Gaussian
    edges = np.arange(81) - 40.0
    cdf = st.norm.cdf(edges / math.sqrt(s2))
    px = np.diff(cdf)
    ref = 10 ** (-4.0) * np.outer(px, px)
    core = ref > 1e-3 * ref.max()
    assert np.max(np.abs(img[core] - ref[core]) / ref[core]) < 2e-2


# ---------------------------------------------------------------- convolution and masks

@pytest.mark.parametrize("k", [1, 3, 7, 25])
def test_brute_convolution_is_2d_convolution(k):
    rng = np.random.default_rng(k)
    src = rng.random((37, 29))
    krn = rng.random((k, k))
    ref = ssig.convolve2d(src, krn, mode="same")
    for conv in ("brute", "brute-old", "fft"):
        assert np.allclose(O.convolve(src, krn, convolver=conv), ref, rtol=1e-13, atol=1e-13)
    mask = (rng.random(src.shape) > 0.5).astype(np.uint8)
    got = O.convolve(src, krn, mask=mask)
    assert np.all(got[mask == 0] == 0)
    assert np.allclose(got[mask == 1], ref[mask == 1], rtol=1e-13, atol=1e-13)


def test_delta_kernel_is_identity_and_small_sources_extend():
    rng = np.random.default_rng(3)
    src = rng.random((20, 20))
    delta = np.zeros((9, 9))
    delta[4, 4] = 1
    assert np.array_equal(O.convolve(src, delta), src)
    small = rng.random((5, 5))
    krn = rng.random((9, 9))
    ref = ssig.convolve2d(small, krn, mode="same")
    assert np.allclose(O.convolve(small, krn), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("pad", [(0, 0), (1, 2), (12, 12), (25, 25)])
def test_expand_by_is_box_dilation(pad):
    rng = np.random.default_rng(sum(pad))
    m = (rng.random((60, 50)) > 0.97).astype(np.uint8)
    ref = ndi.binary_dilation(m, structure=np.ones((2 * pad[1] + 1, 2 * pad[0] + 1), dtype=bool))
    assert np.array_equal(O.expand_by(m, pad), ref.astype(np.uint8))


def test_adjust_mask_padding_decision():
    psf = np.ones((11, 11))
    central = np.zeros((100, 100), dtype=np.uint8)
    central[40:60, 30:70] = 1
    adj = O.adjust_mask(central, (100, 100), psf)
    assert adj.shape == (100, 100)  # mask far from the edges: no padding
    ref = ndi.binary_dilation(central, structure=np.ones((11, 11), dtype=bool))
    assert np.array_equal(adj, ref.astype(np.uint8))
    edge = np.zeros((100, 100), dtype=np.uint8)
    edge[0:10, 0:10] = 1
    adj = O.adjust_mask(edge, (100, 100), psf)
    assert adj.shape == (110, 110)  # padded by psf/2 on each side, then dilated
    ext = np.zeros((110, 110), dtype=np.uint8)
    ext[5:105, 5:105] = edge
    ref = ndi.binary_dilation(ext, structure=np.ones((11, 11), dtype=bool))
    assert np.array_equal(adj, ref.astype(np.uint8))
    wl = W.config2()
    assert O.adjust_mask(wl["region"], (400, 400), wl["psf"]).shape == (450, 450)


def test_finesampling_sums_back():
    p = O.make_profile("sersic", xcen=20.2, ycen=19.9, mag=10, re=3.0, nser=1.0)
    coarse = O.model_evaluate(O.ModelSpec(40, 40, finesampling=3, return_finesampled=False), [p])["image"]
    fine = O.model_evaluate(O.ModelSpec(40, 40, finesampling=3, return_finesampled=True), [p])["image"]
    assert fine.shape == (120, 120) and coarse.shape == (40, 40)
    assert np.allclose(fine.reshape(40, 3, 40, 3).sum(axis=(1, 3)), coarse, rtol=1e-12)
    assert abs(coarse.sum() / 1e-4 - 1) < 2e-3


# ---------------------------------------------------------------- likelihoods

def test_likelihoods_match_scipy():
    rng = np.random.default_rng(7)
    n = 500
    model = rng.random(n) * 10 + 1
    sigma = np.sqrt(model + 4)
    data = model + rng.normal(size=n) * sigma
    z = (data - model) / sigma
    region = (rng.random(n) > 0.2).astype(np.uint8)
    r = region == 1
    assert abs(O.loglike_image("norm", data, sigma, model, region) - st.norm.logpdf(z[r]).sum()) < 1e-9
    X = (z[r] ** 2).sum()
    assert abs(O.loglike_image("chisq", data, sigma, model, region) - st.chi2.logpdf(X, r.sum())) < 1e-9
    v = np.var(z[r], ddof=1)
    dof = max(1.0, 2 * v / (v - 1)) if v > 1 else 1.0
    assert abs(O.loglike_image("t", data, sigma, model, region) - st.t.logpdf(z[r], dof).sum()) < 1e-8
    s = np.max(np.abs(data[r]) / sigma[r] ** 2)
    sc = s if (s < 0.1 or s > 10) else 1.0
    ref = -np.sum(model[r] * sc - data[r] * sc * np.log(model[r] * sc))
    assert abs(O.loglike_image("pois", data, sigma, model, region) - ref) < 1e-9


def test_likelihood_gradients_match_finite_differences():
    rng = np.random.default_rng(8)
    n = 50
    model = rng.random(n) * 10 + 1
    sigma = np.sqrt(model + 4)
    data = model + rng.normal(size=n) * sigma
    for lf in ("norm", "chisq", "pois"):
        _, g = O.loglike_image(lf, data, sigma, model, want_grad=True)
        for k in range(0, n, 7):
            h = 1e-6 * model[k]
            mp, mm = model.copy(), model.copy()
            mp[k] += h
            mm[k] -= h
            fd = (O.loglike_image(lf, data, sigma, mp) - O.loglike_image(lf, data, sigma, mm)) / (2 * h)
            assert abs(g[k] - fd) <= 1e-6 * max(1, abs(fd)), (lf, k)


@pytest.mark.parametrize("noise", [1.5, 0.7])
def test_t_likelihood_gradient_matches_finite_differences(noise):
    """The t likelihood's dof follows the sample variance of the residuals
    (R/profitLikeModel.R:281-286): residuals 1.5 sigma wide give dof = 2v/(v-1)
    ~ 3.6, whose dependence on every pixel enters the gradient; 0.7 sigma wide
    clamp dof at 1, where it does not."""
    rng = np.random.default_rng(9)
    n = 60
    model = rng.random(n) * 10 + 1
    sigma = np.sqrt(model + 4)
    data = model + noise * rng.normal(size=n) * sigma
    z = (data - model) / sigma
    v = np.var(z, ddof=1)
    assert (v > 1) == (noise > 1)
    _, g = O.loglike_image("t", data, sigma, model, want_grad=True)
    for k in range(n):
        h = 1e-6 * model[k]
        mp, mm = model.copy(), model.copy()
        mp[k] += h
        mm[k] -= h
        fd = (O.loglike_image("t", data, sigma, mp) - O.loglike_image("t", data, sigma, mm)) / (2 * h)
        assert abs(g[k] - fd) <= 1e-6 * max(1, abs(fd)), k
    # the dof term matters: dropping it moves the unclamped gradient
    if noise > 1:
        dof = 2 * v / (v - 1)
        fixed = -(-(dof + 1) * z / (dof + z * z)) / sigma
        assert np.max(np.abs(fixed - g)) > 1e-3 * np.max(np.abs(g))


# ---------------------------------------------------------------- gradient

def _small_fit(components, magzero=30.0, size=48, psf=None):
    from tests.helpers import oracle_profiles
    spec = O.ModelSpec(size, size, psf=psf, magzero=magzero, threads=8)
    model = O.model_evaluate(spec, oracle_profiles(components, convolve=psf is not None))["image"]
    sigma = np.sqrt(np.maximum(model, 0) + 25.0)
    data = model + W.normal_draws(5, size * size).reshape(size, size) * sigma
    fit = O.FitSpec(spec, data, sigma, None, "chisq")
    return fit


@pytest.mark.parametrize("family", ["sersic", "moffat", "ferrer", "king"])
def test_oracle_gradient_matches_frozen_central_differences(family):
    from tests.helpers import oracle_profiles
    shapes = {"sersic": dict(re=4.0, nser=2.2), "moffat": dict(fwhm=3.5, con=2.5),
              "ferrer": dict(rout=9.0, a=1.4, b=0.6), "king": dict(rc=2.0, rt=14.0, a=2.2)}
    comp = dict(type=family, xcen=24.3, ycen=23.6, mag=14.0, ang=35.0, axrat=0.7, box=0.0, **shapes[family])
    psf = W.gaussian_psf(9, 2.0)
    fit = _small_fit([comp], psf=psf)
    pert = dict(comp)
    pert.update(xcen=24.5, ycen=23.3, mag=13.9, ang=40.0, axrat=0.66)
    for k in shapes[family]:
        pert[k] = shapes[family][k] * 1.05
    profs = oracle_profiles([pert])
    ll, g = O.loglike_grad(fit, profs)
    nshape = len(shapes[family])
    slots = np.zeros((1, O.NGRAD), dtype=np.uint8)
    slots[0, [0, 1, 2, 3, 4] + [6 + i for i in range(nshape)]] = 1
    llf, gf = O.loglike_grad_fd(fit, profs, rel_step=1e-6, slots=slots, frozen=True)
    assert abs(ll - llf) < 1e-12 * abs(ll)
    for s in np.nonzero(slots[0])[0]:
        assert abs(g[0, s] - gf[0, s]) <= 1e-5 * max(abs(gf[0, s]), 1e-3 * np.abs(gf[0]).max()), \
            (family, O.GRAD_SLOTS[s], g[0, s], gf[0, s])


def test_box_gradient_matches_frozen_central_differences():
    from tests.helpers import oracle_profiles
    comp = dict(type="sersic", xcen=24.3, ycen=23.6, mag=14.0, ang=35.0, axrat=0.7, box=0.3, re=4.0, nser=1.6)
    fit = _small_fit([comp])
    profs = oracle_profiles([dict(comp, box=0.25)], convolve=False)
    ll, g = O.loglike_grad(fit, profs)
    slots = np.zeros((1, O.NGRAD), dtype=np.uint8)
    slots[0, 5] = 1
    _, gf = O.loglike_grad_fd(fit, profs, rel_step=1e-6, slots=slots, frozen=True)
    assert abs(g[0, 5] - gf[0, 5]) <= 1e-5 * abs(gf[0, 5])
