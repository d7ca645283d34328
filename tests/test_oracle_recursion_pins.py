"""Pins of the oracle's sub-pixel recursion (radial.cpp:54-134) and of the
Sersic centre-pixel rule (sersic.cpp:331-347) by what the mathematics fixes,
plus flux conservation and the adjusted settings of the King, core-Sersic and
broken-exponential profiles.

* acc = 0 refines every cell down to max_recursions, so a pixel is the plain
  midpoint rule on a uniform grid of res^(L+1) x res^(L+1) cells: this fixes
  the child-cell geometry, the level weights and the depth limit.
* acc huge refines nothing, so a pixel is the res x res midpoint rule for any
  max_recursions: with the first case it fixes the direction of the test
  |test - val| > acc * val.
* The centre rule: with acc -> 0 at resolution 8 and depth 10 the centre
  pixels converge to the exact pixel integral (scipy dblquad); with
  adjust = true exactly the pixels within one pixel of the centre take
  resolution 8 and depth 10, every other pixel its adjusted resolution and
  max_recursions.

The profile values here come from an independent numpy Sersic, not from the
oracle."""
import math

import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as ss

import oracle as O


def _sersic_numpy(x, y, xcen, ycen, re, n, ang, axrat, bn):
    angrad = math.fmod(ang + 90, 360.0) * math.pi / 180
    c, s = math.cos(angrad), math.sin(angrad)
    xr, yr = x - xcen, y - ycen
    xp = xr * c + yr * s
    yp = (-xr * s + yr * c) / axrat
    return np.exp(-bn * (np.sqrt((xp * xp + yp * yp) / (re * re)) ** (1 / n) - 1))


def _grid_mean(i, j, cells, fn):
    h = 1.0 / cells
    xs = i + h * (np.arange(cells) + 0.5)
    ys = j + h * (np.arange(cells) + 0.5)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return fn(X, Y).mean()


SERSIC = dict(xcen=20.3, ycen=19.8, mag=12.0, re=4.0, nser=1.7, ang=33.0, axrat=0.8)


def _oracle_image(res, maxrec, acc, size=40, **over):
    p = O.make_profile("sersic", adjust=False, resolution=res, max_recursions=maxrec, acc=acc,
                       rscale_switch=100.0, **dict(SERSIC, **over))
    spec = O.ModelSpec(size, size, threads=8)
    return O.model_evaluate(spec, [p], diag=True), O.profile_info(p)


def _numpy_fn(info, **over):
    s = dict(SERSIC, **over)
    return lambda X, Y: _sersic_numpy(X, Y, s["xcen"], s["ycen"], s["re"], s["nser"], s["ang"], s["axrat"],
                                      info["bn"])


@pytest.mark.parametrize("maxrec", [1, 2, 3])
def test_zero_accuracy_refines_every_cell_to_the_depth_limit(maxrec):
    res = 4
    out, info = _oracle_image(res, maxrec, 0.0)
    img, nodes = out["image"], out["nodes"][0]
    fn = _numpy_fn(info)
    cells = res ** (maxrec + 1)
    # every node of a full tree of depth maxrec: 1 + 16 + ... + 16^maxrec
    full_tree = sum((res * res) ** lv for lv in range(maxrec + 1))
    for (i, j) in [(20, 19), (18, 22), (23, 17), (25, 25)]:
        val = _grid_mean(i, j, cells, fn)
        assert nodes[j, i] == full_tree, (i, j, nodes[j, i])
        assert abs(img[j, i] / info["flux_scale"] - val) <= 1e-12 * val, (maxrec, i, j)
    # one level short (a depth limit off by one) is far from the oracle
    i, j = 20, 19
    short = _grid_mean(i, j, res ** maxrec, fn)
    assert abs(img[j, i] / info["flux_scale"] - short) > 1e-8 * short


@pytest.mark.parametrize("maxrec", [1, 3])
def test_huge_accuracy_refines_nothing(maxrec):
    res = 6
    out, info = _oracle_image(res, maxrec, 1e30)
    img, nodes = out["image"], out["nodes"][0]
    fn = _numpy_fn(info)
    for (i, j) in [(20, 19), (18, 22), (5, 7)]:
        val = _grid_mean(i, j, res, fn)
        assert nodes[j, i] == 1
        assert abs(img[j, i] / info["flux_scale"] - val) <= 1e-12 * val


def test_refinement_happens_only_where_the_profile_varies():
    """A moderate accuracy refines the steep cells near the centre and leaves the
    flat outskirts as single nodes: a flipped inequality would invert both."""
    out, info = _oracle_image(4, 3, 0.2)
    nodes = out["nodes"][0]
    assert nodes[19, 20] > 1                        # next to the centre
    assert nodes[2, 2] == 1 and nodes[37, 37] == 1  # far out, where f is flat within a pixel


CENTRE = dict(xcen=10.3, ycen=10.6, mag=12.0, re=2.0, nser=4.0, ang=20.0, axrat=0.9)


def test_centre_rule_depth_converges_to_the_pixel_integral():
    """Resolution 8 and depth 10 (the centre rule's settings) as acc -> 0: the
    centre pixels of an n = 4 cusp converge to the exact pixel integral
    (dblquad), the error falling with acc.  Pixels within 2 px of the centre
    are sub-sampled (rscale_switch = 1 re)."""
    s = CENTRE
    bn = O.profile_info(O.make_profile("sersic", **CENTRE))["bn"]
    f = lambda y, x: float(_sersic_numpy(np.array(x), np.array(y), s["xcen"], s["ycen"], s["re"], s["nser"],
                                          s["ang"], s["axrat"], bn))
    pix = [(9, 10), (10, 10), (9, 11), (10, 11)]
    exact = [si.dblquad(f, i, i + 1, j, j + 1, epsabs=1e-13, epsrel=1e-11)[0] for (i, j) in pix]
    worst = []
    for acc in (0.1, 0.03, 0.01):
        p = O.make_profile("sersic", adjust=False, resolution=8, max_recursions=10, acc=acc, rscale_switch=1.0,
                           **CENTRE)
        img = O.model_evaluate(O.ModelSpec(21, 21, threads=8), [p])["image"]
        fs = O.profile_info(p)["flux_scale"]
        worst.append(max(abs(img[j, i] / fs - e) / e for (i, j), e in zip(pix, exact)))
    print("centre-pixel errors for acc 0.1, 0.03, 0.01:", worst)
    assert worst[0] > worst[1] > worst[2]
    assert worst[2] <= 1e-4


def test_centre_rule_applies_to_the_pixels_within_one_pixel_of_the_centre():
    p_adj = O.make_profile("sersic", **CENTRE)  # adjust = true
    info = O.profile_info(p_adj)
    fixed = dict(acc=info["acc"], rscale_switch=info["rscale_switch"], rscale_max=info["rscale_max"])
    spec = O.ModelSpec(21, 21, threads=8)
    adj = O.model_evaluate(spec, [p_adj], diag=True)
    deep = O.model_evaluate(spec, [O.make_profile("sersic", adjust=False, resolution=8, max_recursions=10,
                                                  **fixed, **CENTRE)])["image"]
    plain = O.model_evaluate(spec, [O.make_profile("sersic", adjust=False, resolution=info["resolution"],
                                                   max_recursions=2, **fixed, **CENTRE)])["image"]
    # pixel centres within one pixel of (10.3, 10.6): x in {9.5, 10.5}, y in {10.5, 11.5}
    centre = {(i, j) for i in range(21) for j in range(21)
              if abs(i + 0.5 - CENTRE["xcen"]) < 1 and abs(j + 0.5 - CENTRE["ycen"]) < 1}
    assert centre == {(9, 10), (10, 10), (9, 11), (10, 11)}
    sub = adj["cls"][0] == 3  # sub-pixel integration
    for j in range(21):
        for i in range(21):
            if not sub[j, i]:
                continue
            want = deep[j, i] if (i, j) in centre else plain[j, i]
            assert adj["image"][j, i] == want, (i, j)
    # the rule matters there: depth 10 and depth 2 differ in the pixel holding the centre
    assert abs(deep[10, 10] - plain[10, 10]) > 1e-6 * deep[10, 10]


@pytest.mark.parametrize("name, shape, tol", [
    ("coresersic", dict(re=10.0, rb=1.5, nser=1.0, a=2.0, b=0.5), 1e-3),
    ("brokenexp", dict(h1=6.0, h2=3.0, rb=10.0, a=2.0), 1e-3),
])
def test_flux_conservation_of_the_integrated_profiles(name, shape, tol):
    """Their luminosity is a QUADPACK integral to infinity (rscale_max = 100
    scale radii): the image holds 10^(-0.4 mag).  Beyond rscale_switch (one
    scale radius) pixels are sampled at their centres, so the scale radii are
    large enough for that midpoint rule to hold the flux to 1e-3 (at re = 3 px
    the curvature of the profile costs 0.6%, in libprofit as here)."""
    p = O.make_profile(name, xcen=100.3, ycen=99.6, mag=10.0, ang=20.0, axrat=0.7, **shape)
    img = O.model_evaluate(O.ModelSpec(200, 200, magzero=0.0, threads=8), [p])["image"]
    total = 10 ** (-0.4 * 10.0)
    assert abs(img.sum() / total - 1) <= tol, (name, img.sum() / total)


def _res(rsw, rscale):
    r = math.ceil(160 / (rsw * rscale))
    r += r % 2
    return min(max(r, 4), 16)


def test_adjusted_settings_of_king_coresersic_brokenexp():
    """king.cpp:99-109, coresersic.cpp:119-129, brokenexponential.cpp:129-139 with
    radial.cpp:158-185: rscale, rscale_switch, resolution, rscale_max; acc
    unchanged."""
    for rt in (3.0, 20.0, 70.0):
        k = O.profile_info(O.make_profile("king", rc=2.0, rt=rt, a=2.0, acc=0.07))
        assert k["rscale"] == rt and k["rscale_switch"] == 0.5 and k["rscale_max"] == 1.5
        assert k["resolution"] == _res(0.5, rt) and k["acc"] == 0.07
    for re in (0.9, 4.0, 30.0):
        c = O.profile_info(O.make_profile("coresersic", re=re, rb=0.5, nser=2.0, a=2.0, b=0.5, acc=0.07))
        assert c["rscale"] == re and c["rscale_switch"] == 1 and c["rscale_max"] == 100
        assert c["resolution"] == _res(1, re) and c["acc"] == 0.07
        assert abs(c["bn"] - ss.gammaincinv(4.0, 0.5)) <= 1e-13 * c["bn"]
    for h1 in (0.7, 3.0, 15.0):
        b = O.profile_info(O.make_profile("brokenexp", h1=h1, h2=2.0, rb=5.0, a=1.0, acc=0.07))
        assert b["rscale"] == 4 * h1 and b["rscale_switch"] == 1 and b["rscale_max"] == 100
        assert b["resolution"] == _res(1, 4 * h1) and b["acc"] == 0.07
