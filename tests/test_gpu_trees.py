"""Pixel classes and sub-pixel recursion trees, GPU against the FP64 oracle,
bit-exact (engine diagnostics through the C-ABI), beyond config 1:

* config 2 (bulge n = 4, re = 3 and disk n = 1 over the adjusted elliptical
  calc region, 450 x 450 canvas);
* a 512 x 512 window of config 5 (the field's own source draws: n 0.5-8,
  re 0.5-20 px, max_recursions = 4) with its 101 x 101 PSF padding;
* King and Ferrer profiles cut off at rt / rout inside the image, where the
  refinement test compares values that cancel towards zero.

Class 0: outside the calc region, 1: beyond rscale_max, 2: direct, 3:
sub-pixel integration; nodes: the number of nodes of each pixel's tree."""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import gpu_profiles, oracle_profiles

pytestmark = pytest.mark.gpu


def _compare(tag, w, h, psf, comps, calc=None):
    spec = O.ModelSpec(w, h, psf=psf, mask=calc, adjust_mask=calc is None, threads=8)
    ref = O.model_evaluate(spec, oracle_profiles(comps), diag=True)
    eng = P.Engine()
    zeros = np.zeros((h, w))
    g = eng.add_galaxy(w, h, zeros, zeros + 1.0, psf=psf, calcregion=calc)
    eng.set_batch([g], [len(comps)])
    cls, nodes = eng.diagnostics(gpu_profiles(comps))
    rc, rn = ref["cls"].astype(np.int32), ref["nodes"].astype(np.int32)
    assert cls.shape == rc.shape
    bad_c = int(np.sum(cls != rc))
    bad_n = int(np.sum(nodes != rn))
    counts = np.bincount(rc.ravel(), minlength=4)
    print(f"[trees] {tag}: {len(comps)} components, classes {counts.tolist()}, nodes {int(rn.sum())}, "
          f"deepest {int(rn.max())}; class mismatches {bad_c}, node mismatches {bad_n}", flush=True)
    assert bad_c == 0 and bad_n == 0


def test_config2_trees_bit_exact():
    wl = W.config2()
    calc = O.adjust_mask(wl["region"], (400, 400), wl["psf"])
    _compare("config 2", 400, 400, wl["psf"], wl["components"], calc=calc)


def test_config5_window_trees_bit_exact():
    wl = W.config5(nsrc=500)
    win = 512
    comps = [dict(c) for c in wl["components"] if c["xcen"] < win and c["ycen"] < win]
    assert len(comps) >= 20
    _compare(f"config 5 window {win}x{win}", win, win, wl["psf"], comps)


# the adjusted settings sub-sample only within half the cut-off radius; with
# adjust = false and rscale_switch beyond the cut, cells straddle rt / rout
EDGE = dict(adjust=False, rscale_switch=1.2, rscale_max=1.5, resolution=8, max_recursions=3, acc=0.1)
TRUNCATED = [
    dict(type="king", xcen=40.3, ycen=39.6, mag=15.0, ang=30.0, axrat=0.7, box=0.0, rc=2.5, rt=17.0, a=2.0),
    dict(type="king", xcen=40.7, ycen=40.2, mag=15.0, ang=75.0, axrat=0.55, box=0.0, rc=1.2, rt=9.0, a=1.4),
    dict(type="ferrer", xcen=40.2, ycen=39.9, mag=15.0, ang=10.0, axrat=0.8, box=0.0, rout=14.0, a=1.0, b=0.5),
    dict(type="ferrer", xcen=39.6, ycen=40.4, mag=15.0, ang=120.0, axrat=0.6, box=0.0, rout=8.0, a=0.3, b=1.2),
    dict(type="king", xcen=40.3, ycen=39.6, mag=15.0, ang=30.0, axrat=0.7, box=0.0, rc=2.5, rt=17.0, a=2.0, **EDGE),
    dict(type="king", xcen=40.7, ycen=40.2, mag=15.0, ang=75.0, axrat=0.55, box=0.0, rc=1.2, rt=9.0, a=1.4, **EDGE),
    dict(type="ferrer", xcen=40.2, ycen=39.9, mag=15.0, ang=10.0, axrat=0.8, box=0.0, rout=14.0, a=1.0, b=0.5,
         **EDGE),
    dict(type="ferrer", xcen=39.6, ycen=40.4, mag=15.0, ang=120.0, axrat=0.6, box=0.0, rout=8.0, a=0.3, b=1.2,
         **EDGE),
]


@pytest.mark.parametrize("k", range(len(TRUNCATED)))
def test_truncated_profile_trees_bit_exact(k):
    comp = dict(TRUNCATED[k])
    _compare(f"{comp['type']} {k}", 80, 80, W.gaussian_psf(9, 2.0), [comp])
