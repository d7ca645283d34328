"""Batches of a single radial family: the engine then renders with the kernel
compiled for that family alone (k_render<TYPE>), which only a like/like_grad
evaluation uses.  Each family's batch of two instances is checked against the
FP64 oracle: logL within 1e-3 and the gradient finite and the same for the
two identical instances up to the grouping of the adjoint walker's FP32
partial sums (per grab of queue entries, whose boundaries differ between
the two instances) and the order of the atomic sums: 1e-6 relative."""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import gpu_profiles, oracle_profiles

pytestmark = pytest.mark.gpu

COMMON = dict(xcen=30.3, ycen=29.6, mag=15.0, ang=35.0, axrat=0.7, box=0.0)
FAMILIES = {
    "sersic": dict(re=5.0, nser=2.5),
    "moffat": dict(fwhm=4.0, con=2.5),
    "ferrer": dict(rout=12.0, a=1.0, b=0.5),
    "king": dict(rc=3.0, rt=20.0, a=2.0),
    "coresersic": dict(re=6.0, rb=1.5, nser=3.0, a=2.0, b=0.5),
    "brokenexp": dict(h1=5.0, h2=2.5, rb=8.0, a=3.0),
}


@pytest.mark.parametrize("kind", sorted(FAMILIES))
def test_single_family_batch(kind):
    size = 60
    psf = W.gaussian_psf(9, 2.0)
    comps = [dict(type=kind, **COMMON, **FAMILIES[kind])]
    spec = O.ModelSpec(size, size, psf=psf, magzero=30.0, threads=8)
    model = O.model_evaluate(spec, oracle_profiles(comps))["image"]
    sigma = np.sqrt(np.maximum(model, 0) + 25.0)
    data = model + W.normal_draws(5, size * size).reshape(size, size) * sigma
    fit = O.FitSpec(spec, data, sigma, np.ones((size, size), dtype=bool))
    oll = O.loglike(fit, oracle_profiles(comps))
    eng = P.Engine()
    g = eng.add_galaxy(size, size, data, sigma, psf=psf, magzero=30.0)
    eng.set_batch([g, g], [1, 1])
    params = gpu_profiles(comps + comps)
    ll = eng.like(params)
    ll2, grad = eng.like_grad(params)
    print(f"[family] {kind}: logL {ll[0]:.6f} / {ll2[0]:.6f} vs oracle {oll:.6f}", flush=True)
    assert abs(ll[0] - oll) <= 1e-3 and abs(ll[1] - oll) <= 1e-3
    assert abs(ll2[0] - oll) <= 1e-3
    assert np.all(np.isfinite(grad)) and np.allclose(grad[0], grad[1], rtol=1e-6, atol=1e-9 * np.abs(grad).max())
