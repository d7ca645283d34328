"""Config 5, the precision stress field: 2048 x 2048, 500 Sersic sources,
max_recursions = 4, 101 x 101 Moffat PSF, rendered in FP32 (the FP32 value
option) and convolved three ways (FP32 brute force, split-BF16x3 brute force
on the matrix cores, FP64 FFT), each against the FP64 oracle: maximum
relative pixel error above 1e-6 of the peak, total-flux error and logL error.

FP32 brute force and FFT are held to the parity gates (1e-5 per pixel, 1e-3
absolute in logL) and to 1e-6 in total flux; split-BF16x3 is recorded, with a
looser pixel bound."""
import numpy as np
import pytest

import oracle as O
import profit_amd as P
from profit_amd import workloads as W
from tests.helpers import gpu_profiles, oracle_model_fn, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def field():
    return W.config5(oracle_model_fn(threads=16))


@pytest.mark.parametrize("method", ["brute", "fft", "brute-bf16x3"])
def test_precision_stress_field(field, method):
    wl = field
    oll = O.loglike_image("chisq", wl["data"], wl["sigma"], wl["model"], wl["region"])
    eng = P.Engine()
    eng.set_render_fp64(False)
    eng.set_conv_method(method)
    g = eng.add_galaxy(2048, 2048, wl["data"], wl["sigma"], psf=wl["psf"], region=wl["region"], magzero=30.0)
    eng.set_batch([g], [len(wl["components"])])
    params = gpu_profiles(wl["components"])
    model = eng.model_images(params)[0]
    ll = eng.like(params)[0]
    err, n = rel_err(model, wl["model"])
    flux = abs(model.astype(np.float64).sum() / wl["model"].sum() - 1)
    print(f"[precision] {method}: max rel pixel err {err:.3e} over {n} px, total-flux err {flux:.3e}, "
          f"logL {ll:.6f} vs {oll:.6f} (diff {ll - oll:.3e})", flush=True)
    if method == "brute-bf16x3":
        assert err <= 1e-4
        return
    assert err <= 1e-5
    assert flux <= 1e-6
    assert abs(ll - oll) <= 1e-3
