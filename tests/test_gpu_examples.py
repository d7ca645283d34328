"""ProFit's example galaxies end to end: the 10 KiDS and 10 SDSS cut-outs it
ships (tests/golden/examples, read with profit_amd.fits), each set up as the
reference's galaxy-fit example does (vignettes/ProFit-Galaxy-Fit-Example.Rmd:
bulge + disk started from data/ExampleInit.tab, read from the file; tied
centres, logged sizes, intervals, profitMakePriors normal priors, the t
likelihood, magzero 0).

* all 20 galaxies as one heterogeneous batch on one engine (profit_amd.fit
  .MultiFit): the log posterior and its gradient in the fitted parameters
  against the same glue over the FP64 oracle, at the initial parameters and
  at a perturbed point.  Their LP reaches -8e4 with bright cores, where the
  FP32 canvas moves LP by up to a few 1e-3, so the batch runs in the FP64
  model mode (FP64 values, canvas and FFT) against the 1e-3 gate;
* KiDS G267525 on its own: the same check at three points, then a
  quasi-Newton fit on the adjoint gradient must raise LP."""
import os

import numpy as np
import pytest

from profit_amd import fit as F
from tests.helpers import OracleEvaluator

pytestmark = pytest.mark.gpu

EXAMPLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples")
INIT = F.read_example_init(os.path.join(EXAMPLES, "ExampleInit.tab"))
IDS = sorted(int(f[1:7]) for f in os.listdir(os.path.join(EXAMPLES, "KiDS")) if f.endswith("fitim.fits"))
GATE_LP = 1e-3
GATE_GRAD = 2e-3


def _check(tag, lp, g, olp, og):
    print(f"[examples] {tag}: LP gpu {lp:.6f} oracle {olp:.6f}", flush=True)
    print(f"[examples]   grad gpu    {np.array2string(g, precision=5)}", flush=True)
    print(f"[examples]   grad oracle {np.array2string(og, precision=5)}", flush=True)
    assert abs(lp - olp) <= GATE_LP, tag
    assert np.all(np.abs(g - og) <= GATE_GRAD * np.maximum(np.abs(og), 1e-2 * np.abs(og).max())), tag


def test_ten_galaxies_per_survey_as_one_batch():
    assert len(IDS) == 10
    keys = [(src, cid) for src in ("KiDS", "SDSS") for cid in IDS]
    gpu = [F.example_fit(EXAMPLES, src, cid, INIT, evaluator=lambda d: None) for src, cid in keys]
    ora = [F.example_fit(EXAMPLES, src, cid, INIT, evaluator=OracleEvaluator) for src, cid in keys]
    batch = F.MultiFit(gpu, precision=2)
    rng = np.random.default_rng(20)
    for point in range(2):
        parms = [d.init + (0 if point == 0 else rng.normal(size=len(d.init)) * 0.02) for d in gpu]
        res = batch.like_grad(parms)
        for (src, cid), o, p, (lp, g) in zip(keys, ora, parms, res):
            olp, og = o.like_grad(p)
            _check(f"{src} G{cid} point {point}", lp, g, olp, og)


def test_kids_g267525_posterior_gradient_and_fit():
    gpu = F.example_fit(EXAMPLES, "KiDS", 267525, INIT)
    ora = F.example_fit(EXAMPLES, "KiDS", 267525, INIT, evaluator=OracleEvaluator)
    assert gpu.N == ora.N > 4000
    rng = np.random.default_rng(3)
    for k in range(3):
        p = gpu.init + (0 if k == 0 else rng.normal(size=len(gpu.init)) * 0.02)
        lp, g = gpu.like_grad(p)
        olp, og = ora.like_grad(p)
        _check(f"KiDS G267525 point {k}", lp, g, olp, og)
    lp0 = gpu.like(gpu.init)
    res = gpu.optimise(maxiter=60)
    lp1 = gpu.like(res.x)
    print(f"[examples] KiDS G267525: LP {lp0:.3f} -> {lp1:.3f} after {res.nit} iterations ({res.nfev} "
          f"evaluations): {dict(zip(gpu.parm_names, np.round(res.x, 4)))}", flush=True)
    assert np.isfinite(lp1) and lp1 > lp0 + 10
