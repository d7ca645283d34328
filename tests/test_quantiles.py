"""The product's gamma quantile (profit_amd/csrc/special.h, host build), as
the Sersic set-up uses it: b_n = qgamma(0.5, 2n) and the switch and maximum
radii's quantiles 1 - n^2/2000 and 0.9999 (sersic.cpp adjust), against
scipy's inverse regularised incomplete gamma.  The upper-tail quantiles stop
once the Newton step is below the resolution of x; they must still agree to
a few ulps."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy import special

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def rows(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("quantile_pin") / "quantile_pin")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O2",
                    "-I" + os.path.join(ROOT, "profit_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "quantile_pin.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return np.array([[float(v) for v in ln.split()] for ln in out.splitlines() if ln.strip()])


def test_quantiles_match_scipy(rows):
    p, a, x, back = rows.T
    ref = special.gammaincinv(a, p)
    rel = np.abs(x - ref) / ref
    assert len(rows) > 600
    assert rel.max() < 2e-14, (rel.max(), rows[np.argmax(rel)])


def test_quantiles_invert_pgamma(rows):
    p, a, x, back = rows.T
    # P(a, x) at the quantile is p up to the conditioning of P near x (scipy's
    # own gammainc at these x is 1.2e-14 from p at a = 36)
    assert np.abs(back - p).max() < 5e-14
