"""The set-up's QUADPACK integrals spread over a wavefront (each lane one
Kronrod point, every lane forming the sums in the serial order): the set-up
(`setup_component<true>`) and the luminosity gradients (`lum_gradient<true>`)
of King, core-Sersic and broken-exponential components, and the set-up of
Sersic components (b_n and the adjust rule's two gamma quantiles on three
lanes), must equal the serial ones bit for bit.  Built against the host stand-in of the HIP thread model
(tests/native/hip_cpu), one block of 64 threads; no GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_wave_quadrature_is_bitwise_the_serial_one(tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "quad_wave")
    r = subprocess.run([gxx, "-std=c++20", "-O2", "-pthread",
                        "-I" + os.path.join(ROOT, "tests", "native", "hip_cpu"),
                        "-I" + os.path.join(ROOT, "profit_amd", "csrc"),
                        os.path.join(ROOT, "tests", "native", "quad_wave.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300).stdout
    rows = [ln.split() for ln in out.splitlines() if ln.strip()]
    assert len(rows) == 13 and {r[0] for r in rows} == {"sersic", "king", "coresersic", "brokenexp"}, out
    for name, case, same, gsame, *_ in rows:
        assert same == "1" and gsame == "1", (name, case, out)
