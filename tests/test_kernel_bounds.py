"""Kernel index arithmetic checked on the CPU: the FFT kernels run through a
host stand-in of the HIP thread model (one thread per GPU thread, a barrier
per block) under AddressSanitizer, with every device buffer at its exact size,
for several geometries.  An access outside a buffer fails the test; with a
unit-impulse PSF the convolution modes must reproduce their input, and with a
three-tap PSF (one tap at a wrapped negative offset) the model image and the
adjoint correlation must equal their direct sums."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "profit_amd", "csrc")
MARK = "// ------------------------------------------------------------------ launchers"


def _build(tmp, name):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    src = open(os.path.join(CSRC, "fft.hip")).read()
    assert MARK in src
    with open(os.path.join(tmp, "fft_kernels.inc"), "w") as f:
        f.write(src[:src.index(MARK)] + "\n}  // namespace profit_amd\n")
    exe = os.path.join(tmp, name)
    r = subprocess.run([gxx, "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-pthread",
                        "-I" + os.path.join(ROOT, "tests", "native", "hip_cpu"), "-I" + CSRC, "-I" + tmp,
                        os.path.join(ROOT, "tests", "native", name + ".cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def test_fft_kernels_stay_in_bounds(tmp_path):
    exe = _build(str(tmp_path), "fft_bounds")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.count("ok ") == 11, r.stdout
