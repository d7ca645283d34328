"""FITS reader (profit_amd.fits): round trips of every BITPIX with scaling
and extensions, and the example cut-outs ProFit ships (fixtures under
tests/golden/examples: 10 KiDS and 10 SDSS galaxies)."""
import glob
import os

import numpy as np
import pytest

from profit_amd import fits

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = os.path.join(HERE, "golden", "examples")
KIDS = os.path.join(EXAMPLES, "KiDS")


@pytest.mark.parametrize("bitpix", [8, 16, 32, 64, -32, -64])
def test_round_trip(tmp_path, bitpix):
    rng = np.random.default_rng(bitpix & 0xff)
    if bitpix > 0:
        hi = {8: 255, 16: 30000, 32: 2 ** 30, 64: 2 ** 40}[bitpix]
        data = rng.integers(0, hi, size=(37, 53))
    else:
        data = rng.normal(size=(37, 53)) * 1e-10
    p = tmp_path / "x.fits"
    fits.write_image(p, data, bitpix=bitpix)
    assert os.path.getsize(p) % 2880 == 0
    back = fits.read_image(p)
    assert back.shape == (37, 53)
    if bitpix == -32:
        assert np.array_equal(back, data.astype(np.float32).astype(np.float64))
    else:
        assert np.array_equal(back, data)


def test_scaling_and_extensions(tmp_path):
    raw = np.arange(12, dtype=np.int64).reshape(3, 4)
    p = tmp_path / "s.fits"
    fits.write_image(p, raw, bitpix=16, extra={"BSCALE": 0.5, "BZERO": 100.0})
    assert np.array_equal(fits.read_image(p), 100.0 + 0.5 * raw)
    # an empty primary HDU followed by two image extensions
    q = tmp_path / "e.fits"
    fits.write_image(q, None, bitpix=-64, extensions=[np.ones((2, 5)), np.full((4, 3), 7.0)])
    hdus = fits.read_hdus(q)
    assert hdus[0][1] is None and hdus[1][1].shape == (2, 5) and hdus[2][1].shape == (4, 3)
    assert np.array_equal(fits.read_image(q), np.ones((2, 5)))
    assert np.array_equal(fits.read_image(q, hdu=2), np.full((4, 3), 7.0))


def test_not_fits(tmp_path):
    p = tmp_path / "n.fits"
    p.write_bytes(b"hello" * 1000)
    with pytest.raises(fits.FitsError):
        fits.read_image(p)


def test_kids_fixture():
    img = fits.read_image(os.path.join(KIDS, "G267525fitim.fits"))
    sig = fits.read_image(os.path.join(KIDS, "G267525sigma.fits"))
    seg = fits.read_image(os.path.join(KIDS, "G267525segim.fits"))
    psf = fits.read_image(os.path.join(KIDS, "G267525psfim.fits"))
    # NAXIS1 = 101, NAXIS2 = 153: [y, x] arrays
    assert img.shape == sig.shape == seg.shape == (153, 101)
    assert psf.shape == (25, 25)
    assert np.all(np.isfinite(img)) and np.all(sig > 0)
    assert np.array_equal(seg, np.round(seg)) and seg.min() >= 0
    # the PSF peaks at its centre pixel
    assert np.unravel_index(np.argmax(psf), psf.shape) == (12, 12)


def test_every_example_image_reads():
    paths = sorted(glob.glob(os.path.join(EXAMPLES, "*", "*.fits")))
    assert len(paths) == 80
    for p in paths:
        hdus = fits.read_hdus(p)
        assert sum(1 for _, d in hdus if d is not None) == 1, p
        assert np.all(np.isfinite(fits.read_image(p))), p
