"""The allocation-free convolution entry point (profit_amd_convolve_noalloc)
against the allocating one (profit_amd_convolve), through the C-ABI, on two
calls in a row, the second with a smaller source: the two share their
convolution kernel and launch shape, so the results must be bit-identical.
Device buffers for its inputs and outputs come from the HIP runtime the
library links (libamdhip64), through ctypes."""
import ctypes as C

import numpy as np
import pytest

import profit_amd as P

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = C.c_int
    return hip


class DeviceArrays:
    """Device copies of host arrays (and device outputs), freed on close."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(1, nbytes)) == 0
        self.ptrs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(dev, convolver, src, krn, mask):
    h, w = src.shape
    kh, kw = krn.shape
    d_src = dev.upload(src.astype(np.float64))
    d_krn = dev.upload(krn.astype(np.float64))
    d_mask = dev.upload(mask.astype(np.uint8)) if mask is not None else None
    d_out = dev.alloc(src.size * 8)
    rc = P.lib().profit_amd_convolve_noalloc(P.CONVOLVERS[convolver], d_src, w, h, d_krn, kw, kh, d_mask, d_out, 0,
                                             None)
    P._check(rc)
    assert dev.hip.hipDeviceSynchronize() == 0
    return dev.download(d_out, src.shape, np.float64)


CASES = [
    # (source height, width, kernel side, masked) -- the second call smaller
    (257, 300, 25, True),
    (70, 90, 9, False),
]


def test_noalloc_matches_the_allocating_entry_point():
    hip = _hip()
    dev = DeviceArrays(hip)
    rng = np.random.default_rng(7)
    try:
        for convolver in ("brute", "brute-old"):
            for h, w, k, masked in CASES:
                src = rng.random((h, w))
                krn = rng.random((k, k))
                mask = (rng.random((h, w)) > 0.3).astype(np.uint8) if masked else None
                print(f"[noalloc] {convolver} {w}x{h} kernel {k} mask {masked}", flush=True)
                ref = P.convolve(src, krn, mask, convolver=convolver)
                got = _noalloc(dev, convolver, src, krn, mask)
                assert np.array_equal(got, ref), (convolver, h, w, k)
                if masked:
                    assert np.all(got[mask == 0] == 0)
        src = rng.random((40, 50))
        got = _noalloc(dev, "null", src, rng.random((5, 5)), None)
        assert np.array_equal(got, src)
    finally:
        dev.close()


def test_noalloc_rejects_what_exceeds_its_scratch():
    hip = _hip()
    dev = DeviceArrays(hip)
    try:
        d = dev.alloc(8)
        L = P.lib()
        # sizes are checked on the host before anything is enqueued
        assert L.profit_amd_convolve_noalloc(P.CONVOLVERS["brute"], d, 5000, 10, d, 3, 3, None, d, 0, None) != 0
        assert b"4096" in L.profit_amd_last_error()
        assert L.profit_amd_convolve_noalloc(P.CONVOLVERS["brute"], d, 2048, 1024, d, 3, 3, None, d, 0, None) != 0
        assert L.profit_amd_convolve_noalloc(P.CONVOLVERS["brute"], d, 10, 10, d, 300, 3, None, d, 0, None) != 0
        assert L.profit_amd_convolve_noalloc(P.CONVOLVERS["brute"], d, 10, 10, d, 11, 3, None, d, 0, None) != 0
        assert L.profit_amd_convolve_noalloc(P.CONVOLVERS["fft"], d, 10, 10, d, 3, 3, None, d, 0, None) != 0
    finally:
        dev.close()
