// Allocation-free brute-force convolution (noalloc.hip): limits and entry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace profit_amd {

// limits of the fixed scratch array (about 8.7 MB per device)
constexpr int kNoAllocMaxSide = 4096;               // source pixels per side
constexpr size_t kNoAllocMaxPixels = size_t(1) << 20;  // source pixels
constexpr int kNoAllocMaxKernel = 255;              // kernel pixels per side
constexpr int kNoAllocMaxTiles = 2048;              // 64 x 16 output tiles (at most ~1300 within the limits above)

enum NoAllocMethod { NOALLOC_BRUTE = 0, NOALLOC_NULL = 1 };

// Enqueues on `st`: out = the brute-force convolution of src (h rows of w
// doubles, device memory) with krn (kh rows of kw, device memory), zero where
// mask (device memory, w*h bytes; may be null) is 0.  Returns after
// enqueueing.  Calls on one device share its scratch array: calls that
// overlap on one device race.
void convolve_noalloc(int method, const double *src, int w, int h, const double *krn, int kw, int kh,
                      const uint8_t *mask, double *out, int device, hipStream_t st);

}  // namespace profit_amd
