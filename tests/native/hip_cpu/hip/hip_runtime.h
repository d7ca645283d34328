// CPU stand-in for the parts of the HIP device API that the kernels use, so
// that kernel code can run on the host (one std::thread per GPU thread, a
// std::barrier per block) under AddressSanitizer: any access outside a
// device buffer (allocated here with malloc, at its exact size) is reported.
// Used by the bounds tests only; never by the library.
#pragma once

#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__

struct dim3 {
	unsigned x = 0, y = 0, z = 0;
};
inline thread_local dim3 threadIdx, blockIdx;
inline dim3 blockDim, gridDim;

struct double2 {
	double x, y;
};
struct float2 {
	float x, y;
};
struct int2 {
	int x, y;
};
struct uint4 {
	unsigned x, y, z, w;
};
inline double2 make_double2(double x, double y) { return double2{x, y}; }
inline int2 make_int2(int x, int y) { return int2{x, y}; }

typedef void *hipStream_t;
typedef int hipError_t;
constexpr int hipSuccess = 0;

namespace hip_cpu {
inline std::barrier<> *block_barrier = nullptr;
inline double xchg_d[1024];

// runs `kernel` over `grid` blocks of `block` threads, blocks one after
// another; the same `block` host threads run every block, with a barrier
// between blocks
inline void launch(unsigned grid, unsigned block, const std::function<void()> &kernel)
{
	gridDim = dim3{grid, 1, 1};
	blockDim = dim3{block, 1, 1};
	std::barrier<> bar((std::ptrdiff_t)block);
	block_barrier = &bar;
	std::vector<std::thread> ts;
	for (unsigned t = 0; t < block; ++t)
		ts.emplace_back([&, t] {
			threadIdx = dim3{t, 0, 0};
			for (unsigned b = 0; b < grid; ++b) {
				blockIdx = dim3{b, 0, 0};
				kernel();
				bar.arrive_and_wait();
			}
		});
	for (auto &t : ts) t.join();
}
}  // namespace hip_cpu

inline void __syncthreads() { hip_cpu::block_barrier->arrive_and_wait(); }
inline unsigned __brev(unsigned v)
{
	unsigned r = 0;
	for (int i = 0; i < 32; ++i) r |= ((v >> i) & 1u) << (31 - i);
	return r;
}
inline int __clz(int v)
{
	int n = 0;
	for (int i = 31; i >= 0 && !(((unsigned)v >> i) & 1u); --i) ++n;
	return n;
}
inline int __ffs(int v) { return __builtin_ffs(v); }
// the transcendental-unit log2 / exp2 of the FP32 profile values
inline float __builtin_amdgcn_logf(float x) { return std::log2(x); }
inline float __builtin_amdgcn_exp2f(float x) { return std::exp2(x); }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline float __builtin_amdgcn_rcpf(float x) { return 1.f / x; }
inline double __shfl_xor(double v, int mask)
{
	const unsigned t = threadIdx.x;
	hip_cpu::xchg_d[t] = v;
	__syncthreads();
	const double r = hip_cpu::xchg_d[t ^ (unsigned)mask];
	__syncthreads();
	return r;
}
// lane of the wavefront and a read of another lane's value
inline int __lane_id() { return (int)(threadIdx.x & 63u); }
inline double __shfl(double v, int src)
{
	const unsigned t = threadIdx.x;
	hip_cpu::xchg_d[t] = v;
	__syncthreads();
	const double r = hip_cpu::xchg_d[(t & ~63u) + ((unsigned)src & 63u)];
	__syncthreads();
	return r;
}
