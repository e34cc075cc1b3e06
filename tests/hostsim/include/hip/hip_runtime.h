// Host stand-in for the HIP runtime header, used only to compile the engine's
// device code (gemini_unified_amd/csrc/gm_device.inc, gm_kernels.inc) as
// ordinary C++ for the CPU tests: every kernel then runs as a plain function,
// one block at a time with one thread per block (blockDim.x = 1), so that
// __syncthreads() and the block-shared scalars need no emulation.
#pragma once
#include <math.h>
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__
#define __launch_bounds__(...)

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint4 { unsigned x, y, z, w; };
typedef void* hipStream_t;

// block and thread index of the kernel that runs; set by the host driver
extern dim3 threadIdx, blockIdx, blockDim, gridDim;

inline void __syncthreads() {}
template <class T> inline T atomicAdd(T* p, T v) { T o = *p; *p = o + v; return o; }
template <class T> inline T atomicMax(T* p, T v) { T o = *p; if (v > o) *p = v; return o; }

// the hardware reciprocal estimate, here exact (the fast variant refines it by
// two Newton steps anyway); a wave-uniform value is the lane's own value
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
#define __builtin_amdgcn_readfirstlane(x) (x)

using std::max;
using std::min;
