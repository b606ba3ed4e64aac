// Wave-uniform loads (device code).
//
// Values read from global memory land in vector registers even when every
// lane reads the same address, and the compiler then treats them as
// divergent: a coefficient's bit masks are rebuilt per lane and loops and
// branches on job fields run under exec masks.  readfirstlane moves such a
// value into a scalar register, which is exact whenever the value is the same
// in every active lane -- true for everything passed through these helpers
// (job descriptors, program steps and coefficient words addressed by
// workgroup-uniform indices).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace trs {

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint64_t uni(uint64_t x) {
  return (uint64_t(uni(uint32_t(x >> 32))) << 32) | uni(uint32_t(x));
}

template <class T>
__device__ __forceinline__ T *uni(T *p) {
  return reinterpret_cast<T *>(uni(uint64_t(reinterpret_cast<uintptr_t>(p))));
}

}  // namespace trs
