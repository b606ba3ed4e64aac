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

// 16-byte global accesses for shard data.  NT: non-temporal (no reuse
// expected in the caches) -- right where every byte is read or written once
// per launch (the generated encode programs); the table kernel re-reads
// inputs per output tile and the tower interpreter shares 128-byte lines
// between neighbouring workgroups through L2, so they use plain accesses.
typedef unsigned int trs_v4u __attribute__((ext_vector_type(4)));

template <bool NT = false>
__device__ __forceinline__ uint4 ld16(const uint8_t *p) {
  trs_v4u v;
  if (NT)
    v = __builtin_nontemporal_load(reinterpret_cast<const trs_v4u *>(p));
  else
    v = *reinterpret_cast<const trs_v4u *>(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}

template <bool NT = false>
__device__ __forceinline__ void st16(uint8_t *p, uint4 x) {
  const trs_v4u v = {x.x, x.y, x.z, x.w};
  if (NT)
    __builtin_nontemporal_store(v, reinterpret_cast<trs_v4u *>(p));
  else
    *reinterpret_cast<trs_v4u *>(p) = v;
}

}  // namespace trs
