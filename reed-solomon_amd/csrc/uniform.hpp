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
//
// Shard pointers reach the kernels through pointer tables in memory, so the
// compiler cannot tell which address space they point to and emits flat
// accesses, which count against both the vector-memory and the scalar-memory
// wait counters: every wait for a scalar load then also drains the shard
// loads in flight.  Shards always live in device global memory, so these
// helpers access them through global (address space 1) pointers.
typedef unsigned int trs_v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) trs_v4u trs_g_v4u;

template <bool NT = false>
__device__ __forceinline__ uint4 ld16(const uint8_t *p) {
  const trs_g_v4u *g = (const trs_g_v4u *)(p);
  trs_v4u v;
  if (NT)
    v = __builtin_nontemporal_load(g);
  else
    v = *g;
  return make_uint4(v.x, v.y, v.z, v.w);
}

template <bool NT = false>
__device__ __forceinline__ void st16(uint8_t *p, uint4 x) {
  trs_g_v4u *g = (trs_g_v4u *)(p);
  const trs_v4u v = {x.x, x.y, x.z, x.w};
  if (NT)
    __builtin_nontemporal_store(v, g);
  else
    *g = v;
}

// 4-byte shard words through global pointers.
__device__ __forceinline__ uint32_t ld4(const uint8_t *p) {
  return *(const __attribute__((address_space(1))) uint32_t *)(p);
}

__device__ __forceinline__ void st4(uint8_t *p, uint32_t v) {
  *(__attribute__((address_space(1))) uint32_t *)(p) = v;
}

// One byte of a shard through a global pointer (tail and unaligned paths).
__device__ __forceinline__ uint8_t ld1(const uint8_t *p) {
  return *(const __attribute__((address_space(1))) uint8_t *)(p);
}

__device__ __forceinline__ void st1(uint8_t *p, uint8_t v) {
  *(__attribute__((address_space(1))) uint8_t *)(p) = v;
}

// A read-only word at a wave-uniform address, through the scalar data cache
// (s_load): the constant address space tells the compiler that nothing writes
// it during the kernel.  For coefficient tables and program words only.
__device__ __forceinline__ uint32_t ld_const(const uint32_t *p) {
  return *(const __attribute__((address_space(4))) uint32_t *)(p);
}

// A read-only struct (a work-list job) at a wave-uniform address; its fields
// load through the scalar data cache.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T &ld_const_ref(const T *p) {
  return *(const __attribute__((address_space(4))) T *)(p);
}

// A copy of a read-only struct at a wave-uniform address, loaded word by word
// through the scalar data cache (T: a multiple of 4 bytes, 4-byte aligned).
template <class T>
__device__ __forceinline__ T ld_const_val(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized struct");
  T v;
  uint32_t *dst = reinterpret_cast<uint32_t *>(&v);
  const __attribute__((address_space(4))) uint32_t *src =
      (const __attribute__((address_space(4))) uint32_t *)(p);
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) dst[i] = src[i];
  return v;
}

// An entry of a read-only pointer table at a wave-uniform address (s_load).
template <class T>
__device__ __forceinline__ T *ld_const_ptr(T *const *p) {
  return reinterpret_cast<T *>(*(const __attribute__((address_space(4))) uint64_t *)(p));
}

}  // namespace trs
