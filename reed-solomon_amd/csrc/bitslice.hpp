// Byte <-> bit-plane conversion for 32-byte lane chunks (device code).
//
// A lane's 32 bytes sit in 8 words w[0..7] (byte 4i+j in byte j of w[i]).
// Three rounds of masked bit swaps between word pairs (distance 1, 2, 4 with
// shifts 1, 2, 4) leave w[b] holding bit b of all 32 bytes: after the first
// round w[even] holds the even bits of its pair and w[odd] the odd bits, and
// so on.  Each swap is its own inverse, so running the rounds in reverse
// order turns planes back into bytes.  The byte order inside a plane is some
// fixed permutation of the 32 bytes; every GF(2^8) operation on planes is
// position-wise, so that permutation is invisible after the inverse.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "uniform.hpp"

namespace trs {

__device__ __forceinline__ void swap_bits(uint32_t &a, uint32_t &b, uint32_t mask, int n) {
  const uint32_t t = ((a >> n) ^ b) & mask;
  b ^= t;
  a ^= t << n;
}

__device__ __forceinline__ void bytes_to_planes(uint32_t (&w)[8]) {
  swap_bits(w[0], w[1], 0x55555555u, 1);
  swap_bits(w[2], w[3], 0x55555555u, 1);
  swap_bits(w[4], w[5], 0x55555555u, 1);
  swap_bits(w[6], w[7], 0x55555555u, 1);
  swap_bits(w[0], w[2], 0x33333333u, 2);
  swap_bits(w[1], w[3], 0x33333333u, 2);
  swap_bits(w[4], w[6], 0x33333333u, 2);
  swap_bits(w[5], w[7], 0x33333333u, 2);
  swap_bits(w[0], w[4], 0x0f0f0f0fu, 4);
  swap_bits(w[1], w[5], 0x0f0f0f0fu, 4);
  swap_bits(w[2], w[6], 0x0f0f0f0fu, 4);
  swap_bits(w[3], w[7], 0x0f0f0f0fu, 4);
}

__device__ __forceinline__ void planes_to_bytes(uint32_t (&w)[8]) {
  swap_bits(w[0], w[4], 0x0f0f0f0fu, 4);
  swap_bits(w[1], w[5], 0x0f0f0f0fu, 4);
  swap_bits(w[2], w[6], 0x0f0f0f0fu, 4);
  swap_bits(w[3], w[7], 0x0f0f0f0fu, 4);
  swap_bits(w[0], w[2], 0x33333333u, 2);
  swap_bits(w[1], w[3], 0x33333333u, 2);
  swap_bits(w[4], w[6], 0x33333333u, 2);
  swap_bits(w[5], w[7], 0x33333333u, 2);
  swap_bits(w[0], w[1], 0x55555555u, 1);
  swap_bits(w[2], w[3], 0x55555555u, 1);
  swap_bits(w[4], w[5], 0x55555555u, 1);
  swap_bits(w[6], w[7], 0x55555555u, 1);
}

// 32 bytes at p into planes; `full` lanes use two 16-byte loads (non-temporal
// when NT), others load the `rem` valid bytes and zero the rest.
template <bool NT = false>
__device__ __forceinline__ void load_planes(const uint8_t *p, bool full, uint64_t rem,
                                            uint32_t (&w)[8]) {
  if (full) {
    const uint4 a = ld16<NT>(p);
    const uint4 b = ld16<NT>(p + 16);
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = 0;
#pragma unroll
    for (int b = 0; b < 32; ++b)
      if (uint64_t(b) < rem) w[b >> 2] |= uint32_t(p[b]) << (8 * (b & 3));
  }
  bytes_to_planes(w);
}

template <bool NT = false>
__device__ __forceinline__ void store_planes(uint8_t *p, bool full, uint64_t rem,
                                             const uint32_t (&planes)[8]) {
  uint32_t w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = planes[i];
  planes_to_bytes(w);
  if (full) {
    st16<NT>(p, make_uint4(w[0], w[1], w[2], w[3]));
    st16<NT>(p + 16, make_uint4(w[4], w[5], w[6], w[7]));
  } else {
#pragma unroll
    for (int b = 0; b < 32; ++b)
      if (uint64_t(b) < rem) p[b] = uint8_t(w[b >> 2] >> (8 * (b & 3)));
  }
}

}  // namespace trs
