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

// Exchanges the bits of b under mask with the bits of a under ~mask, shifted
// by n (~mask == mask << n for the three masks used here).  SEL: two shifts
// and two selects (S0 & S2) | (S1 & ~S2) instead of two shifts and three
// XOR / AND operations; one more live register during the swap.
template <bool SEL = false>
__device__ __forceinline__ void swap_bits(uint32_t &a, uint32_t &b, uint32_t mask, int n) {
  if constexpr (SEL) {
    const uint32_t na = __builtin_amdgcn_bitop3_b32(a, b << n, mask, 0xe4);
    b = __builtin_amdgcn_bitop3_b32(a >> n, b, mask, 0xe4);
    a = na;
  } else {
    const uint32_t t = ((a >> n) ^ b) & mask;
    b ^= t;
    a ^= t << n;
  }
}

template <bool SEL = false>
__device__ __forceinline__ void bytes_to_planes(uint32_t (&w)[8]) {
  swap_bits<SEL>(w[0], w[1], 0x55555555u, 1);
  swap_bits<SEL>(w[2], w[3], 0x55555555u, 1);
  swap_bits<SEL>(w[4], w[5], 0x55555555u, 1);
  swap_bits<SEL>(w[6], w[7], 0x55555555u, 1);
  swap_bits<SEL>(w[0], w[2], 0x33333333u, 2);
  swap_bits<SEL>(w[1], w[3], 0x33333333u, 2);
  swap_bits<SEL>(w[4], w[6], 0x33333333u, 2);
  swap_bits<SEL>(w[5], w[7], 0x33333333u, 2);
  swap_bits<SEL>(w[0], w[4], 0x0f0f0f0fu, 4);
  swap_bits<SEL>(w[1], w[5], 0x0f0f0f0fu, 4);
  swap_bits<SEL>(w[2], w[6], 0x0f0f0f0fu, 4);
  swap_bits<SEL>(w[3], w[7], 0x0f0f0f0fu, 4);
}

template <bool SEL = false>
__device__ __forceinline__ void planes_to_bytes(uint32_t (&w)[8]) {
  swap_bits<SEL>(w[0], w[4], 0x0f0f0f0fu, 4);
  swap_bits<SEL>(w[1], w[5], 0x0f0f0f0fu, 4);
  swap_bits<SEL>(w[2], w[6], 0x0f0f0f0fu, 4);
  swap_bits<SEL>(w[3], w[7], 0x0f0f0f0fu, 4);
  swap_bits<SEL>(w[0], w[2], 0x33333333u, 2);
  swap_bits<SEL>(w[1], w[3], 0x33333333u, 2);
  swap_bits<SEL>(w[4], w[6], 0x33333333u, 2);
  swap_bits<SEL>(w[5], w[7], 0x33333333u, 2);
  swap_bits<SEL>(w[0], w[1], 0x55555555u, 1);
  swap_bits<SEL>(w[2], w[3], 0x55555555u, 1);
  swap_bits<SEL>(w[4], w[5], 0x55555555u, 1);
  swap_bits<SEL>(w[6], w[7], 0x55555555u, 1);
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
      if (uint64_t(b) < rem) w[b >> 2] |= uint32_t(ld1(p + b)) << (8 * (b & 3));
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
      if (uint64_t(b) < rem) st1(p + b, uint8_t(w[b >> 2] >> (8 * (b & 3))));
  }
}

// Split lane layout of an 8 KiB tile (256 lanes x 32 bytes): lane l of wave w
// owns the two 16-byte pieces at w*2048 + l*16 and w*2048 + 1024 + l*16, so
// each 16-byte load or store instruction of a wave covers 1 KiB of contiguous
// bytes (eight full 128-byte lines) instead of half of sixteen lines.  Which
// 32 bytes a lane holds is invisible to position-wise plane arithmetic as
// long as loads and stores use the same layout.
constexpr uint64_t kSplitGap = 1024;

__device__ __forceinline__ uint64_t split_lane_offset(uint32_t tid) {
  return uint64_t(tid >> 6) * 2048 + uint64_t(tid & 63) * 16;
}

// How a tile's lanes reach their bytes; the same for every lane of a tile,
// so a kernel branches on it once and each path is straight-line code whose
// loads can all be in flight together (a per-load branch would put a wait
// after every load, and lanes of one wave taking different paths would run
// the program twice).
enum SplitMode : int {
  kSplitBytes = 0,   // unaligned shard: byte accesses
  kSplitRagged = 1,  // aligned, length not a multiple of 16: 16-byte accesses
                     // except the piece holding the end
  kSplitMasked = 2,  // aligned, length a multiple of 16, tile crosses the end:
                     // a lane's first piece is in bounds whenever the lane is
                     // active; its second is loaded from a clamped address and
                     // zeroed, or not stored, when past the end
  kSplitFull = 3     // aligned, whole 8 KiB tile in bounds
};

__device__ __forceinline__ int split_mode(bool vec_ok, uint64_t tile_off, uint64_t len) {
  if (!vec_ok) return kSplitBytes;
  if (tile_off + 8192 <= len) return kSplitFull;
  return (len & 15) == 0 ? kSplitMasked : kSplitRagged;
}

// One 16-byte piece of a lane: one 16-byte access when the piece is in
// bounds and aligned, bytes otherwise; `rem` = shard bytes from p to the end
// (0 = piece past the end: zeros, nothing stored).
template <bool NT = false>
__device__ __forceinline__ void load_piece(const uint8_t *p, bool vec_ok, uint64_t rem,
                                           uint32_t *w) {
  if (vec_ok && rem >= 16) {
    const uint4 a = ld16<NT>(p);
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b)
      if (uint64_t(b) < rem) w[b >> 2] |= uint32_t(ld1(p + b)) << (8 * (b & 3));
  }
}

template <bool NT = false>
__device__ __forceinline__ void store_piece(uint8_t *p, bool vec_ok, uint64_t rem,
                                            const uint32_t *w) {
  if (vec_ok && rem >= 16) {
    st16<NT>(p, make_uint4(w[0], w[1], w[2], w[3]));
  } else {
#pragma unroll
    for (int b = 0; b < 16; ++b)
      if (uint64_t(b) < rem) st1(p + b, uint8_t(w[b >> 2] >> (8 * (b & 3))));
  }
}

// The lane's 32 bytes of shard `base` (len bytes) at off (first piece) and
// off + kSplitGap as raw words, bytes past the end reading as zero.  The lane
// is active: off < len.
template <bool NT = false>
__device__ __forceinline__ void load_split_raw(const uint8_t *base, uint64_t off, uint64_t len,
                                               int mode, uint32_t (&w)[8]) {
  const uint8_t *p = base + off;
  if (mode == kSplitFull || mode == kSplitMasked) {
    const bool in1 = mode == kSplitFull || off + kSplitGap < len;
    const uint4 a = ld16<NT>(p);
    const uint4 b = ld16<NT>(in1 ? p + kSplitGap : p);
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    w[4] = in1 ? b.x : 0u;
    w[5] = in1 ? b.y : 0u;
    w[6] = in1 ? b.z : 0u;
    w[7] = in1 ? b.w : 0u;
  } else {
    const uint64_t rem = len - off;
    load_piece<NT>(p, mode == kSplitRagged, rem, w);
    load_piece<NT>(p + kSplitGap, mode == kSplitRagged, rem > kSplitGap ? rem - kSplitGap : 0,
                   w + 4);
  }
}

template <bool NT = false, bool SEL = false>
__device__ __forceinline__ void load_split(const uint8_t *base, uint64_t off, uint64_t len,
                                           int mode, uint32_t (&w)[8]) {
  load_split_raw<NT>(base, off, len, mode, w);
  bytes_to_planes<SEL>(w);
}

// Verify epilogue of a program: instead of storing `planes`, compare them
// with the lane's 32 bytes of shard `base` and OR the difference into `bad`.
// Bytes past the end read as zero, and a linear program's output there is
// zero too (its inputs read as zero), so they never differ.
template <bool NT = false>
__device__ __forceinline__ void verify_split(const uint8_t *base, uint64_t off, uint64_t len,
                                             int mode, const uint32_t (&planes)[8],
                                             uint32_t &bad) {
  uint32_t w[8], x[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = planes[i];
  planes_to_bytes(w);
  load_split_raw<NT>(base, off, len, mode, x);
#pragma unroll
  for (int i = 0; i < 8; ++i) bad |= w[i] ^ x[i];
}

template <bool NT = false, bool SEL = false>
__device__ __forceinline__ void store_split(uint8_t *base, uint64_t off, uint64_t len, int mode,
                                            const uint32_t (&planes)[8]) {
  uint32_t w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = planes[i];
  planes_to_bytes<SEL>(w);
  uint8_t *p = base + off;
  if (mode == kSplitFull || mode == kSplitMasked) {
    st16<NT>(p, make_uint4(w[0], w[1], w[2], w[3]));
    if (mode == kSplitFull || off + kSplitGap < len)
      st16<NT>(p + kSplitGap, make_uint4(w[4], w[5], w[6], w[7]));
  } else {
    const uint64_t rem = len - off;
    store_piece<NT>(p, mode == kSplitRagged, rem, w);
    store_piece<NT>(p + kSplitGap, mode == kSplitRagged, rem > kSplitGap ? rem - kSplitGap : 0,
                    w + 4);
  }
}

}  // namespace trs
