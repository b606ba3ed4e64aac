// HIP kernels for gfx950 (MI355X): fused GF(2^8) matrix apply, tower
// (GF((2^8)^2)) Karatsuba apply over split planes, the batched work-list
// kernel, and the seeded fill used by benches and tests.
//
// Arithmetic.  Multiplying a packed word of 4 bytes by 2 in GF(2^8) (poly
// 0x11d) is xtime4: shift every byte left, then fold 0x1d into the bytes whose
// top bit fell out.  A coefficient c is applied as the XOR of the powers
// x*2^b for the set bits b of c; the coefficient is uniform across the
// workgroup, so its bits become scalar masks and each term is one
// `acc ^= p & mask` (a single v_bitop3 on gfx950).  The powers of an input
// word are computed once and reused by every output row.  This is exact
// GF(2^8) arithmetic: it matches galois::mul (src/galois.rs:74) bit for bit.
//
// Layout.  Every lane owns 16 consecutive bytes of every shard (one 16-byte
// global load per input), so a 256-lane workgroup streams a 4 KiB column
// chunk of all inputs once, coalesced, and keeps up to 16 output rows in
// VGPRs.  blockIdx.y tiles larger output counts.
//
// The batched work list runs a bit-sliced variant of the same apply
// (k_batch8_planes, below): 32 bytes per lane as 8 bit planes, where the
// powers x*2^b cost three XORs each instead of a packed xtime4.
#include <hip/hip_runtime.h>

#include "bitslice.hpp"
#include "kernels.hpp"
#include "uniform.hpp"

// 1: apply each coefficient bit as a full-word mask (branch-free);
// 0: skip zero bits with scalar branches.
#ifndef TRS_MASKED_MAC
#define TRS_MASKED_MAC 0
#endif

namespace trs {

namespace {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// The reduction bytes (0 or 0x1d per byte) come from one packed 16-bit
// multiply (full rate), not a 32-bit v_mul_lo (quarter rate).
__device__ __forceinline__ uint32_t xtime4(uint32_t x) {
  const uint32_t hi = (x >> 7) & 0x01010101u;
  const u16x2 red = __builtin_bit_cast(u16x2, hi) * (u16x2){0x1d, 0x1d};
  return ((x << 1) & 0xfefefefeu) ^ __builtin_bit_cast(uint32_t, red);
}

// Loads 16 bytes at p; `full` lanes use one 16-byte load, the others (tail
// or unaligned shards) assemble the `rem` valid bytes and zero the rest.
__device__ __forceinline__ void load_lane(const uint8_t *p, bool full, uint64_t rem,
                                          uint32_t x[4]) {
  if (full) {
    const uint4 v = ld16(p);
    x[0] = v.x;
    x[1] = v.y;
    x[2] = v.z;
    x[3] = v.w;
  } else {
    x[0] = x[1] = x[2] = x[3] = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b)
      if (uint64_t(b) < rem) x[b >> 2] |= uint32_t(ld1(p + b)) << (8 * (b & 3));
  }
}

__device__ __forceinline__ void store_lane(uint8_t *p, bool full, uint64_t rem,
                                           const uint32_t x[4]) {
  if (full) {
    st16(p, make_uint4(x[0], x[1], x[2], x[3]));
  } else {
#pragma unroll
    for (int b = 0; b < 16; ++b)
      if (uint64_t(b) < rem) st1(p + b, uint8_t(x[b >> 2] >> (8 * (b & 3))));
  }
}

#if TRS_MASKED_MAC
// acc ^ (p & m) in one v_bitop3 (truth table over (S0, S1, S2) = (acc, p, m)
// indexed S0 S1 S2 from the high bit: 0x78); the compiler does not fuse it
// when m is a scalar register.
__device__ __forceinline__ uint32_t xor_and(uint32_t acc, uint32_t p, uint32_t m) {
  return __builtin_amdgcn_bitop3_b32(acc, p, m, 0x78);
}
#endif

// acc[j] ^= c_j * x for the NT coefficients at cw (transposed, padded rows:
// byte j of the NT/4 words at cw is output j's coefficient).
template <int NT>
__device__ __forceinline__ void mac_rows(uint32_t (&acc)[NT][4], const uint32_t x[4],
                                         const uint32_t *cw) {
  uint32_t p[8][4];
#pragma unroll
  for (int w = 0; w < 4; ++w) p[0][w] = x[w];
#pragma unroll
  for (int b = 1; b < 8; ++b)
#pragma unroll
    for (int w = 0; w < 4; ++w) p[b][w] = xtime4(p[b - 1][w]);
#pragma unroll
  for (int q = 0; q < NT / 4; ++q) {
    const uint32_t word = ld_const(cw + q);  // same address in every lane: s_load
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const uint32_t c = (word >> (8 * s)) & 0xffu;  // scalar
#if TRS_MASKED_MAC
      if (c) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const uint32_t msk = 0u - ((c >> b) & 1u);
#pragma unroll
          for (int w = 0; w < 4; ++w) acc[q * 4 + s][w] = xor_and(acc[q * 4 + s][w], p[b][w], msk);
        }
      }
#else
      // zero bits cost a scalar branch, not VALU work
#pragma unroll
      for (int b = 0; b < 8; ++b)
        if ((c >> b) & 1u) {
#pragma unroll
          for (int w = 0; w < 4; ++w) acc[q * 4 + s][w] ^= p[b][w];
        }
#endif
    }
  }
}

// One lane's share of a GF(2^8) apply tile: outputs jbase..jbase+nt over the
// 16 bytes at `off`.  ptr_at(i) returns shard pointer i (inputs first).
template <int NT, uint32_t MODE, typename PtrAt>
__device__ __forceinline__ void gf8_tile_body(PtrAt ptr_at, const uint32_t *coef, uint64_t rem,
                                              bool full, uint32_t k, uint32_t nout,
                                              uint32_t ostride, uint32_t jbase, uint64_t off,
                                              uint32_t *flag) {
  const uint32_t nt = min(uint32_t(NT), nout - jbase);
  uint32_t acc[NT][4];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int w = 0; w < 4; ++w) acc[j][w] = 0;

  for (uint32_t i = 0; i < k; ++i) {
    uint32_t x[4];
    load_lane(ptr_at(i) + off, full, rem, x);
    mac_rows<NT>(acc, x, coef + ((size_t)i * ostride + jbase) / 4);
  }

  uint32_t bad = 0;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    if (uint32_t(j) < nt) {
      uint8_t *o = const_cast<uint8_t *>(ptr_at(k + jbase + j)) + off;
      if (MODE == kWrite) {
        store_lane(o, full, rem, acc[j]);
      } else {
        uint32_t y[4];
        load_lane(o, full, rem, y);
        if (MODE == kVerify) {
#pragma unroll
          for (int w = 0; w < 4; ++w) bad |= y[w] ^ acc[j][w];
        } else {
#pragma unroll
          for (int w = 0; w < 4; ++w) y[w] ^= acc[j][w];
          store_lane(o, full, rem, y);
        }
      }
    }
  }
  if (MODE == kVerify && bad) *flag = 1u;
}

// The full / tail choice is made once per tile: the full path is then
// straight-line code whose loads can all be in flight together, where a
// per-load branch would put a wait after every load.
template <int NT, uint32_t MODE, typename PtrAt>
__device__ __forceinline__ void gf8_tile(PtrAt ptr_at, const uint32_t *coef, uint64_t len,
                                         uint32_t k, uint32_t nout, uint32_t ostride,
                                         bool vec_ok, uint32_t jbase, uint64_t tile,
                                         uint64_t off, uint32_t *flag) {
  // Every active lane has 16 bytes in bounds when the shard is aligned and its
  // length a multiple of 16 or the whole tile is in bounds; the choice is the
  // same for all lanes of the tile.
  const uint64_t rem = len - off;
  if (vec_ok && ((len & 15) == 0 || tile + kChunkBytes <= len))
    gf8_tile_body<NT, MODE>(ptr_at, coef, rem, true, k, nout, ostride, jbase, off, flag);
  else
    gf8_tile_body<NT, MODE>(ptr_at, coef, rem, false, k, nout, ostride, jbase, off, flag);
}

// ---- Bit-sliced apply with runtime coefficients ----
//
// A lane holds 32 bytes of a shard as 8 bit planes (bitslice.hpp, split
// layout).  In plane form multiplying by 2 (poly 0x11d: 2*v = v<<1 ^ (v7 ?
// 0x1d : 0)) moves plane b to b+1 and XORs the old top plane into planes 0
// (as a move), 2, 3 and 4: three XORs per 32 bytes, against ~40 operations
// for the packed-byte xtime4.  c*x is the XOR of x*2^b over the set bits b of
// c, so each input walks y = x, 2x, 4x, ... and XORs y into every output
// whose coefficient has bit b set; the bit tests are scalar branches on the
// wave-uniform coefficient words.

__device__ __forceinline__ void planes_times2(uint32_t (&y)[8]) {
  const uint32_t t = y[7];
  y[7] = y[6];
  y[6] = y[5];
  y[5] = y[4];
  y[4] = y[3] ^ t;
  y[3] = y[2] ^ t;
  y[2] = y[1] ^ t;
  y[1] = y[0];
  y[0] = t;
}

// y / 2: the inverse of planes_times2.
__device__ __forceinline__ void planes_half(uint32_t (&y)[8]) {
  const uint32_t t = y[0];
  y[0] = y[1];
  y[1] = y[2] ^ t;
  y[2] = y[3] ^ t;
  y[3] = y[4] ^ t;
  y[4] = y[5];
  y[5] = y[6];
  y[6] = y[7];
  y[7] = t;
}

template <int S>
__device__ __forceinline__ void planes_steps(uint32_t (&y)[8], bool half) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (half) planes_half(y);
    else planes_times2(y);
  }
}

// y *= 2^-t for a wave-uniform t in [-7, 7]: one straight-line sequence of
// |t| steps per value, so the plane renaming costs no moves inside a case.
__device__ __forceinline__ void planes_unscale(uint32_t (&y)[8], int t) {
  switch (t) {
    case 1: planes_steps<1>(y, true); break;
    case 2: planes_steps<2>(y, true); break;
    case 3: planes_steps<3>(y, true); break;
    case 4: planes_steps<4>(y, true); break;
    case 5: planes_steps<5>(y, true); break;
    case 6: planes_steps<6>(y, true); break;
    case 7: planes_steps<7>(y, true); break;
    case -1: planes_steps<1>(y, false); break;
    case -2: planes_steps<2>(y, false); break;
    case -3: planes_steps<3>(y, false); break;
    case -4: planes_steps<4>(y, false); break;
    case -5: planes_steps<5>(y, false); break;
    case -6: planes_steps<6>(y, false); break;
    case -7: planes_steps<7>(y, false); break;
    default: break;
  }
}

// acc[j] ^= c_j * y for the NT coefficients at cw (same layout as mac_rows);
// y is consumed (left at y * 2^7).
template <int NT>
__device__ __forceinline__ void mac_planes(uint32_t (&acc)[NT][8], uint32_t (&y)[8],
                                           const uint32_t *cw) {
  uint32_t word[NT / 4];
#pragma unroll
  for (int q = 0; q < NT / 4; ++q) word[q] = ld_const(cw + q);  // wave-uniform: s_load
#pragma unroll
  for (int b = 0; b < 8; ++b) {
#pragma unroll
    for (int j = 0; j < NT; ++j)
      if ((word[j >> 2] >> (8 * (j & 3) + b)) & 1u) {
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[j][r] ^= y[r];
      }
    if (b < 7) planes_times2(y);
  }
}

// acc ^= the XOR of the powers p[b] over the set bits b of the 4-bit N, two
// terms per three-input XOR: ceil(popcount(N) / 2) operations per plane.
template <int N>
__device__ __forceinline__ void nib_xor(uint32_t (&acc)[8], const uint32_t (&p0)[8],
                                        const uint32_t (&p1)[8], const uint32_t (&p2)[8],
                                        const uint32_t (&p3)[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    uint32_t t[4];
    int n = 0;
    if (N & 1) t[n++] = p0[r];
    if (N & 2) t[n++] = p1[r];
    if (N & 4) t[n++] = p2[r];
    if (N & 8) t[n++] = p3[r];
    uint32_t v = acc[r];
    if (n >= 2) v = __builtin_amdgcn_bitop3_b32(v, t[0], t[1], 0x96);
    else if (n == 1) v ^= t[0];
    if (n == 4) v = __builtin_amdgcn_bitop3_b32(v, t[2], t[3], 0x96);
    else if (n == 3) v ^= t[2];
    acc[r] = v;
  }
}

// The 16-way choice on the wave-uniform nibble at bit s of w as a tree of
// if / else on its bits (B = 3 .. 0), so each leaf updates acc in place.
// Each test reads w through readfirstlane, which the compiler does not move
// across the branches: without it, the tests are hoisted and merged into
// lane-mask conditions or a switch whose joins copy the accumulators.
template <int B, int N>
__device__ __forceinline__ void nib_tree(uint32_t (&acc)[8], uint32_t w, uint32_t s,
                                         const uint32_t (&p0)[8], const uint32_t (&p1)[8],
                                         const uint32_t (&p2)[8], const uint32_t (&p3)[8]) {
  if constexpr (B < 0) {
    if constexpr (N != 0) nib_xor<N>(acc, p0, p1, p2, p3);
  } else {
    if ((__builtin_amdgcn_readfirstlane(w) >> (s + B)) & 1u)
      nib_tree<B - 1, N | (1 << B)>(acc, w, s, p0, p1, p2, p3);
    else
      nib_tree<B - 1, N>(acc, w, s, p0, p1, p2, p3);
  }
}

// acc[j] ^= c_j * y as mac_planes, with each coefficient taken a nibble at a
// time: the four powers of a nibble stay live (they share five of every
// eight planes with their neighbours, so 17 registers hold all four), one
// wave-uniform 16-way branch per (output, nibble) selects the straight-line
// XOR of its set powers, and pairs of terms fold into one three-input XOR.
// y is consumed.
template <int NT>
__device__ __forceinline__ void mac_planes_nib(uint32_t (&acc)[NT][8], uint32_t (&y)[8],
                                               const uint32_t *cw) {
  uint32_t word[NT / 4];
#pragma unroll
  for (int q = 0; q < NT / 4; ++q) word[q] = ld_const(cw + q);  // wave-uniform: s_load
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint32_t p1[8], p2[8], p3[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) p1[r] = y[r];
    planes_times2(p1);
#pragma unroll
    for (int r = 0; r < 8; ++r) p2[r] = p1[r];
    planes_times2(p2);
#pragma unroll
    for (int r = 0; r < 8; ++r) p3[r] = p2[r];
    planes_times2(p3);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      nib_tree<3, 0>(acc[j], word[j >> 2], 8 * (j & 3) + 4 * h, y, p1, p2, p3);
    }
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) y[r] = p3[r];
      planes_times2(y);
    }
  }
}

// One lane's share of a bit-sliced apply tile: outputs jbase..jbase+nt over
// the lane's split 32 bytes whose first piece is at `off` (an active lane:
// off < len).  The next input's bytes are loaded before the current input's
// multiply-accumulate.  coef holds the rows scaled by 2^t_j and, after the
// k * ostride coefficient bytes, the ostride exponents t_j (int8): each
// output is multiplied by 2^-t_j before it is stored.
// FAST (the 16-byte access modes): nibble-wise multiply-accumulate and the
// select form of the transposes.  LNT: the 16-byte input loads are
// non-temporal (inputs that no other output tile reads again).  The byte-access modes keep the bitwise
// multiply-accumulate, whose smaller code and register set leave the
// kernel's register count to the fast paths.
template <int NT, uint32_t MODE, bool FAST = false, bool LNT = false, typename PtrAt>
__device__ __forceinline__ void gf8_planes_body(PtrAt ptr_at, const uint32_t *coef, uint64_t len,
                                                int mode, uint32_t k, uint32_t nout,
                                                uint32_t ostride, uint32_t jbase, uint64_t off,
                                                uint32_t *flag) {
  const uint32_t nt = min(uint32_t(NT), nout - jbase);
  uint32_t acc[NT][8];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int r = 0; r < 8; ++r) acc[j][r] = 0;

  uint32_t nxt[8];
  load_split_raw<LNT>(ptr_at(0), off, len, mode, nxt);
  for (uint32_t i = 0; i < k; ++i) {
    uint32_t y[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) y[r] = nxt[r];
    if (i + 1 < k) load_split_raw<LNT>(ptr_at(i + 1), off, len, mode, nxt);
    bytes_to_planes<FAST>(y);
    if constexpr (FAST) mac_planes_nib<NT>(acc, y, coef + ((size_t)i * ostride + jbase) / 4);
    else mac_planes<NT>(acc, y, coef + ((size_t)i * ostride + jbase) / 4);
  }
  {
    uint32_t tw[NT / 4];
#pragma unroll
    for (int q = 0; q < NT / 4; ++q) tw[q] = ld_const(coef + ((size_t)k * ostride + jbase) / 4 + q);
#pragma unroll
    for (int j = 0; j < NT; ++j)
      if (uint32_t(j) < nt) planes_unscale(acc[j], int(int8_t(tw[j >> 2] >> (8 * (j & 3)))));
  }

  uint32_t bad = 0;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    if (uint32_t(j) < nt) {
      uint8_t *o = const_cast<uint8_t *>(ptr_at(k + jbase + j));
      if (MODE == kWrite) {
        store_split<true, FAST>(o, off, len, mode, acc[j]);
      } else {
        uint32_t y[8];
        load_split<false, FAST>(o, off, len, mode, y);
        if (MODE == kVerify) {
#pragma unroll
          for (int r = 0; r < 8; ++r) bad |= y[r] ^ acc[j][r];
        } else {
#pragma unroll
          for (int r = 0; r < 8; ++r) y[r] ^= acc[j][r];
          store_split<false, FAST>(o, off, len, mode, y);
        }
      }
    }
  }
  if (MODE == kVerify && bad) *flag = 1u;
}

// One branch per tile on its access mode (split_mode), so the common paths
// are straight-line copies of the body.
template <int NT, uint32_t MODE, bool LNT = false, typename PtrAt>
__device__ __forceinline__ void gf8_tile_planes(PtrAt ptr_at, const uint32_t *coef, uint64_t len,
                                                uint32_t k, uint32_t nout, uint32_t ostride,
                                                bool vec_ok, uint32_t jbase, uint64_t tile,
                                                uint64_t off, uint32_t *flag) {
  const int mode = split_mode(vec_ok, tile, len);
  if (mode == kSplitFull)
    gf8_planes_body<NT, MODE, true, LNT>(ptr_at, coef, len, kSplitFull, k, nout, ostride, jbase, off,
                                    flag);
  else if (mode == kSplitMasked)
    gf8_planes_body<NT, MODE, true, LNT>(ptr_at, coef, len, kSplitMasked, k, nout, ostride, jbase, off,
                                    flag);
  else
    gf8_planes_body<NT, MODE>(ptr_at, coef, len, mode, k, nout, ostride, jbase, off, flag);
}

template <int NT, uint32_t MODE>
__global__ __launch_bounds__(kBlockThreads) void k_gf8_apply(const Gf8Call a) {
  const uint64_t tile = uint64_t(blockIdx.x) * kChunkBytes;
  const uint64_t off = tile + uint64_t(threadIdx.x) * kLaneBytes;
  if (off >= a.len) return;
  gf8_tile<NT, MODE>([&](uint32_t idx) { return uni(a.ptr[idx]); }, a.coef, a.len, a.k, a.nout,
                     a.ostride, a.vec_ok != 0, blockIdx.y * NT, tile, off, a.flag);
}

// The single-call apply on the bit-sliced tile: 8 KiB column chunks in x,
// NT-row output tiles in y.
template <int NT, uint32_t MODE>
__global__ __launch_bounds__(kBlockThreads) void k_gf8_planes(const Gf8Call a) {
  const uint64_t tile = uint64_t(blockIdx.x) * kPlanesChunkBytes;
  const uint64_t off = tile + split_lane_offset(threadIdx.x);
  if (off >= a.len) return;
  gf8_tile_planes<NT, MODE>([&](uint32_t idx) { return uni(a.ptr[idx]); }, a.coef, a.len, a.k,
                            a.nout, a.ostride, a.vec_ok != 0, blockIdx.y * NT, tile, off,
                            a.flag);
}

// Batched work list: tiles are (job, 4 KiB column chunk, 16-row output tile),
// numbered job by job; tile_job[t] names the job of tile t.  A fixed grid
// strides over the tiles, which interleaves heavy and light jobs across
// workgroups.
__global__ __launch_bounds__(kBlockThreads) void k_batch8(const BatchJob *jobs,
                                                         const uint32_t *tile_job,
                                                         uint64_t total) {
  for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
    const auto &J = ld_const_ref(jobs + ld_const(tile_job + t));
    const uint64_t len = J.len;
    const uint32_t out_tiles = J.out_tiles;
    const uint64_t local = t - J.tile_begin;
    const uint64_t chunk = local / out_tiles;
    const uint32_t ot = uint32_t(local - chunk * out_tiles);
    const uint64_t tile = chunk * kChunkBytes;
    const uint64_t off = tile + uint64_t(threadIdx.x) * kLaneBytes;
    if (off < len) {
      const uint8_t *const *ptrs = J.ptrs;
      gf8_tile<kOutStride, kWrite>([&](uint32_t idx) { return ld_const_ptr(ptrs + idx); },
                                   J.coef, len, J.k, J.nout, J.ostride, J.vec_ok != 0,
                                   ot * kOutStride, tile, off, nullptr);
    }
  }
}

// The same work list over the bit-sliced tile: tiles are (job, 8 KiB column
// chunk, kPlanesOut-row output tile).  Eight rows (64 accumulator VGPRs, five
// waves per SIMD) measured faster than sixteen (three waves) although jobs
// with more than eight outputs then read their inputs once per tile.  The
// nibble-wise multiply-accumulate needs 96 VGPRs for five waves; the
// attribute holds the compiler to that.
__global__ __launch_bounds__(kBlockThreads)
__attribute__((amdgpu_waves_per_eu(kBatchPlanesOut == 8 ? 5 : 3))) void
k_batch8_planes(const BatchJob *jobs, const uint32_t *tile_job, uint64_t total) {
  for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
    const auto &J = ld_const_ref(jobs + ld_const(tile_job + t));
    const uint64_t len = J.len;
    const uint32_t out_tiles = J.out_tiles;
    const uint64_t local = t - J.tile_begin;
    // Within a job, tiles run in groups of eight chunks, output tile major:
    // the output tiles of one chunk are 8 tile indices apart (fewer in a
    // job's last group), so they land on one XCD (workgroups go to the XCDs
    // round robin) at about the same time and their input re-reads meet in
    // its L2.  One-tile jobs keep chunk order.
    const uint64_t chunks = (len + kPlanesChunkBytes - 1) / kPlanesChunkBytes;
    const uint64_t q = local / (8u * out_tiles);
    const uint64_t r = local - q * 8u * out_tiles;
    const uint64_t g = min(uint64_t(8), chunks - q * 8u);
    const uint32_t ot = uint32_t(r / g);
    const uint64_t chunk = q * 8u + (r - uint64_t(ot) * g);
    const uint64_t tile = chunk * kPlanesChunkBytes;
    const uint64_t off = tile + split_lane_offset(threadIdx.x);
    if (off < len) {
      const uint8_t *const *ptrs = J.ptrs;
      auto ptr_at = [&](uint32_t idx) { return ld_const_ptr(ptrs + idx); };
      // A job with one output tile reads its inputs once: non-temporal
      // loads keep them from evicting the inputs that multi-tile jobs read
      // again per output tile (measured 4086-4133 -> 4398-4404 GB/s; making
      // the last tile of every job non-temporal instead was slower).
      if (out_tiles == 1)
        gf8_tile_planes<kBatchPlanesOut, kWrite, true>(ptr_at, J.coef, len, J.k, J.nout,
                                                       J.ostride, J.vec_ok != 0,
                                                       ot * kBatchPlanesOut, tile, off, nullptr);
      else
        gf8_tile_planes<kBatchPlanesOut, kWrite>(ptr_at, J.coef, len, J.k, J.nout, J.ostride,
                                                 J.vec_ok != 0, ot * kBatchPlanesOut, tile, off,
                                                 nullptr);
    }
  }
}

// Tower apply over split planes with the Karatsuba triple
// [c_lo, c_hi^c_lo, c_hi*L] (src/fft16.rs:176-180): per input symbol
// (a_hi, a_lo), t0 = c_lo*a_lo, t1 = (c_hi^c_lo)*(a_hi^a_lo), t2 = (c_hi*L)*a_hi,
// out_lo ^= t0^t2, out_hi ^= t0^t1 (src/gf/tower_x86.rs:80-93).  The three
// product sums are accumulated separately and combined at the end.
template <int NT, uint32_t MODE, typename PtrAt>
__device__ __forceinline__ void gf16_body(const Gf16Call &a, PtrAt ptr_at, uint64_t off,
                                          uint64_t rem, bool full) {
  const uint32_t jbase = blockIdx.y * NT;
  const uint32_t nt = min(uint32_t(NT), a.nout - jbase);
  const size_t csz = (size_t)a.k * a.ostride / 4;
  uint32_t A[NT][4], B[NT][4], C[NT][4];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int w = 0; w < 4; ++w) A[j][w] = B[j][w] = C[j][w] = 0;

  for (uint32_t i = 0; i < a.k; ++i) {
    const uint8_t *s = ptr_at(i);
    uint32_t lo[4], hi[4], sum[4];
    load_lane(s + off, full, rem, lo);
    load_lane(s + a.plane + off, full, rem, hi);
    const uint32_t *cw = a.coef + ((size_t)i * a.ostride + jbase) / 4;
    mac_rows<NT>(A, lo, cw);
    mac_rows<NT>(C, hi, cw + 2 * csz);
#pragma unroll
    for (int w = 0; w < 4; ++w) sum[w] = lo[w] ^ hi[w];
    mac_rows<NT>(B, sum, cw + csz);
  }

  uint32_t bad = 0;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    if (uint32_t(j) < nt) {
      uint8_t *o = const_cast<uint8_t *>(ptr_at(a.k + jbase + j));
      uint32_t olo[4], ohi[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        olo[w] = A[j][w] ^ C[j][w];
        ohi[w] = A[j][w] ^ B[j][w];
      }
      if (MODE == kWrite) {
        store_lane(o + off, full, rem, olo);
        store_lane(o + a.plane + off, full, rem, ohi);
      } else {
        uint32_t ylo[4], yhi[4];
        load_lane(o + off, full, rem, ylo);
        load_lane(o + a.plane + off, full, rem, yhi);
#pragma unroll
        for (int w = 0; w < 4; ++w) bad |= (ylo[w] ^ olo[w]) | (yhi[w] ^ ohi[w]);
      }
    }
  }
  if (MODE == kVerify && bad) *a.flag = 1u;
}

// The full / tail choice is the same for every lane of a workgroup (as in
// gf8_tile).
template <int NT, uint32_t MODE, typename PtrAt>
__device__ __forceinline__ void gf16_tile(const Gf16Call &a, PtrAt ptr_at) {
  const uint64_t tile = uint64_t(blockIdx.x) * kChunkBytes;
  const uint64_t rel = tile + uint64_t(threadIdx.x) * kLaneBytes;
  if (rel >= a.cols) return;
  if (a.vec_ok && ((a.cols & 15) == 0 || tile + kChunkBytes <= a.cols))
    gf16_body<NT, MODE>(a, ptr_at, a.col0 + rel, a.cols - rel, true);
  else
    gf16_body<NT, MODE>(a, ptr_at, a.col0 + rel, a.cols - rel, false);
}

template <int NT, uint32_t MODE>
__global__ __launch_bounds__(kBlockThreads) void k_gf16_apply(const Gf16Call a) {
  gf16_tile<NT, MODE>(a, [&](uint32_t i) { return ld_const_ptr(a.ptrs + i); });
}

template <int NT, uint32_t MODE>
__global__ __launch_bounds__(kBlockThreads) void k_gf16_apply_inline(const Gf16InlineCall a) {
  gf16_tile<NT, MODE>(a.call, [&](uint32_t i) { return uni(a.ptr[i]); });
}

// Wire-format conversion between interleaved u16 symbols (symbol j in bytes
// 2j (low) and 2j+1 (high), little-endian) and the split-plane shard layout
// [low N | high N] (src/reedsolomon16.rs:4-7).  A lane converts 8 symbols;
// v_perm_b32 picks the even or odd bytes of two words (selector bytes index
// the 8-byte concatenation, the second operand first).
__global__ __launch_bounds__(kBlockThreads) void k_u16_to_split(const uint8_t *in, uint8_t *lo,
                                                               uint8_t *hi, uint64_t n,
                                                               uint32_t vec_ok) {
  const uint64_t s0 = (uint64_t(blockIdx.x) * kBlockThreads + threadIdx.x) * 8;
  if (s0 >= n) return;
  if (vec_ok && s0 + 8 <= n) {
    const uint4 v = ld16(in + 2 * s0);
    const uint32_t l0 = __builtin_amdgcn_perm(v.y, v.x, 0x06040200u);
    const uint32_t l1 = __builtin_amdgcn_perm(v.w, v.z, 0x06040200u);
    const uint32_t h0 = __builtin_amdgcn_perm(v.y, v.x, 0x07050301u);
    const uint32_t h1 = __builtin_amdgcn_perm(v.w, v.z, 0x07050301u);
    st4(lo + s0, l0);
    st4(lo + s0 + 4, l1);
    st4(hi + s0, h0);
    st4(hi + s0 + 4, h1);
  } else {
    for (uint64_t s = s0; s < n && s < s0 + 8; ++s) {
      st1(lo + s, ld1(in + 2 * s));
      st1(hi + s, ld1(in + 2 * s + 1));
    }
  }
}

__global__ __launch_bounds__(kBlockThreads) void k_split_to_u16(const uint8_t *lo,
                                                               const uint8_t *hi, uint8_t *out,
                                                               uint64_t n, uint32_t vec_ok) {
  const uint64_t s0 = (uint64_t(blockIdx.x) * kBlockThreads + threadIdx.x) * 8;
  if (s0 >= n) return;
  if (vec_ok && s0 + 8 <= n) {
    const uint32_t l0 = ld4(lo + s0), l1 = ld4(lo + s0 + 4);
    const uint32_t h0 = ld4(hi + s0), h1 = ld4(hi + s0 + 4);
    st16(out + 2 * s0, make_uint4(__builtin_amdgcn_perm(h0, l0, 0x05010400u),
                                  __builtin_amdgcn_perm(h0, l0, 0x07030602u),
                                  __builtin_amdgcn_perm(h1, l1, 0x05010400u),
                                  __builtin_amdgcn_perm(h1, l1, 0x07030602u)));
  } else {
    for (uint64_t s = s0; s < n && s < s0 + 8; ++s) {
      st1(out + 2 * s, ld1(lo + s));
      st1(out + 2 * s + 1, ld1(hi + s));
    }
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlockThreads) void k_fill(uint8_t *dst, uint64_t len,
                                                       uint64_t seed) {
  const uint64_t words = (len + 7) / 8;
  const bool aligned = (reinterpret_cast<uintptr_t>(dst) & 7) == 0;
  for (uint64_t w = uint64_t(blockIdx.x) * kBlockThreads + threadIdx.x; w < words;
       w += uint64_t(gridDim.x) * kBlockThreads) {
    const uint64_t v = mix64(seed + (w + 1) * 0x9E3779B97F4A7C15ull);
    const uint64_t base = w * 8;
    if (aligned && base + 8 <= len) {
      *reinterpret_cast<uint64_t *>(dst + base) = v;
    } else {
      for (int b = 0; b < 8; ++b)
        if (base + b < len) dst[base + b] = uint8_t(v >> (8 * b));
    }
  }
}

}  // namespace

hipError_t launch_gf8_apply(const Gf8Call &call, uint32_t mode, hipStream_t stream) {
  if (call.len == 0 || call.nout == 0) return hipSuccess;
  const uint64_t lanes = (call.len + kLaneBytes - 1) / kLaneBytes;
  const uint32_t gx = uint32_t((lanes + kBlockThreads - 1) / kBlockThreads);
  const int nt = call.nout <= 4 ? 4 : call.nout <= 8 ? 8 : 16;
  const dim3 grid(gx, (call.nout + nt - 1) / nt);
  const dim3 block(kBlockThreads);
#define TRS_GF8(NT, MODE) hipLaunchKernelGGL((k_gf8_apply<NT, MODE>), grid, block, 0, stream, call)
  if (mode == kWrite) {
    if (nt == 4) TRS_GF8(4, kWrite);
    else if (nt == 8) TRS_GF8(8, kWrite);
    else TRS_GF8(16, kWrite);
  } else if (mode == kVerify) {
    if (nt == 4) TRS_GF8(4, kVerify);
    else if (nt == 8) TRS_GF8(8, kVerify);
    else TRS_GF8(16, kVerify);
  } else {
    if (nt == 4) TRS_GF8(4, kXorInto);
    else if (nt == 8) TRS_GF8(8, kXorInto);
    else TRS_GF8(16, kXorInto);
  }
#undef TRS_GF8
  return hipGetLastError();
}

namespace {
dim3 gf16_grid(const Gf16Call &call, int &nt) {
  const uint64_t lanes = (call.cols + kLaneBytes - 1) / kLaneBytes;
  const uint32_t gx = uint32_t((lanes + kBlockThreads - 1) / kBlockThreads);
  nt = call.nout <= 4 ? 4 : 8;
  return dim3(gx, (call.nout + nt - 1) / nt);
}
}  // namespace

hipError_t launch_gf16_apply(const Gf16Call &call, hipStream_t stream) {
  if (call.cols == 0 || call.nout == 0) return hipSuccess;
  int nt = 4;
  const dim3 grid = gf16_grid(call, nt);
  const dim3 block(kBlockThreads);
#define TRS_GF16(NT, MODE) hipLaunchKernelGGL((k_gf16_apply<NT, MODE>), grid, block, 0, stream, call)
  if (call.mode == kVerify) {
    if (nt == 4) TRS_GF16(4, kVerify);
    else TRS_GF16(8, kVerify);
  } else {
    if (nt == 4) TRS_GF16(4, kWrite);
    else TRS_GF16(8, kWrite);
  }
#undef TRS_GF16
  return hipGetLastError();
}

hipError_t launch_gf16_apply_inline(const Gf16InlineCall &call, hipStream_t stream) {
  if (call.call.cols == 0 || call.call.nout == 0) return hipSuccess;
  if (call.call.k + call.call.nout > uint32_t(kInlinePtrs)) return hipErrorInvalidValue;
  int nt = 4;
  const dim3 grid = gf16_grid(call.call, nt);
  const dim3 block(kBlockThreads);
#define TRS_GF16(NT, MODE) \
  hipLaunchKernelGGL((k_gf16_apply_inline<NT, MODE>), grid, block, 0, stream, call)
  if (call.call.mode == kVerify) {
    if (nt == 4) TRS_GF16(4, kVerify);
    else TRS_GF16(8, kVerify);
  } else {
    if (nt == 4) TRS_GF16(4, kWrite);
    else TRS_GF16(8, kWrite);
  }
#undef TRS_GF16
  return hipGetLastError();
}

hipError_t launch_gf8_planes(const Gf8Call &call, uint32_t mode, hipStream_t stream, int nt) {
  if (call.len == 0 || call.nout == 0) return hipSuccess;
  const uint32_t gx = uint32_t((call.len + kPlanesChunkBytes - 1) / kPlanesChunkBytes);
  // 8-row tiles, unless that grid leaves the chip short of work (at most 640
  // workgroups, 2.5 per CU): then 4-row tiles double the workgroups, which
  // measured 1.2-1.6x faster there (k = 33..200 at 256 KiB - 1 MiB) and
  // 1.2-1.5x slower on full grids.
  const uint64_t wg8 = uint64_t(gx) * ((call.nout + kPlanesOut - 1) / kPlanesOut);
  if (nt != 4 && nt != kPlanesOut) nt = call.nout <= 4 || wg8 <= 640 ? 4 : kPlanesOut;
  const dim3 grid(gx, (call.nout + nt - 1) / nt);
  const dim3 block(kBlockThreads);
#define TRS_GF8P(NT, MODE) hipLaunchKernelGGL((k_gf8_planes<NT, MODE>), grid, block, 0, stream, call)
  if (mode == kWrite) {
    if (nt == 4) TRS_GF8P(4, kWrite);
    else TRS_GF8P(kPlanesOut, kWrite);
  } else if (mode == kVerify) {
    if (nt == 4) TRS_GF8P(4, kVerify);
    else TRS_GF8P(kPlanesOut, kVerify);
  } else {
    if (nt == 4) TRS_GF8P(4, kXorInto);
    else TRS_GF8P(kPlanesOut, kXorInto);
  }
#undef TRS_GF8P
  return hipGetLastError();
}

hipError_t launch_batch8(const BatchJob *jobs, const uint32_t *tile_job, uint64_t total_tiles,
                         uint32_t grid, hipStream_t stream) {
  if (total_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch8, dim3(grid), dim3(kBlockThreads), 0, stream, jobs, tile_job,
                     total_tiles);
  return hipGetLastError();
}

hipError_t launch_batch8_planes(const BatchJob *jobs, const uint32_t *tile_job,
                                uint64_t total_tiles, uint32_t grid, hipStream_t stream) {
  if (total_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch8_planes, dim3(grid), dim3(kBlockThreads), 0, stream, jobs, tile_job,
                     total_tiles);
  return hipGetLastError();
}

hipError_t launch_u16_split(const uint8_t *u16, uint8_t *shard, uint64_t n, bool to_split,
                            hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint8_t *lo = shard, *hi = shard + n;
  const uintptr_t a = reinterpret_cast<uintptr_t>(u16), l = reinterpret_cast<uintptr_t>(lo),
                  h = reinterpret_cast<uintptr_t>(hi);
  const uint32_t vec_ok = (a % 16 == 0 && l % 4 == 0 && h % 4 == 0) ? 1u : 0u;
  const uint64_t blocks = (n + 8ull * kBlockThreads - 1) / (8ull * kBlockThreads);
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  if (to_split)
    hipLaunchKernelGGL(k_u16_to_split, dim3(uint32_t(blocks)), dim3(kBlockThreads), 0, stream, u16,
                       lo, hi, n, vec_ok);
  else
    hipLaunchKernelGGL(k_split_to_u16, dim3(uint32_t(blocks)), dim3(kBlockThreads), 0, stream,
                       lo, hi, const_cast<uint8_t *>(u16), n, vec_ok);
  return hipGetLastError();
}

hipError_t launch_fill(uint8_t *dst, uint64_t len, uint64_t seed, hipStream_t stream) {
  if (len == 0) return hipSuccess;
  const uint64_t words = (len + 7) / 8;
  uint64_t blocks = (words + kBlockThreads - 1) / kBlockThreads;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3(uint32_t(blocks)), dim3(kBlockThreads), 0, stream, dst, len,
                     seed);
  return hipGetLastError();
}

}  // namespace trs
