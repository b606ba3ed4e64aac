// Bit-sliced GF(2^8) matrix apply on int8 MFMA (gfx950).
//
// Multiplying by c in GF(2^8) is GF(2)-linear, so an e x k coefficient matrix
// is an (8e) x (8k) 0/1 matrix A with A[8j+r][8i+b] = bit r of (c_ji * 2^b)
// (the bit-matrix GFNI's affine instruction takes, src/gf/gfni.rs:41-64).
// With B[8i+b][col] = bit b of input i at byte column col, the integer
// product C = A*B counts, for every output bit, how many set terms it sums;
// its low bit is the GF(2) sum.  The int32 accumulation is exact (at most
// 8*256 terms), so the result is byte-identical to galois::mul-based coding.
//
// Tiling (v_mfma_i32_32x32x32_i8: M = N = K = 32):
//  * K = 32 bits = 4 inputs.  Lane l supplies K elements 16*(l/32)+e,
//    e = 0..15: bits of inputs 2*(l/32) and 2*(l/32)+1 of the K step.  A and
//    B use the same (lane half, element) -> K mapping, which is all an MFMA
//    needs.
//  * N = 32 byte columns.  A lane loads one dword (4 consecutive columns) of
//    each of its two inputs and feeds byte t of it to column tile t, so four
//    MFMA column tiles cover 128 consecutive columns per wave.
//  * M = 32 rows = 4 outputs x 8 bits, two M tiles per pass (8 outputs).  The
//    accumulator of lane l, register v holds row 8*(v/4) + 4*(l/32) + v%4 of
//    column l%32; the host orders A's rows so that registers 0..7 of a lane
//    are the 8 bits of one output byte and 8..15 those of the next, letting
//    every lane pack whole output bytes without talking to other lanes.
//  * blockIdx.y = pass (outputs 8p .. 8p+7).  Each wave keeps its A fragments
//    in VGPRs and walks kBlocksPerWave consecutive 128-column blocks.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "uniform.hpp"

namespace trs {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kBlocksPerWave = 4;  // 128-column blocks per wave

// 4 bits -> 4 bytes of 0/1 (bit i to byte i): the partial products of
// t * 0x204081 land on bits i + 7i and never overlap.
__device__ __forceinline__ uint32_t spread4(uint32_t t) {
  return (t * 0x00204081u) & 0x01010101u;  // t < 16: one v_mul_u32_u24
}

__device__ __forceinline__ uint32_t pack8(const v16i &a, int base) {
  uint32_t b = 0;
#pragma unroll
  for (int r = 0; r < 8; ++r) b |= (uint32_t(a[base + r]) & 1u) << r;
  return b;
}

// Loads the dword of columns [col, col+4) of p; lanes at the tail or with an
// unaligned shard assemble the valid bytes.
__device__ __forceinline__ uint32_t load4(const uint8_t *p, uint64_t col, uint64_t len,
                                          bool aligned) {
  if (aligned && col + 4 <= len) return ld4(p + col);
  uint32_t v = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b)
    if (col + b < len) v |= uint32_t(ld1(p + col + b)) << (8 * b);
  return v;
}

__device__ __forceinline__ void store4(uint8_t *p, uint64_t col, uint64_t len, bool aligned,
                                       uint32_t v) {
  if (aligned && col + 4 <= len) {
    st4(p + col, v);
    return;
  }
#pragma unroll
  for (int b = 0; b < 4; ++b)
    if (col + b < len) st1(p + col + b, uint8_t(v >> (8 * b)));
}

// Input dwords of one 128-column block: lane l reads columns 4*(l%32)..+3 of
// inputs 2h and 2h+1 of every K step (h = lane half).  FULL: every column in
// range and every shard 4-byte aligned, so the loads are plain dwords.
// Pointers are read with uniform indices (4 per K step) and each lane half
// selects its pair, so no pointer fetch depends on the lane; an input past k
// reads shard 0 and is zeroed.
template <int KS, bool FULL, typename PtrAt>
__device__ __forceinline__ void mfma_load(PtrAt ptr_at, uint64_t len, uint32_t k,
                                          uint64_t blk_col, bool aligned, uint32_t (&d0s)[KS],
                                          uint32_t (&d1s)[KS]) {
  const int lane = threadIdx.x & 63;
  const bool upper = (lane >> 5) != 0;
  const uint64_t col = blk_col + 4 * (lane & 31);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const uint8_t *q0 = ptr_at(4 * s);
    const uint8_t *q1 = (4 * s + 1 < k) ? ptr_at(4 * s + 1) : q0;
    const uint8_t *q2 = (4 * s + 2 < k) ? ptr_at(4 * s + 2) : q0;
    const uint8_t *q3 = (4 * s + 3 < k) ? ptr_at(4 * s + 3) : q0;
    const uint8_t *pa = upper ? q2 : q0;
    const uint8_t *pb = upper ? q3 : q1;
    const bool va = (4 * s + (upper ? 2u : 0u)) < k;
    const bool vb = (4 * s + (upper ? 3u : 1u)) < k;
    uint32_t x0, x1;
    if (FULL) {
      x0 = ld4(pa + col);
      x1 = ld4(pb + col);
    } else {
      x0 = col < len ? load4(pa, col, len, aligned) : 0u;
      x1 = col < len ? load4(pb, col, len, aligned) : 0u;
    }
    d0s[s] = va ? x0 : 0u;
    d1s[s] = vb ? x1 : 0u;
  }
}

// MFMAs and output packing of one 128-column block from its input dwords.
template <int KS, typename PtrAt>
__device__ __forceinline__ void mfma_compute(PtrAt ptr_at, const v4i (&A0)[KS],
                                             const v4i (&A1)[KS], const uint32_t (&d0s)[KS],
                                             const uint32_t (&d1s)[KS], uint64_t len, uint32_t k,
                                             uint32_t nout, uint32_t pass, uint64_t blk_col,
                                             bool full, bool aligned) {
  const int lane = threadIdx.x & 63;
  const bool upper = (lane >> 5) != 0;
  const uint64_t col = blk_col + 4 * (lane & 31);
  v16i acc[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    acc[t][0] = v16i{};
    acc[t][1] = v16i{};
  }
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const uint32_t d0 = d0s[s], d1 = d1s[s];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      v4i B;
      B[0] = int(spread4((d0 >> (8 * t)) & 0xfu));
      B[1] = int(spread4((d0 >> (8 * t + 4)) & 0xfu));
      B[2] = int(spread4((d1 >> (8 * t)) & 0xfu));
      B[3] = int(spread4((d1 >> (8 * t + 4)) & 0xfu));
      acc[t][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0[s], B, acc[t][0], 0, 0, 0);
      acc[t][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1[s], B, acc[t][1], 0, 0, 0);
    }
  }
  if (!full && col >= len) return;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int oo = 0; oo < 2; ++oo) {
      // lane half h writes output 2h + oo of each M tile
      const uint32_t o_lo = pass * 8 + mt * 4 + oo;
      const uint32_t o_hi = o_lo + 2;
      uint32_t word = 0;
#pragma unroll
      for (int t = 0; t < 4; ++t) word |= pack8(acc[t][mt], 8 * oo) << (8 * t);
      const uint32_t o = upper ? o_hi : o_lo;
      if (o < nout) {
        const uint8_t *lo_ptr = o_lo < nout ? ptr_at(k + o_lo) : nullptr;
        const uint8_t *hi_ptr = o_hi < nout ? ptr_at(k + o_hi) : nullptr;
        uint8_t *dst = const_cast<uint8_t *>(upper ? hi_ptr : lo_ptr);
        if (full)
          st4(dst + col, word);
        else
          store4(dst, col, len, aligned, word);
      }
    }
}

// One wave's share: kBlocksPerWave 128-column blocks from wave_col, outputs
// 8*pass .. 8*pass+7, with the next block's inputs loaded while the current
// block runs its MFMAs.  ptr_at(i) returns shard pointer i (inputs first).
template <int KS, typename PtrAt>
__device__ __forceinline__ void mfma_wave(PtrAt ptr_at, const void *frag_base, uint64_t len,
                                          uint32_t k, uint32_t nout, uint32_t pass,
                                          uint64_t wave_col, bool aligned) {
  const int lane = threadIdx.x & 63;
  v4i A0[KS], A1[KS];
  const v4i *frag = reinterpret_cast<const v4i *>(frag_base) + size_t(pass) * KS * 2 * 64;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    A0[s] = frag[(s * 2 + 0) * 64 + lane];
    A1[s] = frag[(s * 2 + 1) * 64 + lane];
  }
  // Fast path: the wave's whole span is in range and aligned.
  if (aligned && wave_col + 128ull * kBlocksPerWave <= len) {
    uint32_t c0[KS], c1[KS], n0[KS], n1[KS];
    mfma_load<KS, true>(ptr_at, len, k, wave_col, true, c0, c1);
#pragma unroll
    for (int blk = 0; blk < kBlocksPerWave; ++blk) {
      const uint64_t blk_col = wave_col + uint64_t(blk) * 128;
      if (blk + 1 < kBlocksPerWave) mfma_load<KS, true>(ptr_at, len, k, blk_col + 128, true, n0, n1);
      mfma_compute<KS>(ptr_at, A0, A1, c0, c1, len, k, nout, pass, blk_col, true, true);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        c0[s] = n0[s];
        c1[s] = n1[s];
      }
    }
    return;
  }
  for (int blk = 0; blk < kBlocksPerWave; ++blk) {
    const uint64_t blk_col = wave_col + uint64_t(blk) * 128;
    if (blk_col >= len) break;
    uint32_t c0[KS], c1[KS];
    mfma_load<KS, false>(ptr_at, len, k, blk_col, aligned, c0, c1);
    mfma_compute<KS>(ptr_at, A0, A1, c0, c1, len, k, nout, pass, blk_col, false, aligned);
  }
}

constexpr uint64_t kWaveCols = 128ull * kBlocksPerWave;
constexpr uint64_t kGroupCols = uint64_t(kBlockThreads / 64) * kWaveCols;

template <int KS>
__global__ __launch_bounds__(kBlockThreads) void k_gf8_mfma(const Gf8MfmaCall a) {
  const uint64_t wave_col = uint64_t(blockIdx.x) * kGroupCols + (threadIdx.x >> 6) * kWaveCols;
  if (wave_col >= a.len) return;
  mfma_wave<KS>([&](uint32_t i) { return uni(a.ptr[i]); }, a.frag, a.len, a.k, a.nout, blockIdx.y,
                wave_col, a.vec_ok != 0);
}

// Batched form: tiles are (job, pass, 2048-column group), numbered job by job;
// tile_job[t] names the job of tile t and a fixed grid strides over them.
// J.coef carries the A fragments.
__global__ __launch_bounds__(kBlockThreads) void k_batch_mfma(const BatchJob *jobs,
                                                             const uint32_t *tile_job,
                                                             uint64_t total) {
  for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
    // job fields are the same in every lane: scalar loads
    const auto &J = ld_const_ref(jobs + ld_const(tile_job + t));
    const uint32_t nout = J.nout, k = J.k;
    const uint64_t len = J.len;
    const uint32_t passes = (nout + 7) / 8;
    const uint64_t local = t - J.tile_begin;
    const uint64_t group = local / passes;
    const uint32_t pass = uint32_t(local - group * passes);
    const uint64_t wave_col = group * kGroupCols + uni(uint32_t(threadIdx.x >> 6)) * kWaveCols;
    if (wave_col >= len) continue;
    const uint8_t *const *ptrs = J.ptrs;
    const void *frag = J.coef;
    auto at = [&](uint32_t i) { return ld_const_ptr(ptrs + i); };
    const bool al = J.vec_ok != 0;
    switch ((k + 3) / 4) {
      case 1: mfma_wave<1>(at, frag, len, k, nout, pass, wave_col, al); break;
      case 2: mfma_wave<2>(at, frag, len, k, nout, pass, wave_col, al); break;
      case 3: mfma_wave<3>(at, frag, len, k, nout, pass, wave_col, al); break;
      case 4: mfma_wave<4>(at, frag, len, k, nout, pass, wave_col, al); break;
      case 5: mfma_wave<5>(at, frag, len, k, nout, pass, wave_col, al); break;
      case 6: mfma_wave<6>(at, frag, len, k, nout, pass, wave_col, al); break;
      case 7: mfma_wave<7>(at, frag, len, k, nout, pass, wave_col, al); break;
      case 8: mfma_wave<8>(at, frag, len, k, nout, pass, wave_col, al); break;
      default: break;
    }
  }
}

}  // namespace

uint32_t mfma_steps(uint32_t k) { return (k + 3) / 4; }

hipError_t launch_gf8_mfma(const Gf8MfmaCall &a, hipStream_t stream) {
  if (a.len == 0 || a.nout == 0) return hipSuccess;
  const uint32_t ks = mfma_steps(a.k);
  const dim3 grid(uint32_t((a.len + kGroupCols - 1) / kGroupCols), (a.nout + 7) / 8);
  const dim3 block(kBlockThreads);
#define TRS_MFMA(N) \
  case N: hipLaunchKernelGGL((k_gf8_mfma<N>), grid, block, 0, stream, a); break
  switch (ks) {
    TRS_MFMA(1); TRS_MFMA(2); TRS_MFMA(3); TRS_MFMA(4); TRS_MFMA(5); TRS_MFMA(6);
    TRS_MFMA(7); TRS_MFMA(8);
    default: return hipErrorNotSupported;
  }
#undef TRS_MFMA
  return hipGetLastError();
}

bool mfma_available(uint32_t k) { return mfma_steps(k) <= 8 && k > 0; }

uint64_t mfma_group_cols() { return kGroupCols; }

hipError_t launch_batch_mfma(const BatchJob *jobs, const uint32_t *tile_job,
                             uint64_t total_tiles, uint32_t grid, hipStream_t stream) {
  if (total_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_mfma, dim3(grid), dim3(kBlockThreads), 0, stream, jobs, tile_job,
                     total_tiles);
  return hipGetLastError();
}

}  // namespace trs
