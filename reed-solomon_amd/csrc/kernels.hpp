// Launch descriptors shared by the host codec (codec.cpp) and the HIP
// kernels (kernels.hip).  Plain structs only: the host fills them, the
// launchers pass them to the kernels by value.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime_api.h>

namespace trs {

// Pointers carried inline in the kernel arguments of one GF(2^8) apply.
// RS8 stripes have at most 256 shards, so inputs + outputs always fit.
constexpr int kInlinePtrs = 256;

// Output rows per workgroup tile; coefficient rows are padded to this.
constexpr int kOutStride = 16;

// Bytes of every shard handled by one lane of one workgroup.
constexpr int kLaneBytes = 16;
constexpr int kBlockThreads = 256;
constexpr uint64_t kChunkBytes = uint64_t(kLaneBytes) * kBlockThreads;  // 4 KiB

enum ApplyMode : uint32_t {
  kWrite = 0,   // out = M * in
  kVerify = 1,  // flag |= (out != M * in)
  kXorInto = 2  // out ^= M * in
};

// One GF(2^8) matrix apply: nout outputs from k inputs over `len` bytes.
// coef is the transposed, padded coefficient matrix in device memory:
// byte (i * ostride + j) is row j's coefficient for input i.
struct Gf8Call {
  const uint8_t *ptr[kInlinePtrs];  // [0,k) inputs, [k,k+nout) outputs
  const uint32_t *coef;
  uint64_t len;
  uint32_t k, nout, ostride, vec_ok;
  uint32_t *flag;  // verify mismatch flag (kVerify only)
};

// One tower (GF((2^8)^2)) matrix apply over split-plane shards.  ptrs is a
// device array [k inputs, nout outputs]; every shard is `plane` low bytes
// then `plane` high bytes; symbols col0 .. col0+cols-1 are applied.  coef holds three transposed padded byte matrices
// (Karatsuba products c_lo, c_hi^c_lo, c_hi*L), each k * ostride bytes.
struct Gf16Call {
  const uint8_t *const *ptrs;
  const uint32_t *coef;
  uint64_t plane;       // bytes per plane (the high plane starts at this offset)
  uint64_t col0, cols;  // symbol window [col0, col0 + cols) of the planes to apply
  uint32_t k, nout, ostride, vec_ok;
  uint32_t mode;
  uint32_t *flag;
};

// The same apply with up to kInlinePtrs shard pointers in the kernel
// arguments (call.ptrs unused), so a call needs no device pointer table.
struct Gf16InlineCall {
  Gf16Call call;
  const uint8_t *ptr[kInlinePtrs];  // [0,k) inputs, [k,k+nout) outputs
};

// One GF(2^8) apply on int8 MFMA (mfma.hip): frag holds the bit-expanded
// coefficient matrix as per-lane A fragments, [pass][K step][2 M tiles][64
// lanes][16 bytes]; pass p covers outputs 8p..8p+7, K step s inputs 4s..4s+3.
// vec_ok: every pointer 4-byte aligned.
struct Gf8MfmaCall {
  const uint8_t *ptr[kInlinePtrs];  // [0,k) inputs, [k,k+nout) outputs
  const void *frag;
  uint64_t len;
  uint32_t k, nout, vec_ok, pad;
};

// One entry of a batched GF(2^8) work list (device memory).
struct BatchJob {
  const uint8_t *const *ptrs;  // device array [k inputs, nout outputs]
  const uint32_t *coef;        // transposed padded coefficients (as Gf8Call); MFMA jobs:
                               // the A fragments of mfma.hip
  uint64_t len;
  uint64_t tile_begin;         // first global tile of this job
  uint32_t k, nout, ostride, vec_ok;
  uint32_t out_tiles;          // ceil(nout / kOutStride)
  uint32_t prog;               // generated bit-sliced program id (k_batch_bs), 0 = none
};


// Host-side launchers (kernels.hip).  All return hipError_t of the launch.
hipError_t launch_gf8_apply(const Gf8Call &call, uint32_t mode, hipStream_t stream);
// The same apply on the bit-sliced tile (8 KiB chunks; rows per tile `nt`, 4
// or kPlanesOut, 0 = the launch rule below).
hipError_t launch_gf8_planes(const Gf8Call &call, uint32_t mode, hipStream_t stream, int nt = 0);
hipError_t launch_gf16_apply(const Gf16Call &call, hipStream_t stream);
hipError_t launch_gf16_apply_inline(const Gf16InlineCall &call, hipStream_t stream);
// Batched persistent kernels: tile_job[t] is the job index of tile t.
hipError_t launch_batch8(const BatchJob *jobs, const uint32_t *tile_job, uint64_t total_tiles,
                         uint32_t grid, hipStream_t stream);
// The same work list on the bit-sliced runtime-coefficient tile (8 KiB
// column chunks, 32 bytes per lane as bit planes).
constexpr uint64_t kPlanesChunkBytes = 8192;
constexpr int kPlanesOut = 8;  // output rows per planes tile (a multiple of 4; 8 measured best)
#ifndef TRS_BATCH_PLANES_OUT
#define TRS_BATCH_PLANES_OUT 8
#endif
// Output rows per tile of the batched reconstruct kernel (k_batch8_planes).
constexpr int kBatchPlanesOut = TRS_BATCH_PLANES_OUT;
hipError_t launch_batch8_planes(const BatchJob *jobs, const uint32_t *tile_job,
                                uint64_t total_tiles, uint32_t grid, hipStream_t stream);
hipError_t launch_fill(uint8_t *dst, uint64_t len, uint64_t seed, hipStream_t stream);
// n interleaved u16 symbols at `u16` <-> the split-plane shard at `shard`
// ([low n | high n]); to_split selects the direction.
hipError_t launch_u16_split(const uint8_t *u16, uint8_t *shard, uint64_t n, bool to_split,
                            hipStream_t stream);

// int8-MFMA apply (mfma.hip); available for k <= 32.
hipError_t launch_gf8_mfma(const Gf8MfmaCall &call, hipStream_t stream);
bool mfma_available(uint32_t k);
uint32_t mfma_steps(uint32_t k);
uint64_t mfma_group_cols();  // columns per workgroup tile (4 waves x 512)
// Persistent work-list kernel over MFMA jobs (BatchJob::coef = A fragments).
hipError_t launch_batch_mfma(const BatchJob *jobs, const uint32_t *tile_job,
                             uint64_t total_tiles, uint32_t grid, hipStream_t stream);

// Bit-sliced encode with compile-time coefficients (gen_bitsliced.hip,
// generated by gf8_codegen.py).  route: kRouteBsMatrix or kRouteBsFft.
// kRouteStaged8: the GF(2^8) staged FFT program on the tower interpreter
// (codec.cpp), not a bit-sliced kernel.
enum EncodeRoute : int { kRouteTable = 0, kRouteBsMatrix = 1, kRouteBsFft = 2, kRouteStaged8 = 3 };
hipError_t launch_bs_encode(int route, const Gf8Call &call, hipStream_t stream);
// The same programs comparing their outputs with call.ptr[k..k+m) instead of
// storing them: *call.flag = 1 on a mismatch (verify).
hipError_t launch_bs_verify(int route, const Gf8Call &call, hipStream_t stream);
bool bs_encode_available(int route, uint32_t k, uint32_t m);
uint32_t bs_program_id(int route, uint32_t k, uint32_t m);
// Persistent work-list kernels over jobs with a generated program (8 KiB
// tiles), one kernel per program id 1..bs_program_count(); every job of a
// launch runs program `prog`.
constexpr uint64_t kBsChunkBytes = 8192;
uint32_t bs_program_count();
hipError_t launch_batch_bs(uint32_t prog, const BatchJob *jobs, const uint32_t *tile_job,
                           uint64_t total_tiles, uint32_t grid, hipStream_t stream);

}  // namespace trs
