// Additive-FFT (Lin-Chung-Han) encode for the GF((2^8)^2) tower code.
//
// Host side: the staged program builder (a C++ restatement of
// src/fft16.rs:182-621) and its lowering to the device step list.  Device
// side (fft16.hip): an interpreter that keeps one workgroup's register file in
// LDS, bit-sliced (16 planes of 32 symbol columns per register), and runs each
// program step over all of its columns between barriers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include <hip/hip_runtime_api.h>

namespace trs {

// src/fft16.rs:30-38
constexpr size_t kMaxFftDataShards = 1024;
constexpr size_t kMaxStagedRegisters16 = 2048;

// kOpDeriv (decoder programs only): dst ^= XOR of the registers dst | 2^j
// for every bit j < c clear in dst (one layer of the formal derivative).
enum Fft16OpKind : uint8_t {
  kOpXor = 0,
  kOpCopy = 1,
  kOpMulXor = 2,
  kOpMulCopy = 3,
  kOpZero = 4,
  kOpDeriv = 5
};

struct Fft16Op {
  uint8_t kind;
  uint16_t dst, src, c;
};

// The lowered program (src/fft16.rs:138-174), same fields and meaning.
struct Fft16Program {
  struct Stage {
    bool transform;
    bool inverse;
    uint32_t start, size, consts_start;  // transform
    uint32_t glue_start, glue_len;       // glue run
  };
  size_t k = 0, m = 0;
  std::vector<Stage> stages;
  std::vector<uint16_t> consts;
  std::vector<uint8_t> consts_skip;
  std::vector<Fft16Op> glue;
  std::vector<uint16_t> zero_regs, parity_regs;
  size_t register_count = 0;
  size_t mult_count = 0;
};

// src/fft16.rs:357-376 (cheaper of the two base strategies).  Requires
// 1 <= k <= kMaxFftDataShards.
Fft16Program build_staged_program16(size_t k, size_t m);
// src/fft16.rs:85-106, with the multiply-count rule of x86 hosts with GFNI.
bool fft16_profitable(const Fft16Program &p, size_t k, size_t m);

// The GF(2^8) staged program (src/fft.rs:393-592) in the same form, its
// constants bytes of the subfield.  With `liveness` the lowering also drops
// dead glue and marks dead transform blocks (the tower lowering's passes,
// never changing the parity); without it the multiply and register counts
// are src/fft.rs:472-592's.  Requires 1 <= k, k + m <= 256.
constexpr size_t kMaxFftDataShards8 = 256;
constexpr size_t kMaxStagedRegisters8 = 160;  // src/fft.rs:130
Fft16Program build_staged_program8(size_t k, size_t m, bool liveness = true);
// Whether the reference's x86 build takes the staged route for (k, m)
// (src/reedsolomon.rs:204-234): power-of-two k >= 2, not a generated shape,
// at most 160 registers, 2 * multiplies < k * m.
bool staged8_reference_route(size_t k, size_t m);

// ---- device form ---------------------------------------------------------

// One barrier-separated step.  kind 0: one level of a transform, `count`
// butterflies at stride 2^b over registers [a, a + 2*count), block constants
// at consts[c + block], d = inverse.  kind 1: `count` independent glue ops at
// ops[a].
struct Fft16Step {
  uint32_t kind, count, a, b, c, d, pad0, pad1;
};

// Per-constant word: Karatsuba bytes [c_lo, c_hi ^ c_lo, c_hi * L] in bytes
// 0..2 (src/fft16.rs:176-180), flags in byte 3.
constexpr uint32_t kConstSkip = 1u << 24;  // block is a dead subtree
constexpr uint32_t kConstZero = 2u << 24;  // c == 0: butterfly without multiply
constexpr uint32_t kConstOne = 4u << 24;   // c == 1

struct Fft16Device;  // device buffers of one lowered program
void fft16_device_free(Fft16Device *d);
struct Fft16DeviceDeleter {
  void operator()(Fft16Device *d) const { fft16_device_free(d); }
};
using Fft16DevicePtr = std::unique_ptr<Fft16Device, Fft16DeviceDeleter>;

struct Fft16Call {
  const uint8_t *const *ptrs;  // device [stripes][nload + nstore]: loaded then stored shards
  uint64_t plane;              // symbols per shard (bytes per plane)
  uint32_t stripes;
  uint32_t nstore;             // stored registers (a prefix of the program's store list)
  uint32_t vec_ok;             // every shard pointer 16-byte aligned
  uint64_t cols;               // columns from each pointer (0: plane); pointers may sit at a
                               // column offset within their low plane
  uint32_t *flag;              // non-null: verify -- compare the stored registers with the
                               // shards instead of writing them, *flag = 1 on a mismatch
  uint64_t lds_target;         // LDS bytes per workgroup the column groups are sized for
                               // (0: 64 KiB for tower registers, 32 KiB for byte planes)
};

// Lowers p for the device (glue scheduled into independent sub-steps, whole
// range copies coalesced) and uploads it.  hipErrorInvalidValue: the register
// file does not fit one workgroup's LDS.
hipError_t fft16_device_new(const Fft16Program &p, Fft16DevicePtr *out);

// p lowered for the device (host only): glue scheduled into independent
// sub-steps, whole-range copies coalesced, registers renumbered.  ops holds
// two words per op: dst | src << 16, kind | Karatsuba bytes << 8.
// load[i]: register loaded from pointer slot i; store[j]: register written to
// slot load.size() + j.
struct Fft16Lowered {
  std::vector<Fft16Step> steps;
  std::vector<uint32_t> consts;
  std::vector<uint32_t> ops;
  std::vector<uint32_t> load, store;
  size_t regs = 0, max_count = 0;
};
Fft16Lowered fft16_lower(const Fft16Program &p);
// Whether p's device register file fits one workgroup's LDS.
bool fft16_fits_device(const Fft16Program &p);

// WHAT[level][x] (src/fft16.rs:42-67), level < 11.
uint16_t fft16_what(int level, size_t x);
// The byte field's WHAT[level][x] (src/fft.rs:28-79), level < 8, x < 256.
uint16_t fft8_what(int level, size_t x);

// Erasure decoding by the additive FFT (formal-derivative method; no reference
// counterpart, the reference decodes through matrix plans).  The codeword of
// length n is the evaluation of the data polynomial P (degree < k) at points
// 0..n-1; with N = next power of two >= n, E = [0, N) minus `valid` (k known
// points) and Lambda(x) = prod_{e in E} (x - e):
//   Q = P * Lambda is known at all N points (zero on E), so an inverse
//   transform gives its coefficients; at e in E, Q'(e) = P(e) Lambda'(e), so a
//   formal derivative, a forward transform and a scale by 1 / Lambda'(e)
//   recover P(e) for every requested output e.
// Registers: one per point (N <= kMaxDecodePoints); load = valid, store =
// outputs (in the given order).
constexpr size_t kMaxDecodePoints = 2048;
Fft16Lowered fft16_decode_lowered(size_t n, const std::vector<uint32_t> &valid,
                                  const std::vector<uint32_t> &outputs);
// The same decoder over GF(2^8) (n <= 256 points; no reference counterpart),
// for the byte-plane interpreter (launch_fft16 with byte_planes).
Fft16Lowered fft8_decode_lowered(size_t n, const std::vector<uint32_t> &valid,
                                 const std::vector<uint32_t> &outputs);
hipError_t fft16_device_upload(const Fft16Lowered &L, Fft16DevicePtr *out);
size_t fft16_device_loads(const Fft16Device &d);
size_t fft16_device_stores(const Fft16Device &d);
// Shard pointers of one stripe passed in the kernel arguments: call.ptrs is
// ignored, nload + nstore <= kFftInlinePtrs, stripes == 1.
constexpr size_t kFftInlinePtrs = 256;
struct Fft16Ptrs {
  const uint8_t *p[kFftInlinePtrs];
};
// byte_planes: the program's constants are all bytes (a GF(2^8) program,
// build_staged_program8) and the shards are plain byte shards: one byte per
// column, call.plane = shard length, registers of 8 bit planes.  Otherwise
// tower symbols of split-plane shards (16 planes, high plane `plane` bytes
// after the low).
hipError_t launch_fft16(const Fft16Device &d, const Fft16Call &call, hipStream_t stream,
                        const Fft16Ptrs *inline_ptrs = nullptr, bool byte_planes = false);
// Column groups per workgroup (log2) a launch of d over `cols` columns takes
// with LDS target `lds_target` (0: the default); launches with equal values
// are the same launch.
uint32_t fft16_launch_gshift(const Fft16Device &d, uint64_t cols, bool byte_planes,
                             uint64_t lds_target);

// One stripe of one program; launch_fft16_jobs runs many in one launch.
struct Fft16JobSpec {
  const Fft16Device *prog;
  const uint8_t *const *ptrs;  // device table: nload loaded then nstore stored shards
  uint64_t plane;
  uint32_t nstore;
  uint32_t vec_ok;
  // Indexed addressing instead of ptrs (when idx is set): slot i is shard
  // idx[i] (device array) of the stripe at base, shards stride bytes apart.
  const uint32_t *idx = nullptr;
  const uint8_t *base = nullptr;
  uint64_t stride = 0;
};
hipError_t launch_fft16_jobs(const std::vector<Fft16JobSpec> &jobs, hipStream_t stream,
                             bool byte_planes = false);

// The same launch in two parts, for callers that upload once and launch
// many times (batch work lists, graph replays): the host builds `blob` (job
// table and workgroup -> job map), the caller copies it to device memory,
// and fft16_jobs_launch launches the kernel over that copy.
struct Fft16JobsPlan {
  std::vector<uint8_t> blob;
  size_t map_offset = 0, lds = 0;
  uint32_t total = 0, threads = 0;
  bool byte_planes = false;
};
hipError_t fft16_jobs_plan(const std::vector<Fft16JobSpec> &jobs, bool byte_planes,
                           Fft16JobsPlan *plan);
hipError_t fft16_jobs_launch(const Fft16JobsPlan &plan, const void *dev_blob, hipStream_t stream);

// Host-to-device copy of a small per-call table (pointer tables, job
// tables) through a per-thread ring of pinned staging buffers, so the copy
// is asynchronous: the host copies `src` into a staging buffer whose previous
// transfer has finished and returns after enqueueing the transfer on `stream`.
hipError_t staged_upload(void *dst, const void *src, size_t bytes, hipStream_t stream);
// Frees the calling thread's staging ring (after its transfers finish); for
// threads that end before the process does.
void staged_upload_release();

// Scratch of the allocation-free entry point: a file-scope device array of
// kNoAllocScratchBytes in fft16.hip.  noalloc_scratch returns its address on
// the current device (looked up once per device); every call on one device
// shares it.
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
hipError_t noalloc_scratch(void **out);

}  // namespace trs
