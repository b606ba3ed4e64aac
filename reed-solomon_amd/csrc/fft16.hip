// Tower-FFT encode on the GPU: device lowering of a staged program and the
// LDS interpreter kernel.
//
// Work decomposition.  Every symbol column of a stripe runs the same program
// independently (src/fft16.rs:627-690 replays it per column), so a workgroup
// owns G groups of 32 consecutive columns of one stripe and keeps their whole
// register file in LDS.  A register of one group is 16 bit planes of 32
// columns (planes 0-7: bits of the low bytes, 8-15: bits of the high bytes),
// 64 bytes, so a (1024,1024) program of 1024 registers takes 64 KiB for one
// group.  Each program step (one butterfly level of a transform, or one set
// of independent glue ops) is spread over the threads; steps are separated by
// barriers.
//
// Arithmetic.  A tower product c * x with a per-lane constant c runs on the
// planes as the three GF(2^8) products of the Karatsuba split
// (src/fft16.rs:176-180): t0 = c_lo x_lo, t1 = (c_hi ^ c_lo)(x_hi ^ x_lo),
// t2 = (c_hi L) x_hi; (c x)_lo = t0 ^ t2, (c x)_hi = t0 ^ t1.  A GF(2^8)
// product a * v of 32 columns is sum over set bits b of a of v * 2^b, where
// v * 2 on planes is a rotation plus three XORs (poly 0x11d), and the bit
// test is a full-word mask: 85 VALU operations per product and 32 columns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "bitslice.hpp"
#include "fft16.hpp"
#include "field.hpp"

namespace trs {

struct Fft16Device {
  size_t nload = 0, nstore = 0, regs = 0;
  size_t nsteps = 0, max_count = 0;
  void *mem = nullptr;  // one allocation: steps | consts | ops | load regs | store regs
  const Fft16Step *steps = nullptr;
  const uint32_t *consts = nullptr;
  const uint2 *ops = nullptr;
  const uint32_t *load = nullptr;
  const uint32_t *store = nullptr;
};

size_t fft16_device_loads(const Fft16Device &d) { return d.nload; }
size_t fft16_device_stores(const Fft16Device &d) { return d.nstore; }

void fft16_device_free(Fft16Device *d) {
  if (!d) return;
  if (d->mem) (void)hipFree(d->mem);
  delete d;
}

namespace {

constexpr uint32_t kGroupCols = 32;     // columns of one bit-sliced group
constexpr uint32_t kRegBytes = 64;      // 16 planes x 4 bytes
constexpr size_t kLdsTarget = 65536;    // two workgroups per CU
constexpr size_t kLdsMax = 160 * 1024;  // gfx950 LDS per workgroup
constexpr uint32_t kMaxGroupsPerWg = 8;

// Device memory of the allocation-free entry point: every GPU the library runs
// on holds one copy of this array until the process ends.  Host code reaches
// it only through the per-device address noalloc_scratch returns (the array's
// name is a host-side stand-in, not a device address).
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

uint32_t kara_word(uint16_t c) {
  const uint8_t hi = uint8_t(c >> 8), lo = uint8_t(c & 0xff);
  const uint8_t a0 = lo, a1 = uint8_t(hi ^ lo), a2 = gf8().mul[hi][kTowerL];
  uint32_t w = uint32_t(a0) | (uint32_t(a1) << 8) | (uint32_t(a2) << 16);
  if (c == 0) w |= kConstZero;
  if (c == 1) w |= kConstOne;
  return w;
}

// ---- host lowering ----------------------------------------------------------

// Merges the destination of a COPY into its source when the source is dead
// after the copy and the destination is first written there, as long as every
// transform range still maps to a contiguous register range.  A full-size
// folded chunk (src/fft16.rs:339-341) copies the whole base this way, so the
// (1024,1024) program needs 1024 registers instead of 2048.
void coalesce_and_compact(Fft16Program &p) {
  const size_t nreg = p.register_count;
  constexpr size_t kNever = ~size_t(0);
  std::vector<size_t> first(nreg, kNever), last(nreg, 0);
  // Event positions in program order: one per glue op, one per transform.
  std::vector<size_t> op_pos(p.glue.size());
  size_t pos = 0;
  auto touch = [&](size_t r, size_t at) {
    if (first[r] == kNever) first[r] = at;
    last[r] = std::max(last[r], at);
  };
  for (size_t s = 0; s < p.stages.size(); ++s) {
    const auto &st = p.stages[s];
    if (st.transform) {
      ++pos;
      for (size_t r = st.start; r < size_t(st.start) + st.size; ++r) touch(r, pos);
    } else {
      for (size_t j = st.glue_start; j < size_t(st.glue_start) + st.glue_len; ++j) {
        op_pos[j] = ++pos;
        touch(p.glue[j].dst, pos);
        touch(p.glue[j].src, pos);
      }
    }
  }
  for (uint16_t r : p.parity_regs) touch(r, ~size_t(0) - 1);
  for (uint16_t r : p.zero_regs) first[r] = 0;  // read before any write

  std::vector<size_t> parent(nreg);
  std::iota(parent.begin(), parent.end(), size_t(0));
  for (size_t j = 0; j < p.glue.size(); ++j) {
    const Fft16Op &op = p.glue[j];
    if (op.kind != kOpCopy || op.dst == op.src) continue;
    if (first[op.dst] == op_pos[j] && last[op.src] == op_pos[j]) parent[op.dst] = op.src;
  }
  auto root = [&](size_t r) {
    while (parent[r] != r) r = parent[r];
    return r;
  };
  for (bool changed = true; changed;) {
    changed = false;
    for (const auto &st : p.stages) {
      if (!st.transform) continue;
      const size_t base = root(st.start);
      bool ok = true;
      for (size_t i = 0; i < st.size && ok; ++i) ok = root(st.start + i) == base + i;
      if (ok) continue;
      for (size_t i = 0; i < st.size; ++i)
        if (parent[st.start + i] != st.start + i) {
          parent[st.start + i] = st.start + i;
          changed = true;
        }
    }
  }
  // Rewrite, dropping copies that became self-copies.
  std::vector<size_t> r2(nreg);
  for (size_t r = 0; r < nreg; ++r) r2[r] = root(r);
  std::vector<uint8_t> used(nreg, 0);
  std::vector<Fft16Op> glue;
  std::vector<Fft16Program::Stage> stages;
  for (const auto &st : p.stages) {
    if (st.transform) {
      Fft16Program::Stage t = st;
      t.start = uint32_t(r2[st.start]);
      for (size_t i = 0; i < st.size; ++i) used[t.start + i] = 1;
      stages.push_back(t);
      continue;
    }
    const uint32_t g0 = uint32_t(glue.size());
    for (size_t j = st.glue_start; j < size_t(st.glue_start) + st.glue_len; ++j) {
      Fft16Op op = p.glue[j];
      op.dst = uint16_t(r2[op.dst]);
      op.src = uint16_t(r2[op.src]);
      if (op.kind == kOpCopy && op.dst == op.src) continue;
      used[op.dst] = used[op.src] = 1;
      glue.push_back(op);
    }
    if (glue.size() > g0) stages.push_back({false, false, 0, 0, 0, g0, uint32_t(glue.size() - g0)});
  }
  for (auto &r : p.parity_regs) used[r = uint16_t(r2[r])] = 1;
  for (auto &r : p.zero_regs) used[r = uint16_t(r2[r])] = 1;
  for (size_t r = 0; r < std::min(p.k, nreg); ++r) used[r] = 1;  // data registers stay put
  std::vector<size_t> idx(nreg);
  size_t n = 0;
  for (size_t r = 0; r < nreg; ++r) idx[r] = used[r] ? n++ : 0;
  for (auto &st : stages)
    if (st.transform) st.start = uint32_t(idx[st.start]);
  for (auto &op : glue) {
    op.dst = uint16_t(idx[op.dst]);
    op.src = uint16_t(idx[op.src]);
  }
  for (auto &r : p.parity_regs) r = uint16_t(idx[r]);
  for (auto &r : p.zero_regs) r = uint16_t(idx[r]);
  p.glue.swap(glue);
  p.stages.swap(stages);
  p.register_count = n;
}

std::pair<uint32_t, uint32_t> encode_op(const Fft16Op &op) {
  uint32_t w = 0;
  if (op.kind == kOpMulXor || op.kind == kOpMulCopy) w = kara_word(op.c) & 0xffffffu;
  if (op.kind == kOpDeriv) w = op.c;  // the transform's level count
  return {uint32_t(op.dst) | (uint32_t(op.src) << 16), uint32_t(op.kind) | (w << 8)};
}

}  // namespace

namespace {

using OpWords = std::pair<uint32_t, uint32_t>;

void glue_step(Fft16Lowered &L, const std::vector<OpWords> &ops) {
  if (ops.empty()) return;
  L.steps.push_back({1, uint32_t(ops.size()), uint32_t(L.ops.size() / 2), 0, 0, 0, 0, 0});
  for (const auto &o : ops) {
    L.ops.push_back(o.first);
    L.ops.push_back(o.second);
  }
  L.max_count = std::max(L.max_count, ops.size());
}

// Schedules a run of glue ops (in program order) into steps of mutually
// independent ops: no op of a step reads or writes a register that another op
// of the step writes.
void schedule_glue(Fft16Lowered &L, const Fft16Op *ops, size_t n) {
  std::vector<int> lw(L.regs, -1), lr(L.regs, -1);
  std::vector<std::vector<OpWords>> levels;
  for (size_t j = 0; j < n; ++j) {
    const Fft16Op &op = ops[j];
    const bool reads_dst = op.kind == kOpXor || op.kind == kOpMulXor;
    const int lvl = std::max({lw[op.src] + 1, lw[op.dst] + 1, lr[op.dst] + 1});
    if (size_t(lvl) >= levels.size()) levels.resize(lvl + 1);
    levels[lvl].push_back(encode_op(op));
    lw[op.dst] = std::max(lw[op.dst], lvl);
    lr[op.src] = std::max(lr[op.src], lvl);
    if (reads_dst) lr[op.dst] = std::max(lr[op.dst], lvl);
  }
  for (const auto &lv : levels) glue_step(L, lv);
}

// One step per transform level with at least one live block.  consts[b] /
// skip[b] run over the transform's blocks in the order of src/fft16.rs:
// 593-607 (levels low to high for the inverse, high to low forward).
void emit_transform(Fft16Lowered &L, uint32_t start, uint32_t size, bool inverse,
                    const uint16_t *consts, const uint8_t *skip) {
  int levels = 0;
  while ((size_t(1) << levels) < size) ++levels;
  uint32_t coff = 0;
  for (int step = 0; step < levels; ++step) {
    const int level = inverse ? step : levels - 1 - step;
    const uint32_t nblocks = size >> (level + 1);
    bool any = false;
    for (uint32_t b = 0; b < nblocks; ++b) any |= !skip[coff + b];
    if (any) {
      const uint32_t c0 = uint32_t(L.consts.size());
      for (uint32_t b = 0; b < nblocks; ++b)
        L.consts.push_back(kara_word(consts[coff + b]) | (skip[coff + b] ? kConstSkip : 0u));
      L.steps.push_back({0, size / 2, start, uint32_t(level), c0, inverse ? 1u : 0u, 0, 0});
      L.max_count = std::max<size_t>(L.max_count, size / 2);
    }
    coff += nblocks;
  }
}

}  // namespace

Fft16Lowered fft16_lower(const Fft16Program &program) {
  Fft16Program p = program;
  coalesce_and_compact(p);
  Fft16Lowered L;
  L.regs = p.register_count;
  {
    std::vector<OpWords> z;
    for (uint16_t r : p.zero_regs) z.push_back(encode_op({kOpZero, r, r, 0}));
    glue_step(L, z);
  }
  for (const auto &st : p.stages) {
    if (st.transform)
      emit_transform(L, st.start, st.size, st.inverse, p.consts.data() + st.consts_start,
                     p.consts_skip.data() + st.consts_start);
    else
      schedule_glue(L, p.glue.data() + st.glue_start, st.glue_len);
  }
  for (size_t r = 0; r < p.k; ++r) L.load.push_back(uint32_t(r));
  for (uint16_t r : p.parity_regs) L.store.push_back(r);
  return L;
}

namespace {

// The field a decoder is built over: the tower (2^16 points) or the byte
// field (its subfield, 2^8 points; constants < 256 lower to the Karatsuba
// words [c, c, 0] the byte-plane kernel reads the low byte of).
struct TowerField {
  static constexpr uint32_t kOrder = 65535, kLevels = 11;
  static uint16_t mul(uint16_t a, uint16_t b) { return gf16().mul(a, b); }
  static uint16_t inv(uint16_t a) { return gf16().inv(a); }
  static uint32_t log(uint32_t x) { return gf16().log[x]; }
  static uint16_t exp(uint32_t i) { return gf16().exp[i]; }
  static uint16_t what(int level, size_t x) { return fft16_what(level, x); }
};

struct ByteField {
  static constexpr uint32_t kOrder = 255, kLevels = 8;
  static uint16_t mul(uint16_t a, uint16_t b) { return gf8().mul[a & 0xff][b & 0xff]; }
  static uint16_t inv(uint16_t a) { return gf8().inv(uint8_t(a)); }
  static uint32_t log(uint32_t x) { return gf8().log[x]; }
  static uint16_t exp(uint32_t i) { return gf8().exp[i]; }
  static uint16_t what(int level, size_t x) { return fft8_what(level, x); }
};

// Formal-derivative constants of the normalised basis: D_j is the linear
// coefficient of W_j / W_j(2^j), i.e. prod_{0<u<2^j} u / prod_{2^j<=u<2^(j+1)} u
// (the linear coefficient of the subspace polynomial W_j is the product of the
// nonzero points of its subspace).
template <class F>
const std::vector<uint16_t> &derivative_consts() {
  static std::vector<uint16_t> d;
  static std::once_flag once;
  std::call_once(once, [] {
    for (uint32_t j = 0; j < F::kLevels; ++j) {
      uint16_t num = 1, den = 1;
      for (uint32_t u = 1; u < (1u << j); ++u) num = F::mul(num, uint16_t(u));
      for (uint32_t u = 1u << j; u < (2u << j); ++u) den = F::mul(den, uint16_t(u));
      d.push_back(F::mul(num, F::inv(den)));
    }
  });
  return d;
}

template <class F>
Fft16Lowered decode_lowered(size_t n, const std::vector<uint32_t> &valid,
                            const std::vector<uint32_t> &outputs) {
  size_t N = 2;
  int levels = 1;
  while (N < n) {
    N <<= 1;
    ++levels;
  }
  Fft16Lowered L;
  L.regs = N;
  std::vector<uint8_t> known(N, 0);
  for (uint32_t r : valid) known[r] = 1;
  std::vector<uint32_t> erased;
  for (uint32_t r = 0; r < N; ++r)
    if (!known[r]) erased.push_back(r);
  // 1. unknown points start at zero (they hold Q = P * Lambda = 0 there)
  {
    std::vector<OpWords> z;
    for (uint32_t r : erased) z.push_back(encode_op({kOpZero, uint16_t(r), uint16_t(r), 0}));
    glue_step(L, z);
  }
  // log Lambda(x) = sum over e of log(x ^ e), mod the multiplicative order
  auto log_prod = [&](uint32_t x, bool skip_self) {
    uint64_t acc = 0;
    for (uint32_t e : erased) {
      if (skip_self && e == x) continue;
      acc += F::log(x ^ e);
    }
    return uint32_t(acc % F::kOrder);
  };
  // 2. known values times Lambda(x)
  {
    std::vector<uint32_t> lg(valid.size());
    parallel_rows(valid.size(), erased.size(), [&](size_t a, size_t b) {
      for (size_t i = a; i < b; ++i) lg[i] = log_prod(valid[i], false);
    });
    std::vector<OpWords> mul;
    for (size_t i = 0; i < valid.size(); ++i) {
      const uint16_t c = F::exp(lg[i]);
      if (c != 1) mul.push_back(encode_op({kOpMulCopy, uint16_t(valid[i]), uint16_t(valid[i]), c}));
    }
    glue_step(L, mul);
  }
  // 3. values on [0, N) to coefficients
  std::vector<uint16_t> consts;
  std::vector<uint8_t> skip;
  auto transform_consts = [&](bool inverse, const std::vector<uint8_t> *live) {
    consts.clear();
    skip.clear();
    for (int step = 0; step < levels; ++step) {
      const int level = inverse ? step : levels - 1 - step;
      const size_t stride = size_t(1) << level;
      for (size_t block = 0; block < N; block += 2 * stride) {
        consts.push_back(F::what(level, block));
        bool any = live == nullptr;
        for (size_t j = block; !any && j < block + 2 * stride; ++j) any = (*live)[j] != 0;
        skip.push_back(!any);
      }
    }
  };
  transform_consts(true, nullptr);
  emit_transform(L, 0, uint32_t(N), true, consts.data(), skip.data());
  // 4. coefficient t += sum over zero bits j of t of D_j * coefficient (t | 2^j)
  // (the added Q itself vanishes on E).  With s_t = prod_{j in t} D_j and
  // u_t = s_t q_t this is s_t Q'_t = sum_j u_(t | 2^j): scale by s_t, XOR the
  // derivative in place, scale back by 1 / s_t.  u_t reads only registers
  // with one more set bit, so updating t in order of increasing popcount
  // reads every source before it is overwritten, and each popcount layer is
  // one step of independent kOpDeriv ops (levels + 1 steps in all).
  {
    const auto &D = derivative_consts<F>();
    std::vector<uint16_t> sc(N, 1);
    for (size_t t = 1; t < N; ++t) {
      int j = 0;
      while (!((t >> j) & 1)) ++j;
      sc[t] = F::mul(sc[t & (t - 1)], D[j]);
    }
    std::vector<OpWords> pre, post;
    for (size_t t = 0; t < N; ++t) {
      if (sc[t] == 1) continue;
      pre.push_back(encode_op({kOpMulCopy, uint16_t(t), uint16_t(t), sc[t]}));
      post.push_back(encode_op({kOpMulCopy, uint16_t(t), uint16_t(t), F::inv(sc[t])}));
    }
    glue_step(L, pre);
    for (int pc = 0; pc < levels; ++pc) {
      std::vector<OpWords> layer;
      for (size_t t = 0; t < N; ++t)
        if (__builtin_popcountll(t) == pc)
          layer.push_back(encode_op({kOpDeriv, uint16_t(t), uint16_t(t), uint16_t(levels)}));
      glue_step(L, layer);
    }
    glue_step(L, post);
  }
  // 5. coefficients to values, pruned to the blocks that reach an output
  std::vector<uint8_t> live(N, 0);
  for (uint32_t e : outputs) live[e] = 1;
  transform_consts(false, &live);
  emit_transform(L, 0, uint32_t(N), false, consts.data(), skip.data());
  // 6. P(e) = Q'(e) / Lambda'(e)
  {
    std::vector<uint32_t> lg(outputs.size());
    parallel_rows(outputs.size(), erased.size(), [&](size_t a, size_t b) {
      for (size_t i = a; i < b; ++i) lg[i] = log_prod(outputs[i], true);
    });
    std::vector<OpWords> mul;
    for (size_t i = 0; i < outputs.size(); ++i) {
      const uint16_t c = F::exp((F::kOrder - lg[i]) % F::kOrder);
      if (c != 1) mul.push_back(encode_op({kOpMulCopy, uint16_t(outputs[i]), uint16_t(outputs[i]), c}));
    }
    glue_step(L, mul);
  }
  L.load = valid;
  L.store = outputs;
  return L;
}

}  // namespace

Fft16Lowered fft16_decode_lowered(size_t n, const std::vector<uint32_t> &valid,
                                  const std::vector<uint32_t> &outputs) {
  return decode_lowered<TowerField>(n, valid, outputs);
}

Fft16Lowered fft8_decode_lowered(size_t n, const std::vector<uint32_t> &valid,
                                 const std::vector<uint32_t> &outputs) {
  return decode_lowered<ByteField>(n, valid, outputs);
}

namespace {

// ---- device code ------------------------------------------------------------

__device__ __forceinline__ uint32_t bit_mask(uint32_t a, int b) {
  return uint32_t(int32_t(a << (31 - b)) >> 31);
}

// acc ^= a * v over GF(2^8) on 8 planes; a is a per-lane byte.
__device__ __forceinline__ void gf8_mul_acc(uint32_t a, const uint32_t *v, uint32_t *acc) {
  uint32_t p[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) p[j] = v[j];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint32_t mk = bit_mask(a, b);
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] ^= p[j] & mk;
    if (b < 7) {
      // p *= 2: bit j <- bit j-1, overflow bit 7 folds into bits 0,2,3,4 (0x1d)
      const uint32_t t = p[7];
      p[7] = p[6];
      p[6] = p[5];
      p[5] = p[4];
      p[4] = p[3] ^ t;
      p[3] = p[2] ^ t;
      p[2] = p[1] ^ t;
      p[1] = p[0];
      p[0] = t;
    }
  }
}

// acc ^= a * v for a wave-uniform byte a (scalar register): zero bits are
// skipped with scalar branches and the doubling chain stops at the top bit.
__device__ __forceinline__ void gf8_mul_acc_uniform(uint32_t a, const uint32_t *v,
                                                    uint32_t *acc) {
  uint32_t p[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) p[j] = v[j];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    if ((a >> b) == 0) break;
    if ((a >> b) & 1u) {
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] ^= p[j];
    }
    if (b < 7) {
      const uint32_t t = p[7];
      p[7] = p[6];
      p[6] = p[5];
      p[5] = p[4];
      p[4] = p[3] ^ t;
      p[3] = p[2] ^ t;
      p[2] = p[1] ^ t;
      p[1] = p[0];
      p[0] = t;
    }
  }
}

// tower_mul_acc for a wave-uniform constant word.
__device__ __forceinline__ void tower_mul_acc_uniform(uint32_t cw, const uint32_t (&x)[16],
                                                      uint32_t (&out)[16]) {
  const uint32_t a0 = cw & 0xffu, a1 = (cw >> 8) & 0xffu, a2 = (cw >> 16) & 0xffu;
  if (a2 == 0) {
    // subfield constant: c * x = (c_lo x_lo, c_lo x_hi)
    gf8_mul_acc_uniform(a0, x, out);
    gf8_mul_acc_uniform(a0, x + 8, out + 8);
    return;
  }
  uint32_t t0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  gf8_mul_acc_uniform(a0, x, t0);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    out[j] ^= t0[j];
    out[8 + j] ^= t0[j];
  }
  uint32_t s[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = x[j] ^ x[8 + j];
  gf8_mul_acc_uniform(a1, s, out + 8);
  gf8_mul_acc_uniform(a2, x + 8, out);
}

// out ^= c * x over the tower, c given as its Karatsuba word.
__device__ __forceinline__ void tower_mul_acc(uint32_t cw, const uint32_t (&x)[16],
                                              uint32_t (&out)[16]) {
  const uint32_t a0 = cw & 0xffu, a1 = (cw >> 8) & 0xffu, a2 = (cw >> 16) & 0xffu;
  uint32_t t0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  gf8_mul_acc(a0, x, t0);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    out[j] ^= t0[j];
    out[8 + j] ^= t0[j];
  }
  uint32_t s[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = x[j] ^ x[8 + j];
  gf8_mul_acc(a1, s, out + 8);
  if (a2) gf8_mul_acc(a2, x + 8, out);
}

// Register slot x (= register * G + group) holds P/4 quads of 16 bytes (P
// planes: 16 for tower symbols, 8 for bytes).  Tower: quad q sits at position
// (q + x/2) mod 4; bytes: (q + x/4) mod 2; either way lanes reading
// consecutive slots spread over all banks.
template <int P>
__device__ __forceinline__ uint32_t rf_quad(uint32_t x, int q) {
  return P == 16 ? x * 4 + ((q + (x >> 1)) & 3) : x * 2 + ((q + (x >> 2)) & 1);
}

template <int P>
__device__ __forceinline__ void rf_load(const uint4 *rf, uint32_t x, uint32_t (&v)[P]) {
#pragma unroll
  for (int q = 0; q < P / 4; ++q) {
    const uint4 t = rf[rf_quad<P>(x, q)];
    v[4 * q] = t.x;
    v[4 * q + 1] = t.y;
    v[4 * q + 2] = t.z;
    v[4 * q + 3] = t.w;
  }
}

template <int P>
__device__ __forceinline__ void rf_store(uint4 *rf, uint32_t x, const uint32_t (&v)[P]) {
#pragma unroll
  for (int q = 0; q < P / 4; ++q)
    rf[rf_quad<P>(x, q)] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

// out ^= c * x for a per-lane constant word: the tower's Karatsuba word
// (P = 16) or a byte (P = 8, the word's low byte).
template <int P>
__device__ __forceinline__ void field_mul_acc(uint32_t cw, const uint32_t (&x)[P],
                                              uint32_t (&out)[P]) {
  if constexpr (P == 16)
    tower_mul_acc(cw, x, out);
  else
    gf8_mul_acc(cw & 0xffu, x, out);
}

template <int P>
__device__ __forceinline__ void field_mul_acc_uniform(uint32_t cw, const uint32_t (&x)[P],
                                                      uint32_t (&out)[P]) {
  if constexpr (P == 16)
    tower_mul_acc_uniform(cw, x, out);
  else
    gf8_mul_acc_uniform(cw & 0xffu, x, out);
}

struct Fft16Args {
  const uint8_t *const *ptrs;
  const Fft16Step *steps;
  const uint32_t *consts;
  const uint2 *ops;
  const uint32_t *load;
  const uint32_t *store;
  uint64_t plane;  // offset of the high plane from a shard pointer
  uint64_t cols;   // symbol columns processed (from each pointer)
  uint32_t nload, nstore, nsteps, gshift;
  uint32_t groups;   // column groups per shard
  uint32_t cblocks;  // workgroups per stripe
  uint32_t total;    // workgroups (single-program launch); first workgroup (job table)
  uint32_t vec_ok;
  uint32_t *flag;    // verify: compare the stored registers with the shards, set on a mismatch
  // Indexed addressing (ptrs unused when idx is set): pointer slot i is shard
  // idx[i] of a stripe laid out at `base`, shards `stride` bytes apart.
  const uint32_t *idx;
  const uint8_t *base;
  uint64_t stride;
};

__device__ __forceinline__ const uint8_t *slot_ptr(const Fft16Args &a, const uint8_t *const *ptr,
                                                   uint32_t slot) {
  return a.idx ? a.base + uint64_t(a.idx[slot]) * a.stride : ptr[slot];
}

// XCD-aware order: consecutive workgroups of one XCD take consecutive
// column blocks, so the 32-byte segments of one 128-byte line meet in one L2.
__device__ __forceinline__ uint32_t xcd_order(uint32_t wg, uint32_t total) {
  return (total & 7u) == 0 ? (wg & 7u) * (total >> 3) + (wg >> 3) : wg;
}

template <int T, int P>
__device__ __forceinline__ void run_program(const Fft16Args &a, uint32_t g0,
                                            const uint8_t *const *ptr, uint4 *rf);

// PL = planes per register: 16 (tower symbols of split-plane shards) or 8
// (bytes of GF(2^8) shards: a program whose constants are all bytes).
template <int T, int PL>
__global__ __launch_bounds__(T) void k_fft16(Fft16Args a) {
  extern __shared__ uint4 rf[];
  const uint32_t wg = xcd_order(blockIdx.x, a.total);
  const uint32_t stripe = wg / a.cblocks;
  const uint32_t g0 = (wg - stripe * a.cblocks) << a.gshift;
  run_program<T, PL>(a, g0, a.ptrs + uint64_t(stripe) * (a.nload + a.nstore), rf);
}

// One stripe whose shard pointers travel in the kernel arguments (no device
// table to allocate and copy per call).
template <int T, int PL>
__global__ __launch_bounds__(T) void k_fft16_inline(Fft16Args a, Fft16Ptrs P) {
  extern __shared__ uint4 rf[];
  const uint32_t wg = xcd_order(blockIdx.x, a.total);
  run_program<T, PL>(a, wg << a.gshift, P.p, rf);
}

// Many programs (one per job, e.g. one decode pattern per stripe) in one
// launch: workgroup w runs job wg_job[w] on its column block w - first.
template <int T, int PL>
__global__ __launch_bounds__(T) void k_fft16_jobs(const Fft16Args *jobs, const uint32_t *wg_job,
                                                  uint32_t total) {
  extern __shared__ uint4 rf[];
  const uint32_t wg = xcd_order(blockIdx.x, total);
  const Fft16Args a = ld_const_val(jobs + ld_const(wg_job + wg));  // scalar loads
  run_program<T, PL>(a, (wg - a.total) << a.gshift, a.ptrs, rf);
}

template <int T, int P>
__device__ __forceinline__ void run_program(const Fft16Args &a, uint32_t g0,
                                            const uint8_t *const *ptr, uint4 *rf) {
  const uint32_t G = 1u << a.gshift;

  for (uint32_t q = threadIdx.x; q < (a.nload << a.gshift); q += T) {
    const uint32_t slot = q >> a.gshift, g = q & (G - 1);
    uint32_t v[P];
    const uint32_t cg = g0 + g;
    if (cg < a.groups) {
      const uint64_t col = uint64_t(cg) * kGroupCols;
      const uint64_t rem = a.cols - col;
      const bool full = a.vec_ok && rem >= kGroupCols;
      const uint8_t *s = slot_ptr(a, ptr, slot);
      uint32_t lo[8];
      load_planes(s + col, full, rem, lo);
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = lo[j];
      if constexpr (P == 16) {
        uint32_t hi[8];
        load_planes(s + a.plane + col, full, rem, hi);
#pragma unroll
        for (int j = 0; j < 8; ++j) v[8 + j] = hi[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < P; ++j) v[j] = 0;
    }
    rf_store<P>(rf, (a.load[slot] << a.gshift) + g, v);
  }
  __syncthreads();

  for (uint32_t si = 0; si < a.nsteps; ++si) {
    // step descriptors are wave-uniform: scalar loads, not a vector-memory
    // round trip at the head of every step
    const Fft16Step st = ld_const_val(a.steps + si);
    const uint32_t n = st.count << a.gshift;
    if (st.kind == 0) {
      const uint32_t sl = st.b, smask = (1u << sl) - 1;
      for (uint32_t q = threadIdx.x; q < n; q += T) {
        const uint32_t o = q >> a.gshift, g = q & (G - 1);
        const uint32_t block = o >> sl;
        const uint32_t cw = a.consts[st.c + block];
        if (cw & kConstSkip) continue;
        const uint32_t lo = st.a + (block << (sl + 1)) + (o & smask);
        const uint32_t hi = lo + (1u << sl);
        const uint32_t xl = (lo << a.gshift) + g, xh = (hi << a.gshift) + g;
        uint32_t L[P], H[P];
        rf_load<P>(rf, xl, L);
        rf_load<P>(rf, xh, H);
        if (st.d) {
#pragma unroll
          for (int j = 0; j < P; ++j) H[j] ^= L[j];
        }
        if (cw & kConstOne) {
#pragma unroll
          for (int j = 0; j < P; ++j) L[j] ^= H[j];
        } else if (!(cw & kConstZero)) {
          // One block spans the wave at strides >= 64 (one column group per
          // workgroup): then the constant sits in a scalar register.
          const uint32_t cw0 = __builtin_amdgcn_readfirstlane(cw);
          if (__builtin_amdgcn_ballot_w64(cw != cw0) == 0)
            field_mul_acc_uniform<P>(cw0, H, L);
          else
            field_mul_acc<P>(cw, H, L);
        }
        if (!st.d) {
#pragma unroll
          for (int j = 0; j < P; ++j) H[j] ^= L[j];
        }
        rf_store<P>(rf, xl, L);
        rf_store<P>(rf, xh, H);
      }
    } else {
      for (uint32_t q = threadIdx.x; q < n; q += T) {
        const uint32_t o = q >> a.gshift, g = q & (G - 1);
        const uint2 op = a.ops[st.a + o];
        const uint32_t kind = op.y & 0xffu;
        const uint32_t xd = ((op.x & 0xffffu) << a.gshift) + g;
        const uint32_t xs = ((op.x >> 16) << a.gshift) + g;
        uint32_t D[P], S[P];
        if (kind == kOpDeriv) {
          // dst ^= registers dst | 2^j for the clear bits j below the level count
          const uint32_t t = op.x & 0xffffu;
          rf_load<P>(rf, xd, D);
          for (uint32_t zb = ~t & ((1u << (op.y >> 8)) - 1u); zb; zb &= zb - 1u) {
            rf_load<P>(rf, ((t | (zb & (~zb + 1u))) << a.gshift) + g, S);
#pragma unroll
            for (int j = 0; j < P; ++j) D[j] ^= S[j];
          }
        } else if (kind == kOpZero) {
#pragma unroll
          for (int j = 0; j < P; ++j) D[j] = 0;
        } else {
          rf_load<P>(rf, xs, S);
          if (kind == kOpXor || kind == kOpMulXor) {
            rf_load<P>(rf, xd, D);
          } else {
#pragma unroll
            for (int j = 0; j < P; ++j) D[j] = 0;
          }
          if (kind == kOpXor || kind == kOpCopy) {
#pragma unroll
            for (int j = 0; j < P; ++j) D[j] ^= S[j];
          } else {
            field_mul_acc<P>(op.y >> 8, S, D);
          }
        }
        rf_store<P>(rf, xd, D);
      }
    }
    __syncthreads();
  }

  uint32_t bad = 0;
  for (uint32_t q = threadIdx.x; q < (a.nstore << a.gshift); q += T) {
    const uint32_t j = q >> a.gshift, g = q & (G - 1);
    const uint32_t cg = g0 + g;
    if (cg >= a.groups) continue;
    uint32_t v[P];
    rf_load<P>(rf, (a.store[j] << a.gshift) + g, v);
    const uint64_t col = uint64_t(cg) * kGroupCols;
    const uint64_t rem = a.cols - col;
    const bool full = a.vec_ok && rem >= kGroupCols;
    uint8_t *d = const_cast<uint8_t *>(slot_ptr(a, ptr, a.nload + j));
    if (a.flag) {
      // Compare in plane form: both sides are the same transpose of their
      // bytes, and columns past the end are zero on both sides (the program
      // is linear and its inputs read as zero there).
      uint32_t x[8];
      load_planes(d + col, full, rem, x);
#pragma unroll
      for (int i = 0; i < 8; ++i) bad |= x[i] ^ v[i];
      if constexpr (P == 16) {
        load_planes(d + a.plane + col, full, rem, x);
#pragma unroll
        for (int i = 0; i < 8; ++i) bad |= x[i] ^ v[8 + i];
      }
      continue;
    }
    uint32_t lo[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) lo[i] = v[i];
    store_planes(d + col, full, rem, lo);
    if constexpr (P == 16) {
      uint32_t hi[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) hi[i] = v[8 + i];
      store_planes(d + a.plane + col, full, rem, hi);
    }
  }
  if (bad) *a.flag = 1u;
}

}  // namespace

bool fft16_fits_device(const Fft16Program &p) {
  return fft16_lower(p).regs * kRegBytes <= kLdsMax;
}

hipError_t fft16_device_upload(const Fft16Lowered &L, Fft16DevicePtr *out) {
  if (L.regs * kRegBytes > kLdsMax) return hipErrorInvalidValue;
  auto d = Fft16DevicePtr(new Fft16Device);
  d->nload = L.load.size();
  d->nstore = L.store.size();
  d->regs = L.regs;
  d->nsteps = L.steps.size();
  d->max_count = L.max_count;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t parts[5] = {L.steps.size() * sizeof(Fft16Step), L.consts.size() * 4,
                           L.ops.size() * 4, L.load.size() * 4, L.store.size() * 4};
  const void *src[5] = {L.steps.data(), L.consts.data(), L.ops.data(), L.load.data(),
                        L.store.data()};
  size_t total = 256;
  for (size_t b : parts) total += al(b);
  hipError_t e = hipMalloc(&d->mem, total);
  if (e != hipSuccess) return e;
  std::vector<uint8_t> host(total, 0);
  uint8_t *dev[5];
  size_t off = 0;
  for (int i = 0; i < 5; ++i) {
    if (parts[i]) std::memcpy(host.data() + off, src[i], parts[i]);
    dev[i] = static_cast<uint8_t *>(d->mem) + off;
    off += al(parts[i]);
  }
  d->steps = reinterpret_cast<const Fft16Step *>(dev[0]);
  d->consts = reinterpret_cast<const uint32_t *>(dev[1]);
  d->ops = reinterpret_cast<const uint2 *>(dev[2]);
  d->load = reinterpret_cast<const uint32_t *>(dev[3]);
  d->store = reinterpret_cast<const uint32_t *>(dev[4]);
  e = hipMemcpy(d->mem, host.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  *out = std::move(d);
  return hipSuccess;
}

hipError_t fft16_device_new(const Fft16Program &p, Fft16DevicePtr *out) {
  return fft16_device_upload(fft16_lower(p), out);
}

namespace {
// Column groups per workgroup: as many as fit `target` bytes of LDS (up to 8
// and the shard's groups).  Tower programs aim at 64 KiB (two workgroups per
// CU); byte-plane programs at 32 KiB, which measured 5-14 % faster on
// (100,156) and (200,55) at 1-4 MiB than 64 KiB (16 KiB: up to 40 % slower).
uint32_t choose_gshift(const Fft16Device &d, uint64_t groups, uint32_t reg_bytes = kRegBytes,
                       size_t target = kLdsTarget) {
  uint32_t gs = 0;
  while ((1u << (gs + 1)) <= kMaxGroupsPerWg && (d.regs * reg_bytes << (gs + 1)) <= target &&
         (uint64_t(1) << (gs + 1)) <= groups)
    ++gs;
  return gs;
}
}  // namespace

namespace {
// Raises the dynamic LDS limit of every interpreter kernel once per device and
// host thread when a launch needs more than the default 64 KiB.
hipError_t raise_lds_limit(size_t lds) {
  if (lds <= 65536) return hipSuccess;
  static thread_local int raised = -1;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || raised == dev) return e;
  const void *kernels[18] = {reinterpret_cast<const void *>(&k_fft16<1024, 16>),
                             reinterpret_cast<const void *>(&k_fft16<512, 16>),
                             reinterpret_cast<const void *>(&k_fft16<256, 16>),
                             reinterpret_cast<const void *>(&k_fft16_inline<1024, 16>),
                             reinterpret_cast<const void *>(&k_fft16_inline<512, 16>),
                             reinterpret_cast<const void *>(&k_fft16_inline<256, 16>),
                             reinterpret_cast<const void *>(&k_fft16<1024, 8>),
                             reinterpret_cast<const void *>(&k_fft16<512, 8>),
                             reinterpret_cast<const void *>(&k_fft16<256, 8>),
                             reinterpret_cast<const void *>(&k_fft16_inline<1024, 8>),
                             reinterpret_cast<const void *>(&k_fft16_inline<512, 8>),
                             reinterpret_cast<const void *>(&k_fft16_inline<256, 8>),
                             reinterpret_cast<const void *>(&k_fft16_jobs<1024, 16>),
                             reinterpret_cast<const void *>(&k_fft16_jobs<512, 16>),
                             reinterpret_cast<const void *>(&k_fft16_jobs<256, 16>),
                             reinterpret_cast<const void *>(&k_fft16_jobs<1024, 8>),
                             reinterpret_cast<const void *>(&k_fft16_jobs<512, 8>),
                             reinterpret_cast<const void *>(&k_fft16_jobs<256, 8>)};
  for (const void *k : kernels)
    if ((e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsMax))) !=
        hipSuccess)
      return e;
  raised = dev;
  return hipSuccess;
}
// Workgroup size: 1024 threads when the register file takes more than half
// the LDS (one workgroup per CU, e.g. the 128 KiB decoders) and the widest
// step has that many items -- the CU then runs 16 waves instead of 8 -- else
// 512 or 256 by the widest step.
uint32_t choose_threads(size_t lds, uint64_t widest) {
  if (lds > kLdsMax / 2 && widest >= 1024) return 1024;
  return widest >= 512 ? 512 : 256;
}
}  // namespace

namespace {
template <int PL>
void launch_planes(uint32_t threads, const Fft16Args &a, const Fft16Ptrs *inline_ptrs, size_t lds,
                   hipStream_t stream) {
  const dim3 grid(a.total);
  if (inline_ptrs) {
    const Fft16Ptrs &P = *inline_ptrs;
    if (threads == 1024)
      hipLaunchKernelGGL((k_fft16_inline<1024, PL>), grid, dim3(1024), lds, stream, a, P);
    else if (threads == 512)
      hipLaunchKernelGGL((k_fft16_inline<512, PL>), grid, dim3(512), lds, stream, a, P);
    else
      hipLaunchKernelGGL((k_fft16_inline<256, PL>), grid, dim3(256), lds, stream, a, P);
  } else if (threads == 1024) {
    hipLaunchKernelGGL((k_fft16<1024, PL>), grid, dim3(1024), lds, stream, a);
  } else if (threads == 512) {
    hipLaunchKernelGGL((k_fft16<512, PL>), grid, dim3(512), lds, stream, a);
  } else {
    hipLaunchKernelGGL((k_fft16<256, PL>), grid, dim3(256), lds, stream, a);
  }
}
}  // namespace

uint32_t fft16_launch_gshift(const Fft16Device &d, uint64_t cols, bool byte_planes,
                             uint64_t lds_target) {
  const uint64_t groups = (cols + kGroupCols - 1) / kGroupCols;
  const uint32_t reg_bytes = byte_planes ? kRegBytes / 2 : kRegBytes;
  const size_t target =
      lds_target ? size_t(lds_target) : byte_planes ? kLdsTarget / 2 : kLdsTarget;
  return choose_gshift(d, groups, reg_bytes, std::min(target, kLdsMax));
}

hipError_t launch_fft16(const Fft16Device &d, const Fft16Call &call, hipStream_t stream,
                        const Fft16Ptrs *inline_ptrs, bool byte_planes) {
  if (call.plane == 0 || call.stripes == 0) return hipSuccess;
  if (inline_ptrs && (call.stripes != 1 || d.nload + call.nstore > kFftInlinePtrs))
    return hipErrorInvalidValue;
  const uint64_t cols = call.cols ? call.cols : call.plane;
  const uint64_t groups = (cols + kGroupCols - 1) / kGroupCols;
  const uint32_t reg_bytes = byte_planes ? kRegBytes / 2 : kRegBytes;
  const size_t target =
      call.lds_target ? size_t(call.lds_target) : byte_planes ? kLdsTarget / 2 : kLdsTarget;
  const uint32_t gs = choose_gshift(d, groups, reg_bytes, std::min(target, kLdsMax));
  const uint64_t cblocks = (groups + (1u << gs) - 1) >> gs;
  const uint64_t total = cblocks * call.stripes;
  if (groups > 0xffffffffu || total > 0x7fffffffu) return hipErrorInvalidValue;
  Fft16Args a{};
  if (call.nstore > d.nstore) return hipErrorInvalidValue;
  a.ptrs = call.ptrs;
  a.steps = d.steps;
  a.consts = d.consts;
  a.ops = d.ops;
  a.load = d.load;
  a.store = d.store;
  a.plane = call.plane;
  a.cols = cols;
  a.nload = uint32_t(d.nload);
  a.nstore = call.nstore;
  a.nsteps = uint32_t(d.nsteps);
  a.gshift = gs;
  a.groups = uint32_t(groups);
  a.cblocks = uint32_t(cblocks);
  a.total = uint32_t(total);
  // pointers carry any column offset; the high plane sits `plane` bytes on
  a.vec_ok = (call.vec_ok && (byte_planes || call.plane % 16 == 0)) ? 1u : 0u;
  a.flag = call.flag;
  const size_t lds = (d.regs * reg_bytes) << gs;
  const uint32_t threads = choose_threads(lds, uint64_t(d.max_count) << gs);
  hipError_t e = raise_lds_limit(lds);
  if (e != hipSuccess) return e;
  if (byte_planes)
    launch_planes<8>(threads, a, inline_ptrs, lds, stream);
  else
    launch_planes<16>(threads, a, inline_ptrs, lds, stream);
  return hipGetLastError();
}

hipError_t fft16_jobs_plan(const std::vector<Fft16JobSpec> &jobs, bool byte_planes,
                           Fft16JobsPlan *plan) {
  *plan = Fft16JobsPlan{};
  plan->byte_planes = byte_planes;
  const uint32_t reg_bytes = byte_planes ? kRegBytes / 2 : kRegBytes;
  const size_t target = byte_planes ? kLdsTarget / 2 : kLdsTarget;
  std::vector<Fft16Args> table;
  std::vector<uint32_t> wg_job;
  uint64_t widest = 0;
  for (const Fft16JobSpec &j : jobs) {
    const Fft16Device &d = *j.prog;
    if (j.plane == 0 || j.nstore == 0) continue;
    if (j.nstore > d.nstore) return hipErrorInvalidValue;
    const uint64_t groups = (j.plane + kGroupCols - 1) / kGroupCols;
    const uint32_t gs = choose_gshift(d, groups, reg_bytes, target);
    const uint64_t cblocks = (groups + (1u << gs) - 1) >> gs;
    if (groups > 0xffffffffu || wg_job.size() + cblocks > 0x7fffffffu) return hipErrorInvalidValue;
    Fft16Args a{};
    a.ptrs = j.ptrs;
    a.steps = d.steps;
    a.consts = d.consts;
    a.ops = d.ops;
    a.load = d.load;
    a.store = d.store;
    a.plane = j.plane;
    a.cols = j.plane;
    a.nload = uint32_t(d.nload);
    a.nstore = j.nstore;
    a.nsteps = uint32_t(d.nsteps);
    a.gshift = gs;
    a.groups = uint32_t(groups);
    a.cblocks = uint32_t(cblocks);
    a.total = uint32_t(wg_job.size());
    a.vec_ok = (j.vec_ok && (byte_planes || j.plane % 16 == 0)) ? 1u : 0u;
    a.idx = j.idx;
    a.base = j.base;
    a.stride = j.stride;
    wg_job.insert(wg_job.end(), cblocks, uint32_t(table.size()));
    table.push_back(a);
    plan->lds = std::max(plan->lds, (d.regs * reg_bytes) << gs);
    widest = std::max<uint64_t>(widest, uint64_t(d.max_count) << gs);
  }
  if (table.empty()) return hipSuccess;
  const size_t tb = table.size() * sizeof(Fft16Args), mb = wg_job.size() * sizeof(uint32_t);
  plan->map_offset = (tb + 255) & ~size_t(255);
  plan->blob.assign(plan->map_offset + mb, 0);
  std::memcpy(plan->blob.data(), table.data(), tb);
  std::memcpy(plan->blob.data() + plan->map_offset, wg_job.data(), mb);
  plan->total = uint32_t(wg_job.size());
  plan->threads = choose_threads(plan->lds, widest);
  return hipSuccess;
}

hipError_t fft16_jobs_launch(const Fft16JobsPlan &plan, const void *dev_blob, hipStream_t stream) {
  if (plan.total == 0) return hipSuccess;
  hipError_t e = raise_lds_limit(plan.lds);
  if (e != hipSuccess) return e;
  const auto *jt = static_cast<const Fft16Args *>(dev_blob);
  const auto *wj =
      reinterpret_cast<const uint32_t *>(static_cast<const uint8_t *>(dev_blob) + plan.map_offset);
  const dim3 grid(plan.total);
  const uint32_t total = plan.total;
  const size_t lds = plan.lds;
#define TRS_JOBS(T, PL) \
  hipLaunchKernelGGL((k_fft16_jobs<T, PL>), grid, dim3(T), lds, stream, jt, wj, total)
  if (plan.byte_planes) {
    if (plan.threads == 1024) TRS_JOBS(1024, 8);
    else if (plan.threads == 512) TRS_JOBS(512, 8);
    else TRS_JOBS(256, 8);
  } else {
    if (plan.threads == 1024) TRS_JOBS(1024, 16);
    else if (plan.threads == 512) TRS_JOBS(512, 16);
    else TRS_JOBS(256, 16);
  }
#undef TRS_JOBS
  return hipGetLastError();
}

namespace {
// Per-thread pinned staging buffers, kept for the life of the process
// (freeing pinned memory from a thread-exit destructor could run after the
// HIP runtime has shut down).
struct StagingRing {
  static constexpr int N = 4;
  void *buf[N] = {};
  size_t cap[N] = {};
  hipEvent_t done[N] = {};
  bool used[N] = {};
  int next = 0;
};
// One ring per host thread, made on first use.  Threads of the process's own
// lifetime keep theirs (freeing pinned memory during exit is unsafe); a
// thread that ends before the process releases its ring first.
thread_local StagingRing *t_ring = nullptr;
}  // namespace

void staged_upload_release() {
  StagingRing *R = t_ring;
  if (!R) return;
  t_ring = nullptr;
  for (int j = 0; j < StagingRing::N; ++j) {
    if (R->done[j]) {
      if (R->used[j]) (void)hipEventSynchronize(R->done[j]);
      (void)hipEventDestroy(R->done[j]);
    }
    if (R->buf[j]) (void)hipHostFree(R->buf[j]);
  }
  delete R;
}

hipError_t staged_upload(void *dst, const void *src, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return hipSuccess;
  if (!t_ring) t_ring = new StagingRing();
  StagingRing &R = *t_ring;
  const int i = R.next;
  R.next = (R.next + 1) % StagingRing::N;
  hipError_t e;
  if (!R.done[i] && (e = hipEventCreateWithFlags(&R.done[i], hipEventDisableTiming)) != hipSuccess)
    return e;
  if (R.used[i] && (e = hipEventSynchronize(R.done[i])) != hipSuccess) return e;
  if (bytes > R.cap[i]) {
    // Grow every idle slot at once (at least 1 MiB), so the pinned
    // allocations happen on one call rather than one per slot.
    const size_t want = std::max<size_t>(bytes, size_t(1) << 20);
    for (int j = 0; j < StagingRing::N; ++j) {
      if (R.cap[j] >= want || (j != i && R.used[j] && hipEventQuery(R.done[j]) != hipSuccess))
        continue;
      if (R.buf[j]) (void)hipHostFree(R.buf[j]);
      R.buf[j] = nullptr;
      R.cap[j] = 0;
      if ((e = hipHostMalloc(&R.buf[j], want, hipHostMallocDefault)) != hipSuccess) return e;
      R.cap[j] = want;
    }
  }
  std::memcpy(R.buf[i], src, bytes);
  if ((e = hipMemcpyAsync(dst, R.buf[i], bytes, hipMemcpyHostToDevice, stream)) != hipSuccess)
    return e;
  R.used[i] = true;
  return hipEventRecord(R.done[i], stream);
}

hipError_t noalloc_scratch(void **out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {  // once per device, with that device current
    void *p = nullptr;
    if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
    addr[dev] = p;
  }
  *out = addr[dev];
  return hipSuccess;
}

hipError_t launch_fft16_jobs(const std::vector<Fft16JobSpec> &jobs, hipStream_t stream,
                             bool byte_planes) {
  Fft16JobsPlan plan;
  hipError_t e = fft16_jobs_plan(jobs, byte_planes, &plan);
  if (e != hipSuccess || plan.total == 0) return e;
  void *mem = nullptr;
  if ((e = hipMallocAsync(&mem, plan.blob.size(), stream)) != hipSuccess) return e;
  e = staged_upload(mem, plan.blob.data(), plan.blob.size(), stream);
  if (e == hipSuccess) e = fft16_jobs_launch(plan, mem, stream);
  const hipError_t f = hipFreeAsync(mem, stream);
  return e != hipSuccess ? e : f;
}

}  // namespace trs
