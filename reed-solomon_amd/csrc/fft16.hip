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
#include <numeric>
#include <vector>

#include "bitslice.hpp"
#include "fft16.hpp"
#include "field.hpp"

namespace trs {

struct Fft16Device {
  size_t k = 0, m = 0, regs = 0;
  size_t nsteps = 0, max_count = 0;
  void *mem = nullptr;  // one allocation: steps | consts | ops | parity regs
  const Fft16Step *steps = nullptr;
  const uint32_t *consts = nullptr;
  const uint2 *ops = nullptr;
  const uint32_t *parity = nullptr;
};

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
  return {uint32_t(op.dst) | (uint32_t(op.src) << 16), uint32_t(op.kind) | (w << 8)};
}

}  // namespace

Fft16Lowered fft16_lower(const Fft16Program &program) {
  Fft16Program p = program;
  coalesce_and_compact(p);
  Fft16Lowered L;
  L.regs = p.register_count;
  using OpWords = std::pair<uint32_t, uint32_t>;
  auto glue_step = [&](const std::vector<OpWords> &ops) {
    if (ops.empty()) return;
    L.steps.push_back({1, uint32_t(ops.size()), uint32_t(L.ops.size() / 2), 0, 0, 0, 0, 0});
    for (const auto &o : ops) {
      L.ops.push_back(o.first);
      L.ops.push_back(o.second);
    }
    L.max_count = std::max(L.max_count, ops.size());
  };
  {
    std::vector<OpWords> z;
    for (uint16_t r : p.zero_regs) z.push_back(encode_op({kOpZero, r, r, 0}));
    glue_step(z);
  }
  std::vector<int> lw(L.regs), lr(L.regs);
  for (const auto &st : p.stages) {
    if (st.transform) {
      int levels = 0;
      while ((size_t(1) << levels) < st.size) ++levels;
      uint32_t coff = st.consts_start;
      for (int step = 0; step < levels; ++step) {
        const int level = st.inverse ? step : levels - 1 - step;
        const uint32_t nblocks = st.size >> (level + 1);
        bool any = false;
        for (uint32_t b = 0; b < nblocks; ++b) any |= !p.consts_skip[coff + b];
        if (any) {
          const uint32_t c0 = uint32_t(L.consts.size());
          for (uint32_t b = 0; b < nblocks; ++b)
            L.consts.push_back(kara_word(p.consts[coff + b]) |
                               (p.consts_skip[coff + b] ? kConstSkip : 0u));
          L.steps.push_back({0, st.size / 2, st.start, uint32_t(level), c0,
                             st.inverse ? 1u : 0u, 0, 0});
          L.max_count = std::max<size_t>(L.max_count, st.size / 2);
        }
        coff += nblocks;
      }
      continue;
    }
    // Schedule the run into levels of mutually independent ops (no op in a
    // level reads or writes a register another op of the level writes).
    std::fill(lw.begin(), lw.end(), -1);
    std::fill(lr.begin(), lr.end(), -1);
    std::vector<std::vector<OpWords>> levels;
    for (size_t j = st.glue_start; j < size_t(st.glue_start) + st.glue_len; ++j) {
      const Fft16Op &op = p.glue[j];
      const bool reads_dst = op.kind == kOpXor || op.kind == kOpMulXor;
      const int lvl = std::max({lw[op.src] + 1, lw[op.dst] + 1, lr[op.dst] + 1});
      if (size_t(lvl) >= levels.size()) levels.resize(lvl + 1);
      levels[lvl].push_back(encode_op(op));
      lw[op.dst] = std::max(lw[op.dst], lvl);
      lr[op.src] = std::max(lr[op.src], lvl);
      if (reads_dst) lr[op.dst] = std::max(lr[op.dst], lvl);
    }
    for (const auto &lv : levels) glue_step(lv);
  }
  for (uint16_t r : p.parity_regs) L.parity.push_back(r);
  return L;
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

// Register slot x (= register * G + group) holds 4 quads of 16 bytes; quad q
// sits at position (q + x/2) mod 4 so that lanes reading consecutive slots
// spread over all banks.
__device__ __forceinline__ void rf_load(const uint4 *rf, uint32_t x, uint32_t (&v)[16]) {
  const uint32_t sw = x >> 1;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 t = rf[x * 4 + ((q + sw) & 3)];
    v[4 * q] = t.x;
    v[4 * q + 1] = t.y;
    v[4 * q + 2] = t.z;
    v[4 * q + 3] = t.w;
  }
}

__device__ __forceinline__ void rf_store(uint4 *rf, uint32_t x, const uint32_t (&v)[16]) {
  const uint32_t sw = x >> 1;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    rf[x * 4 + ((q + sw) & 3)] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

struct Fft16Args {
  const uint8_t *const *ptrs;
  const Fft16Step *steps;
  const uint32_t *consts;
  const uint2 *ops;
  const uint32_t *parity;
  uint64_t plane;
  uint32_t k, m, nsteps, gshift;
  uint32_t groups;   // column groups per shard
  uint32_t cblocks;  // workgroups per stripe
  uint32_t total;    // workgroups
  uint32_t vec_ok;
};

template <int T>
__global__ __launch_bounds__(T) void k_fft16(Fft16Args a) {
  extern __shared__ uint4 rf[];
  const uint32_t G = 1u << a.gshift;
  // XCD-aware order: consecutive workgroups of one XCD take consecutive
  // column blocks, so the 32-byte segments of one 128-byte line meet in one L2.
  uint32_t wg = blockIdx.x;
  if ((a.total & 7u) == 0) wg = (wg & 7u) * (a.total >> 3) + (wg >> 3);
  const uint32_t stripe = wg / a.cblocks;
  const uint32_t g0 = (wg - stripe * a.cblocks) << a.gshift;
  const uint8_t *const *ptr = a.ptrs + uint64_t(stripe) * (a.k + a.m);

  for (uint32_t q = threadIdx.x; q < (a.k << a.gshift); q += T) {
    const uint32_t r = q >> a.gshift, g = q & (G - 1);
    uint32_t v[16];
    const uint32_t cg = g0 + g;
    if (cg < a.groups) {
      const uint64_t col = uint64_t(cg) * kGroupCols;
      const uint64_t rem = a.plane - col;
      const bool full = a.vec_ok && rem >= kGroupCols;
      const uint8_t *s = ptr[r];
      uint32_t lo[8], hi[8];
      load_planes(s + col, full, rem, lo);
      load_planes(s + a.plane + col, full, rem, hi);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v[j] = lo[j];
        v[8 + j] = hi[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = 0;
    }
    rf_store(rf, q, v);
  }
  __syncthreads();

  for (uint32_t si = 0; si < a.nsteps; ++si) {
    const Fft16Step st = a.steps[si];
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
        uint32_t L[16], H[16];
        rf_load(rf, xl, L);
        rf_load(rf, xh, H);
        if (st.d) {
#pragma unroll
          for (int j = 0; j < 16; ++j) H[j] ^= L[j];
        }
        if (cw & kConstOne) {
#pragma unroll
          for (int j = 0; j < 16; ++j) L[j] ^= H[j];
        } else if (!(cw & kConstZero)) {
          tower_mul_acc(cw, H, L);
        }
        if (!st.d) {
#pragma unroll
          for (int j = 0; j < 16; ++j) H[j] ^= L[j];
        }
        rf_store(rf, xl, L);
        rf_store(rf, xh, H);
      }
    } else {
      for (uint32_t q = threadIdx.x; q < n; q += T) {
        const uint32_t o = q >> a.gshift, g = q & (G - 1);
        const uint2 op = a.ops[st.a + o];
        const uint32_t kind = op.y & 0xffu;
        const uint32_t xd = ((op.x & 0xffffu) << a.gshift) + g;
        const uint32_t xs = ((op.x >> 16) << a.gshift) + g;
        uint32_t D[16], S[16];
        if (kind == kOpZero) {
#pragma unroll
          for (int j = 0; j < 16; ++j) D[j] = 0;
        } else {
          rf_load(rf, xs, S);
          if (kind == kOpXor || kind == kOpMulXor) {
            rf_load(rf, xd, D);
          } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) D[j] = 0;
          }
          if (kind == kOpXor || kind == kOpCopy) {
#pragma unroll
            for (int j = 0; j < 16; ++j) D[j] ^= S[j];
          } else {
            tower_mul_acc(op.y >> 8, S, D);
          }
        }
        rf_store(rf, xd, D);
      }
    }
    __syncthreads();
  }

  for (uint32_t q = threadIdx.x; q < (a.m << a.gshift); q += T) {
    const uint32_t j = q >> a.gshift, g = q & (G - 1);
    const uint32_t cg = g0 + g;
    if (cg >= a.groups) continue;
    uint32_t v[16];
    rf_load(rf, (a.parity[j] << a.gshift) + g, v);
    const uint64_t col = uint64_t(cg) * kGroupCols;
    const uint64_t rem = a.plane - col;
    const bool full = a.vec_ok && rem >= kGroupCols;
    uint8_t *d = const_cast<uint8_t *>(ptr[a.k + j]);
    uint32_t lo[8], hi[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      lo[i] = v[i];
      hi[i] = v[8 + i];
    }
    store_planes(d + col, full, rem, lo);
    store_planes(d + a.plane + col, full, rem, hi);
  }
}

}  // namespace

hipError_t fft16_device_new(const Fft16Program &p, Fft16DevicePtr *out) {
  const Fft16Lowered L = fft16_lower(p);
  if (L.regs * kRegBytes > kLdsMax) return hipErrorInvalidValue;
  auto d = Fft16DevicePtr(new Fft16Device);
  d->k = p.k;
  d->m = p.m;
  d->regs = L.regs;
  d->nsteps = L.steps.size();
  d->max_count = L.max_count;
  const size_t s_bytes = L.steps.size() * sizeof(Fft16Step);
  const size_t c_bytes = L.consts.size() * sizeof(uint32_t);
  const size_t o_bytes = L.ops.size() * sizeof(uint32_t);
  const size_t p_bytes = L.parity.size() * sizeof(uint32_t);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t total = al(s_bytes) + al(c_bytes) + al(o_bytes) + al(p_bytes) + 256;
  hipError_t e = hipMalloc(&d->mem, total);
  if (e != hipSuccess) return e;
  std::vector<uint8_t> host(total, 0);
  size_t off = 0;
  auto put = [&](const void *src, size_t n) {
    std::memcpy(host.data() + off, src, n);
    const size_t at = off;
    off += al(n);
    return static_cast<uint8_t *>(d->mem) + at;
  };
  d->steps = reinterpret_cast<const Fft16Step *>(put(L.steps.data(), s_bytes));
  d->consts = reinterpret_cast<const uint32_t *>(put(L.consts.data(), c_bytes));
  d->ops = reinterpret_cast<const uint2 *>(put(L.ops.data(), o_bytes));
  d->parity = reinterpret_cast<const uint32_t *>(put(L.parity.data(), p_bytes));
  e = hipMemcpy(d->mem, host.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  *out = std::move(d);
  return hipSuccess;
}

namespace {
uint32_t choose_gshift(const Fft16Device &d, uint64_t groups) {
  uint32_t gs = 0;
  while ((1u << (gs + 1)) <= kMaxGroupsPerWg && (d.regs * kRegBytes << (gs + 1)) <= kLdsTarget &&
         (uint64_t(1) << (gs + 1)) <= groups)
    ++gs;
  return gs;
}
}  // namespace

hipError_t launch_fft16(const Fft16Device &d, const Fft16Call &call, hipStream_t stream) {
  if (call.plane == 0 || call.stripes == 0) return hipSuccess;
  const uint64_t groups = (call.plane + kGroupCols - 1) / kGroupCols;
  const uint32_t gs = choose_gshift(d, groups);
  const uint64_t cblocks = (groups + (1u << gs) - 1) >> gs;
  const uint64_t total = cblocks * call.stripes;
  if (groups > 0xffffffffu || total > 0x7fffffffu) return hipErrorInvalidValue;
  Fft16Args a{};
  a.ptrs = call.ptrs;
  a.steps = d.steps;
  a.consts = d.consts;
  a.ops = d.ops;
  a.parity = d.parity;
  a.plane = call.plane;
  a.k = uint32_t(d.k);
  a.m = uint32_t(d.m);
  a.nsteps = uint32_t(d.nsteps);
  a.gshift = gs;
  a.groups = uint32_t(groups);
  a.cblocks = uint32_t(cblocks);
  a.total = uint32_t(total);
  a.vec_ok = (call.vec_ok && call.plane % 16 == 0) ? 1u : 0u;
  const size_t lds = (d.regs * kRegBytes) << gs;
  const bool wide = (d.max_count << gs) >= 512;
  if (lds > 65536) {
    // Raise the dynamic LDS limit once per kernel instance and device.
    static thread_local int raised = -1;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (raised != dev) {
      e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fft16<512>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsMax));
      if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fft16<256>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsMax));
      if (e != hipSuccess) return e;
      raised = dev;
    }
  }
  if (wide)
    hipLaunchKernelGGL(k_fft16<512>, dim3(uint32_t(total)), dim3(512), lds, stream, a);
  else
    hipLaunchKernelGGL(k_fft16<256>, dim3(uint32_t(total)), dim3(256), lds, stream, a);
  return hipGetLastError();
}

}  // namespace trs
