// Staged additive-FFT program builders (host).  The tower builder restates
// src/fft16.rs:42-621: WHAT tables, the interpolation / folded-chunk builder,
// the two base strategies, and the three-pass lowering (liveness, zero
// normalisation, per-block skip flags).  The GF(2^8) builder restates
// src/fft.rs:28-79 (WHAT) and :393-592 (build_staged_program, lower_staged)
// on the same builder with the byte field's tables.  The device lowering
// (fft16_device_new) is separate and lives with the kernel in fft16.hip.
#include <algorithm>
#include <mutex>

#include "fft16.hpp"
#include "field.hpp"

namespace trs {
namespace {

constexpr int kLevels = 11;  // transforms up to 2^11 points (src/fft16.rs:30-34)

// WHAT[j][x] = W_j(x) / W_j(2^j), W_0(x) = x, W_{j+1}(x) = W_j(x) (W_j(x) + W_j(2^j))
// (src/fft16.rs:42-67).
const std::vector<uint16_t> &what_table() {
  static std::vector<uint16_t> t;
  static std::once_flag once;
  std::call_once(once, [] {
    const Gf16 &f = gf16();
    t.assign(size_t(kLevels) << 16, 0);
    std::vector<uint16_t> w(65536), nxt(65536);
    for (uint32_t x = 0; x < 65536; ++x) w[x] = uint16_t(x);
    for (int j = 0; j < kLevels; ++j) {
      const uint16_t piv = w[size_t(1) << j];
      const uint16_t inv = f.inv(piv);
      uint16_t *row = t.data() + (size_t(j) << 16);
      for (uint32_t x = 0; x < 65536; ++x) row[x] = f.mul(w[x], inv);
      if (j + 1 < kLevels) {
        for (uint32_t x = 0; x < 65536; ++x) nxt[x] = f.mul(w[x], uint16_t(w[x] ^ piv));
        w.swap(nxt);
      }
    }
  });
  return t;
}

inline uint16_t what(int level, size_t x) { return what_table()[(size_t(level) << 16) | x]; }

}  // namespace

uint16_t fft16_what(int level, size_t x) { return what(level, x); }

namespace {

// The byte field's WHAT[8][256], same recurrence over GF(2^8)
// (src/fft.rs:28-79).
const std::vector<uint8_t> &what8_table() {
  static std::vector<uint8_t> t;
  static std::once_flag once;
  std::call_once(once, [] {
    const Gf8 &f = gf8();
    t.assign(size_t(8) << 8, 0);
    std::vector<uint8_t> w(256), nxt(256);
    for (uint32_t x = 0; x < 256; ++x) w[x] = uint8_t(x);
    for (int j = 0; j < 8; ++j) {
      const uint8_t piv = w[size_t(1) << j];
      const uint8_t inv = f.inv(piv);
      for (uint32_t x = 0; x < 256; ++x) t[(size_t(j) << 8) | x] = f.mul[w[x]][inv];
      for (uint32_t x = 0; x < 256; ++x) nxt[x] = f.mul[w[x]][w[x] ^ piv];
      w.swap(nxt);
    }
  });
  return t;
}

}  // namespace

uint16_t fft8_what(int level, size_t x) { return what8_table()[(size_t(level) << 8) | x]; }

namespace {

// The field a program is built over: its WHAT rows and product.
struct FieldOps {
  uint16_t (*what)(int level, size_t x);
  uint16_t (*mul)(uint16_t a, uint16_t b);
};

const FieldOps kTower{[](int level, size_t x) { return what(level, x); },
                      [](uint16_t a, uint16_t b) { return gf16().mul(a, b); }};
const FieldOps kByte{
    [](int level, size_t x) { return uint16_t(what8_table()[(size_t(level) << 8) | x]); },
    [](uint16_t a, uint16_t b) { return uint16_t(gf8().mul[a & 0xff][b & 0xff]); }};

inline int log2_exact(size_t n) {
  int l = 0;
  while ((size_t(1) << l) < n) ++l;
  return l;
}

inline size_t next_pow2(size_t n) {
  size_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

struct Item {
  bool transform;
  bool inverse;
  uint16_t start, size;
  size_t beta;
  Fft16Op op;  // glue: kOpXor or kOpMulXor only
};

struct Builder {
  const FieldOps *f = &kTower;
  // Folded chunks end in fft_with_zero_tail (tower, src/fft16.rs:341) or a
  // plain forward transform (byte field, src/fft.rs:449).
  bool zero_tail = true;
  std::vector<Item> items;
  size_t next = 0;

  std::vector<uint16_t> alloc(size_t n) {
    std::vector<uint16_t> r(n);
    for (size_t i = 0; i < n; ++i) r[i] = uint16_t(next + i);
    next += n;
    return r;
  }
  void transform(bool inverse, const uint16_t *regs, size_t n, size_t beta) {
    if (n > 1) items.push_back({true, inverse, regs[0], uint16_t(n), beta, {}});
  }
  void glue(uint8_t kind, uint16_t dst, uint16_t src, uint16_t c) {
    items.push_back({false, false, 0, 0, 0, {kind, dst, src, c}});
  }
  // dst ^= c * src with the c == 1 / c == 0 cases of the builder.
  void mul_xor(uint16_t dst, uint16_t src, uint16_t c) {
    if (c == 1)
      glue(kOpXor, dst, src, 0);
    else if (c)
      glue(kOpMulXor, dst, src, c);
  }

  // src/fft16.rs:225-265
  void interpolate(const uint16_t *regs, size_t n, size_t beta, size_t k) {
    if (k == n) {
      transform(true, regs, n, beta);
      return;
    }
    const size_t half = n / 2;
    if (k <= half) {
      interpolate(regs, half, beta, k);
      return;
    }
    transform(true, regs, half, beta);
    const uint16_t c = f->what(log2_exact(n) - 1, beta);
    const std::vector<uint16_t> scratch = alloc(half);
    for (size_t j = 0; j < half; ++j) glue(kOpXor, scratch[j], regs[j], 0);
    transform(false, scratch.data(), half, beta ^ half);
    for (size_t i = 0; i < k - half; ++i) glue(kOpXor, regs[half + i], scratch[i], 0);
    interpolate(regs + half, half, beta ^ half, k - half);
    for (size_t j = 0; j < k - half; ++j) mul_xor(regs[j], regs[half + j], c);
  }

  // src/fft16.rs:272-293
  void fft_with_zero_tail(const uint16_t *regs, size_t size, size_t beta, size_t defined) {
    if (defined >= size || size == 1) {
      transform(false, regs, size, beta);
      return;
    }
    const size_t half = size / 2;
    const uint16_t c = f->what(log2_exact(size) - 1, beta);
    for (size_t i = 0; i < half; ++i) {
      mul_xor(regs[i], regs[half + i], c);
      glue(kOpXor, regs[half + i], regs[i], 0);
    }
    transform(false, regs, half, beta);
    transform(false, regs + half, half, beta ^ half);
  }

  // src/fft16.rs:300-343
  std::vector<uint16_t> folded_chunk(const std::vector<uint16_t> &base, size_t k,
                                     size_t position, size_t size) {
    const size_t transform_n = base.size();
    const int tlevels = log2_exact(transform_n);
    const std::vector<uint16_t> chunk = alloc(size);
    const int shift = log2_exact(size);
    for (size_t j = 0; j < size; ++j) {
      for (size_t h = 0; h < transform_n / size; ++h) {
        const size_t index = h * size + j;
        if (index >= k) continue;
        uint16_t factor = 1;
        for (int level = shift; level < tlevels; ++level)
          if (index & (size_t(1) << level)) factor = f->mul(factor, f->what(level, position));
        if (factor == 0) continue;
        if (factor == 1)
          glue(kOpXor, chunk[j], base[index], 0);
        else
          glue(kOpMulXor, chunk[j], base[index], factor);
      }
    }
    if (zero_tail)
      fft_with_zero_tail(chunk.data(), size, position, std::min(k, size));
    else
      transform(false, chunk.data(), size, position);
    return chunk;
  }
};

// src/fft16.rs:438-621.  Without `liveness` every item is kept and every
// block counts as live, which is src/fft.rs:472-592 (same multiply count).
Fft16Program lower(const Builder &b, const std::vector<uint16_t> &parity, size_t k, size_t m,
                   bool liveness = true) {
  const FieldOps &f = *b.f;
  const size_t nreg = b.next;
  std::vector<uint8_t> live(nreg, liveness ? 0 : 1), keep(b.items.size(), liveness ? 0 : 1);
  std::vector<std::vector<uint8_t>> live_out(b.items.size());
  for (uint16_t r : parity) live[r] = 1;
  for (size_t i = b.items.size(); liveness && i-- > 0;) {
    const Item &it = b.items[i];
    if (!it.transform) {
      if (live[it.op.dst]) {
        keep[i] = 1;
        live[it.op.src] = 1;
      }
    } else {
      live_out[i].assign(live.begin() + it.start, live.begin() + it.start + it.size);
      std::fill(live.begin() + it.start, live.begin() + it.start + it.size, 1);
    }
  }
  std::vector<uint8_t> is_zero(nreg, 1), needs_zero(nreg, 0);
  std::fill(is_zero.begin(), is_zero.begin() + std::min(k, nreg), 0);
  Fft16Program p;
  p.k = k;
  p.m = m;
  p.register_count = nreg;
  p.parity_regs = parity;
  long open = -1;
  auto close_glue = [&] {
    if (open >= 0) {
      p.stages.push_back({false, false, 0, 0, 0, uint32_t(open), uint32_t(p.glue.size() - open)});
      open = -1;
    }
  };
  for (size_t i = 0; i < b.items.size(); ++i) {
    const Item &it = b.items[i];
    if (!it.transform) {
      if (!keep[i]) continue;
      Fft16Op op = it.op;
      if (is_zero[op.src]) continue;
      if (is_zero[op.dst]) {
        is_zero[op.dst] = 0;
        op.kind = op.kind == kOpMulXor ? kOpMulCopy : kOpCopy;
      }
      if (op.kind == kOpMulXor || op.kind == kOpMulCopy) ++p.mult_count;
      if (open < 0) open = long(p.glue.size());
      p.glue.push_back(op);
    } else {
      close_glue();
      for (size_t r = it.start; r < size_t(it.start) + it.size; ++r) {
        if (is_zero[r]) needs_zero[r] = 1;
        is_zero[r] = 0;
      }
      const auto &alive = live_out[i];
      const uint32_t cstart = uint32_t(p.consts.size());
      const int levels = log2_exact(it.size);
      for (int step = 0; step < levels; ++step) {
        const int level = it.inverse ? step : levels - 1 - step;
        const size_t stride = size_t(1) << level;
        for (size_t block = 0; block < it.size; block += 2 * stride) {
          const uint16_t c = f.what(level, it.beta ^ block);
          bool any = !liveness;
          for (size_t j = block; liveness && j < block + 2 * stride; ++j) any |= alive[j] != 0;
          p.consts.push_back(c);
          p.consts_skip.push_back(!any);
          if (c && any) p.mult_count += stride;
        }
      }
      p.stages.push_back({true, it.inverse, it.start, it.size, cstart, 0, 0});
    }
  }
  close_glue();
  for (size_t r = 0; r < nreg; ++r)
    if (needs_zero[r]) p.zero_regs.push_back(uint16_t(r));
  return p;
}

enum class Strategy { kInPlace, kWindow };

// src/fft16.rs:378-436
Fft16Program build_with(size_t k, size_t m, Strategy strategy) {
  const size_t n = k + m;
  const size_t transform_n = next_pow2(k);
  Builder b;
  const std::vector<uint16_t> base = b.alloc(transform_n);
  b.interpolate(base.data(), transform_n, 0, k);
  std::vector<uint16_t> parity(m, 0);
  const size_t base_end = std::min(n, transform_n);
  auto place = [&](const std::vector<uint16_t> &chunk, size_t position) {
    for (size_t off = 0; off < chunk.size(); ++off) {
      const size_t point = position + off;
      if (point >= k && point < n) parity[point - k] = chunk[off];
    }
  };
  if (base_end > k && strategy == Strategy::kWindow) {
    size_t size = 1;
    while ((k & ~(size - 1)) + size < base_end) size *= 2;
    const size_t position = k & ~(size - 1);
    place(b.folded_chunk(base, k, position, size), position);
  }
  size_t position = transform_n;
  while (position < n) {
    const size_t alignment = position & (~position + 1);
    const size_t size = std::min({alignment, next_pow2(n - position), transform_n});
    place(b.folded_chunk(base, k, position, size), position);
    position += size;
  }
  if (base_end > k && strategy == Strategy::kInPlace) {
    b.fft_with_zero_tail(base.data(), transform_n, 0, k);
    for (size_t i = 0; i < base_end - k; ++i) parity[i] = base[k + i];
  }
  return lower(b, parity, k, m);
}

// src/fft.rs:393-467: interpolation, folded trailing chunks with plain
// forward transforms, then the base transform last.
Fft16Program build_byte(size_t k, size_t m, bool liveness) {
  const size_t n = k + m;
  const size_t transform_n = next_pow2(k);
  Builder b;
  b.f = &kByte;
  b.zero_tail = false;
  const std::vector<uint16_t> base = b.alloc(transform_n);
  b.interpolate(base.data(), transform_n, 0, k);
  std::vector<uint16_t> parity(m, 0);
  size_t position = transform_n;
  while (position < n) {
    const size_t alignment = position & (~position + 1);
    const size_t size = std::min({alignment, next_pow2(n - position), transform_n});
    const std::vector<uint16_t> chunk = b.folded_chunk(base, k, position, size);
    for (size_t off = 0; off < size; ++off) {
      const size_t point = position + off;
      if (point >= k && point < n) parity[point - k] = chunk[off];
    }
    position += size;
  }
  b.transform(false, base.data(), transform_n, 0);
  const size_t base_end = std::min(n, transform_n);
  for (size_t i = k; i < base_end; ++i) parity[i - k] = base[i];
  return lower(b, parity, k, m, liveness);
}

}  // namespace

Fft16Program build_staged_program8(size_t k, size_t m, bool liveness) {
  return build_byte(k, m, liveness);
}

bool staged8_reference_route(size_t k, size_t m) {
  static const size_t kGenerated[][2] = {{7, 13}, {10, 10}, {14, 14}, {16, 16}, {18, 6}, {32, 32}};
  for (const auto &g : kGenerated)
    if (g[0] == k && g[1] == m) return false;
  if (k < 2 || (k & (k - 1)) || k > kMaxFftDataShards8 || m == 0) return false;
  const Fft16Program p = build_byte(k, m, false);
  if (p.register_count > kMaxStagedRegisters8) return false;
  return p.mult_count * 2 < k * m;
}

Fft16Program build_staged_program16(size_t k, size_t m) {
  const size_t n = k + m;
  const size_t transform_n = next_pow2(k);
  const size_t in_base = std::min(n, transform_n) > k ? std::min(n, transform_n) - k : 0;
  if (in_base == 0) return build_with(k, m, Strategy::kWindow);
  Fft16Program w = build_with(k, m, Strategy::kWindow);
  Fft16Program ip = build_with(k, m, Strategy::kInPlace);
  return w.mult_count < ip.mult_count ? std::move(w) : std::move(ip);
}

bool fft16_profitable(const Fft16Program &p, size_t k, size_t m) {
  if (p.register_count > kMaxStagedRegisters16) return false;
  return 9 * p.mult_count < 2 * k * m;
}

}  // namespace trs
