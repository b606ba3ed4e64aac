// Host side of the MI355X Reed-Solomon library: the C ABI of trs_gpu.h.
//
// This layer keeps the reference codec's semantics -- validation order,
// error codes, the systematic generator, first-k-present decode plans and
// the 64-entry most-recent-first plan cache -- and hands every byte of shard
// traffic to the HIP kernels in kernels.hip.  Nothing here touches shard
// bytes on the host.
//
// Reconstruct runs as ONE fused apply: the plan's matrix maps the k
// survivors straight to every missing shard.  Rows for missing data are rows
// of inv(G_valid) (src/reedsolomon.rs:539-571); rows for missing parity are
// G_parity[p] * inv(G_valid), i.e. the parity row applied to the data that
// inv(G_valid) rebuilds.  The reference applies those two steps in sequence
// (src/reedsolomon.rs:866-889); over a field the composition is exact, so the
// fused result is byte-identical while the data re-read of the second pass
// disappears.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <shared_mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/trs_gpu.h"
#include "fft16.hpp"
#include "field.hpp"
#include "kernels.hpp"

namespace trs {

const Gf8 &gf8() {
  static const Gf8 t;
  return t;
}

Gf16::Gf16() : log(65536, 0), exp(2 * 65535, 0) {
  const Gf8 &f = gf8();
  auto pw = [&](uint16_t base, unsigned e) {
    uint16_t r = 1;
    while (e) {
      if (e & 1) r = tower_mul_def(f, r, base);
      base = tower_mul_def(f, base, base);
      e >>= 1;
    }
    return r;
  };
  // Smallest element of full multiplicative order 65535 = 3*5*17*257.
  uint16_t g = 0;
  for (unsigned c = 2; c < 65536 && !g; ++c) {
    const uint16_t x = uint16_t(c);
    if (pw(x, 65535 / 3) != 1 && pw(x, 65535 / 5) != 1 && pw(x, 65535 / 17) != 1 &&
        pw(x, 65535 / 257) != 1)
      g = x;
  }
  uint16_t x = 1;
  for (unsigned i = 0; i < 65535; ++i) {
    exp[i] = x;
    exp[i + 65535] = x;
    log[x] = uint16_t(i);
    x = tower_mul_def(f, x, g);
  }
}

const Gf16 &gf16() {
  static const Gf16 t;
  return t;
}

namespace {

int hip_status(hipError_t e) {
  if (e == hipSuccess) return TRS_OK;
  return e == hipErrorOutOfMemory ? TRS_ERR_OUT_OF_MEMORY : TRS_ERR_DEVICE;
}

#define TRS_TRY(expr)                      \
  do {                                     \
    const hipError_t e_ = (expr);          \
    if (e_ != hipSuccess) return hip_status(e_); \
  } while (0)

// Owned device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t bytes) {
    reset();
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) n = bytes;
    else p = nullptr;
    return e;
  }
  hipError_t ensure(size_t bytes) { return bytes <= n ? hipSuccess : alloc(bytes); }
};

size_t pad_rows(size_t nout) {
  const size_t s = size_t(kOutStride);
  return ((nout ? nout : 1) + s - 1) / s * s;
}

// Coefficients in the kernels' layout: for each plane, byte (i*ostride + j)
// is output row j's coefficient for input i; rows past nout are zero.
// GF(2^8) coefficients are followed by the bit-sliced kernels' form of the
// same rows (planes_ptr): each row j scaled by 2^t_j, then the ostride bytes
// t_j (int8).
struct DevCoef {
  DevBuf buf;
  uint32_t k = 0, nout = 0, ostride = 0;
  bool planes = false;
  const uint32_t *ptr() const { return static_cast<const uint32_t *>(buf.p); }
  const uint32_t *planes_ptr() const {
    return planes ? ptr() + size_t(k) * ostride / 4 : nullptr;
  }
};

// The bit-sliced kernels apply a coefficient a nibble at a time, XORing the
// nibble's set powers x*2^b into the output two per three-input XOR, so a
// coefficient costs 8 * (ceil(s_lo / 2) + ceil(s_hi / 2)) operations for s_lo
// and s_hi set bits in its nibbles.  Scaling output row j by 2^t (t in
// [-7, 7]) and multiplying the result by 2^-t once, at 3 XORs per step in
// plane form, changes the bits but not the bytes written; t_j minimises that
// cost over the scaled row plus 3 |t|.
int nibble_cost(uint8_t c) {
  return (__builtin_popcount(c & 15u) + 1) / 2 + (__builtin_popcount(c >> 4) + 1) / 2;
}

int8_t planes_scale(const uint8_t *row, size_t k, size_t stride) {
  const Gf8 &g = gf8();
  int best_t = 0;
  long best = -1;
  for (int t = -7; t <= 7; ++t) {
    const uint8_t s = g.exp[(t + 255) % 255];
    long cost = 3L * (t < 0 ? -t : t);
    for (size_t i = 0; i < k; ++i) cost += 8L * nibble_cost(g.mul[row[i * stride]][s]);
    if (best < 0 || cost < best) {
      best = cost;
      best_t = t;
    }
  }
  return int8_t(best_t);
}

// rows: nout x k GF(2^8) coefficients, row-major.
int upload_coef8(DevCoef &dc, const uint8_t *rows, size_t nout, size_t k) {
  const size_t os = pad_rows(nout);
  std::vector<uint8_t> t(2 * k * os + os, 0);
  for (size_t j = 0; j < nout; ++j)
    for (size_t i = 0; i < k; ++i) t[i * os + j] = rows[j * k + i];
  const Gf8 &g = gf8();
  uint8_t *scaled = t.data() + k * os;
  for (size_t j = 0; j < nout; ++j) {
    const int8_t sc = planes_scale(rows + j * k, k, 1);
    const uint8_t s = g.exp[(sc + 255) % 255];
    for (size_t i = 0; i < k; ++i) scaled[i * os + j] = g.mul[rows[j * k + i]][s];
    scaled[k * os + j] = uint8_t(sc);
  }
  dc.planes = true;
  TRS_TRY(dc.buf.alloc(t.size()));
  TRS_TRY(hipMemcpy(dc.buf.p, t.data(), t.size(), hipMemcpyHostToDevice));
  dc.k = uint32_t(k);
  dc.nout = uint32_t(nout);
  dc.ostride = uint32_t(os);
  return TRS_OK;
}

// rows: nout x k tower coefficients; three planes of Karatsuba products
// [c_lo, c_hi^c_lo, c_hi*L] (src/fft16.rs:176-180).
int upload_coef16(DevCoef &dc, const uint16_t *rows, size_t nout, size_t k) {
  const size_t os = pad_rows(nout);
  const size_t plane = k * os;
  std::vector<uint8_t> t(3 * plane, 0);
  const Gf8 &f = gf8();
  for (size_t j = 0; j < nout; ++j)
    for (size_t i = 0; i < k; ++i) {
      const uint16_t c = rows[j * k + i];
      const uint8_t hi = c >> 8, lo = c & 0xff;
      t[i * os + j] = lo;
      t[plane + i * os + j] = hi ^ lo;
      t[2 * plane + i * os + j] = f.mul[hi][kTowerL];
    }
  dc.planes = false;
  TRS_TRY(dc.buf.alloc(t.size()));
  TRS_TRY(hipMemcpy(dc.buf.p, t.data(), t.size(), hipMemcpyHostToDevice));
  dc.k = uint32_t(k);
  dc.nout = uint32_t(nout);
  dc.ostride = uint32_t(os);
  return TRS_OK;
}

// A fragments of the int8-MFMA apply (mfma.hip) for an nout x k GF(2^8)
// matrix: [pass][K step][M tile][lane][16 bytes].  Lane l, element e of M tile
// mt holds A[row][kk] with row = l%32 and kk = 16*(l/32) + e, where row maps to
// output bit (o, r) through the accumulator layout (register v of lane half h
// holds row 8*(v/4) + 4*h + v%4; registers 0..7 / 8..15 are the 8 bits of
// outputs 2h / 2h+1 of the tile) and kk to input bit (i, b) = (4s + kk/8, kk%8).
// bit(o, r, in, b) is the GF(2) matrix entry of output bit (o, r), input bit
// (in, b).
template <class Bit>
std::vector<uint8_t> build_mfma_frags_bits(size_t nout, size_t k, Bit bit) {
  const size_t ks = mfma_steps(uint32_t(k));
  const size_t passes = (nout + 7) / 8;
  std::vector<uint8_t> f(passes * ks * 2 * 64 * 16, 0);
  for (size_t p = 0; p < passes; ++p)
    for (size_t s = 0; s < ks; ++s)
      for (size_t mt = 0; mt < 2; ++mt)
        for (size_t l = 0; l < 64; ++l) {
          const size_t row = l % 32;
          const size_t half = (row / 4) % 2;
          const size_t v = 4 * (row / 8) + row % 4;
          const size_t o = p * 8 + mt * 4 + 2 * half + v / 8;
          const size_t r = v % 8;
          uint8_t *dst = &f[(((p * ks + s) * 2 + mt) * 64 + l) * 16];
          for (size_t e = 0; e < 16; ++e) {
            const size_t in = 4 * s + 2 * (l / 32) + e / 8;
            const size_t b = e % 8;
            if (o < nout && in < k) dst[e] = bit(o, r, in, b);
          }
        }
  return f;
}

// GF(2^8): bit r of c_oi * 2^b.
std::vector<uint8_t> build_mfma_frags(const uint8_t *rows, size_t nout, size_t k) {
  const Gf8 &g = gf8();
  return build_mfma_frags_bits(nout, k, [&](size_t o, size_t r, size_t in, size_t b) {
    return uint8_t((g.mul[rows[o * k + in]][1u << b] >> r) & 1u);
  });
}

// The tower over split-plane shards as a byte problem: the low and high
// planes of tower shard i are byte inputs 2i and 2i+1 (bits 0..7 and 8..15
// of its symbols), likewise for outputs, and multiplying by a tower constant
// is GF(2)-linear on the 16 bits, so an nout x k tower matrix is a (2 nout) x
// (2 k) byte problem whose bit matrix holds bit 8p + r of c_ji * x^(8q + b)
// at output bit (2j + p, r), input bit (2i + q, b).
std::vector<uint8_t> build_mfma_frags16(const uint16_t *rows, size_t nout, size_t k) {
  const Gf16 &g = gf16();
  return build_mfma_frags_bits(2 * nout, 2 * k, [&](size_t o, size_t r, size_t in, size_t b) {
    const uint16_t prod = g.mul(rows[(o / 2) * k + in / 2], uint16_t(1u << (8 * (in % 2) + b)));
    return uint8_t((prod >> (8 * (o % 2) + r)) & 1u);
  });
}

using Mask = std::vector<uint64_t>;

Mask make_mask(const uint8_t *present, size_t n) {
  Mask m((n + 63) / 64, 0);
  for (size_t i = 0; i < n; ++i)
    if (present[i]) m[i / 64] |= uint64_t(1) << (i % 64);
  return m;
}

bool mask_get(const Mask &m, size_t i) { return (m[i / 64] >> (i % 64)) & 1; }

// Presence mask and count of n flags (nonzero = present), eight flags at a
// time: each byte is folded to its bit 0 (the OR of its bits), and the
// multiply gathers bit 0 of byte i into bit 56 + i.
void flags_mask(const uint8_t *f, size_t n, Mask &mask, size_t &count) {
  mask.assign((n + 63) / 64, 0);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, f + i, 8);
    w |= w >> 4;
    w |= w >> 2;
    w |= w >> 1;
    w &= 0x0101010101010101ull;
    mask[i / 64] |= ((w * 0x0102040810204080ull) >> 56) << (i % 64);
  }
  for (; i < n; ++i)
    if (f[i]) mask[i / 64] |= uint64_t(1) << (i % 64);
  count = 0;
  for (uint64_t w : mask) count += size_t(__builtin_popcountll(w));
}

// Everything one erasure pattern needs (src/reedsolomon.rs:117-130).
struct Plan {
  Mask presence;
  std::vector<uint32_t> valid;           // first k present indices, ascending
  std::vector<uint32_t> missing_data;    // ascending
  std::vector<uint32_t> missing_parity;  // ascending
  DevCoef coef;                          // [missing data rows; missing parity rows] x k
  std::vector<uint8_t> rows8;            // host copy (GF(2^8) plans)
  DevBuf frag;                           // MFMA A fragments of the same rows (GF(2^8) k <= 32,
                                         // tower k <= 16)
  // Tower plans build their matrix rows (coef) on first use of the matrix
  // route, and the tower-FFT erasure decoder on first use of the FFT route
  // (it loads the valid shards and stores [missing data; missing parity]).
  std::once_flag coef_once;
  int coef_status = 0;
  std::once_flag fft_once;
  hipError_t fft_status = hipSuccess;
  Fft16DevicePtr fft;
  // The decoder's pointer slots as shard indices (valid, missing data,
  // missing parity), for stripes addressed as base + index * stride.
  DevBuf fft_idx;
};

size_t log2_ceil(size_t n) {
  size_t l = 0;
  while ((size_t(1) << l) < n) ++l;
  return l;
}

// Decode route for a tower reconstruct: 1 = matrix, 2 = FFT, 0 = FFT when the
// matrix apply (k * nout products per symbol) costs more than one transform
// over the N = 2^ceil(log2 n) points (N log2 N) and the shape fits.
bool use_fft_decode(int route, size_t n, size_t k, size_t nout) {
  if (route == 1 || nout == 0 || n > kMaxDecodePoints) return false;
  if (route == 2) return true;
  const size_t lg = std::max<size_t>(1, log2_ceil(n));
  return k * nout >= (size_t(1) << lg) * lg;
}

// The plan's FFT erasure decoder (tower, or byte field for GF(2^8) plans),
// built on first use.
// Decode route for a GF(2^8) reconstruct (route as trs_rs8_set_decode_route).
// Automatic takes the FFT decoder where it measured faster than the matrix
// apply, per call at full size (14 shapes x 1..m erasures x 16 KiB - 4 MiB),
// only for k > 32 (MFMA covers k <= 32):
// - below 128 KiB, where the matrix route is the table kernel and leaves the
//   chip mostly idle (the decoder is 1.4-8x faster);
// - up to 256 KiB when the decoder's N-point transforms (N = 2^ceil(log2 n),
//   about 0.47 N us there) undercut the planes kernel (about 0.72 k +
//   0.25 nout us): 100 N < 170 k + 60 nout;
// - beyond, when k * nout products per column outweigh the transforms
//   (about 0.9 N log2 N against 0.133 k nout us per 4 MiB):
//   k nout > 7 N log2 N.
bool use_fft_decode8(int route, size_t n, size_t k, size_t nout, size_t len) {
  if (route == 1 || nout == 0 || n > 256) return false;
  if (route == 2) return true;
  if (k <= 32) return false;
  if (len < (size_t(128) << 10)) return true;
  size_t N = 2, lg = 1;
  while (N < n) {
    N <<= 1;
    ++lg;
  }
  if (len <= (size_t(256) << 10)) return 100 * N < 170 * k + 60 * nout;
  return k * nout > 7 * N * lg;
}

// The same choice for a batched reconstruct, where the work list fills the
// chip whatever the shard length: the full-size rule, k nout > 7 N log2 N.
bool use_fft_decode8_batched(int route, size_t n, size_t k, size_t nout) {
  if (route == 1 || nout == 0 || n > 256) return false;
  if (route == 2) return true;
  size_t N = 2, lg = 1;
  while (N < n) {
    N <<= 1;
    ++lg;
  }
  return k > 32 && k * nout > 7 * N * lg;
}

int plan_fft16(Plan &p, size_t n, const Fft16Device **out, bool byte_field = false) {
  std::call_once(p.fft_once, [&p, n, byte_field] {
    std::vector<uint32_t> outputs(p.missing_data);
    outputs.insert(outputs.end(), p.missing_parity.begin(), p.missing_parity.end());
    p.fft_status = fft16_device_upload(byte_field ? fft8_decode_lowered(n, p.valid, outputs)
                                                  : fft16_decode_lowered(n, p.valid, outputs),
                                       &p.fft);
    if (p.fft_status != hipSuccess) return;
    std::vector<uint32_t> idx(p.valid);
    idx.insert(idx.end(), outputs.begin(), outputs.end());
    p.fft_status = p.fft_idx.alloc(idx.size() * sizeof(uint32_t));
    if (p.fft_status == hipSuccess)
      p.fft_status = hipMemcpy(p.fft_idx.p, idx.data(), idx.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice);
  });
  if (p.fft_status == hipErrorInvalidValue) return TRS_ERR_UNSUPPORTED;
  if (p.fft_status != hipSuccess) return hip_status(p.fft_status);
  *out = p.fft.get();
  return TRS_OK;
}

int upload_frags(DevBuf &dst, const uint8_t *rows, size_t nout, size_t k) {
  if (!mfma_available(uint32_t(k)) || nout == 0) return TRS_OK;
  const auto f = build_mfma_frags(rows, nout, k);
  TRS_TRY(dst.alloc(f.size()));
  TRS_TRY(hipMemcpy(dst.p, f.data(), f.size(), hipMemcpyHostToDevice));
  return TRS_OK;
}

// Tower shapes whose byte problem fits the MFMA kernel (2k <= 32 byte inputs)
// and the kernel arguments' pointer list (2 (k + nout) <= 256).
bool mfma16_available(size_t k, size_t nout) {
  return mfma_available(uint32_t(2 * k)) && 2 * (k + nout) <= size_t(kInlinePtrs);
}

int upload_frags16(DevBuf &dst, const uint16_t *rows, size_t nout, size_t k) {
  if (!mfma16_available(k, nout) || nout == 0) return TRS_OK;
  const auto f = build_mfma_frags16(rows, nout, k);
  TRS_TRY(dst.alloc(f.size()));
  TRS_TRY(hipMemcpy(dst.p, f.data(), f.size(), hipMemcpyHostToDevice));
  return TRS_OK;
}

constexpr size_t kPlanCacheLimit = 64;  // src/reedsolomon.rs:24

struct PlanCache {
  mutable std::shared_mutex mu;
  mutable std::vector<std::pair<Mask, std::shared_ptr<Plan>>> entries;

  std::shared_ptr<Plan> find(const Mask &m) const {
    std::shared_lock<std::shared_mutex> lk(mu);
    for (auto &e : entries)
      if (e.first == m) return e.second;
    return nullptr;
  }
  // Insert at the front unless another thread already did (src/reedsolomon.rs:606-612).
  std::shared_ptr<Plan> insert(const Mask &m, std::shared_ptr<Plan> p) const {
    std::unique_lock<std::shared_mutex> lk(mu);
    for (auto &e : entries)
      if (e.first == m) return e.second;
    entries.insert(entries.begin(), {m, p});
    if (entries.size() > kPlanCacheLimit) entries.resize(kPlanCacheLimit);
    return p;
  }
};

// Splits indices by presence the way build_plan does (src/reedsolomon.rs:550-566).
void split_indices(const Mask &m, size_t k, size_t n, Plan &p) {
  for (size_t i = 0; i < n; ++i) {
    if (mask_get(m, i)) {
      if (p.valid.size() < k) p.valid.push_back(uint32_t(i));
    } else if (i < k) {
      p.missing_data.push_back(uint32_t(i));
    } else {
      p.missing_parity.push_back(uint32_t(i));
    }
  }
}

// Fused decode rows for the missing shards of plan p: row for missing index e
// is V_e * inv(V_P), P the first k present indices.  With G = V * inv(V_top)
// (src/reedsolomon.rs:154-159) the reference's rows are exactly these: the
// missing-data rows of inv(G_P) = V_top * inv(V_P) (src/reedsolomon.rs:
// 539-571) and the parity rows G_e * inv(G_P) = V_e * inv(V_P) of its second
// pass (:866-889).  V_e * inv(V_P) is the Lagrange basis over the nodes P
// evaluated at e, L_j(e) = prod_l (e - p_l) * w_j / (e - p_j) with barycentric
// weights w_j = 1 / prod_{l != j} (p_j - p_l), so the matrix comes out in
// O(k (k + e)) field operations instead of a k x k Gauss-Jordan.  The inverse is
// unique, so the rows are identical to the elimination's.
template <class F>
bool fused_rows(const Matrix<F> &G, size_t k, const Plan &p, std::vector<typename F::T> &rows) {
  using T = typename F::T;
  (void)G;
  const std::vector<uint32_t> &nodes = p.valid;
  if (nodes.size() != k) return false;
  std::vector<uint32_t> targets(p.missing_data);
  targets.insert(targets.end(), p.missing_parity.begin(), p.missing_parity.end());
  std::vector<T> w(k);
  parallel_rows(k, k, [&](size_t j0, size_t j1) {
    for (size_t j = j0; j < j1; ++j) {
      T d = 1;
      for (size_t l = 0; l < k; ++l)
        if (l != j) d = F::mul(d, T(nodes[j] ^ nodes[l]));
      w[j] = F::inv(d);
    }
  });
  rows.assign(targets.size() * k, 0);
  parallel_rows(targets.size(), 3 * k, [&](size_t t0, size_t t1) {
    for (size_t t = t0; t < t1; ++t) {
      const uint32_t e = targets[t];
      T lam = 1;
      for (size_t l = 0; l < k; ++l) lam = F::mul(lam, T(e ^ nodes[l]));
      T *o = rows.data() + t * k;
      for (size_t j = 0; j < k; ++j) o[j] = F::mul(F::mul(lam, w[j]), F::inv(T(e ^ nodes[j])));
    }
  });
  return true;
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

// Device flag per host thread for verify (verify waits for its answer, so one
// thread never has two in flight).  Kept for the life of the process: freeing
// it from a thread-exit or static destructor could run after the HIP runtime
// has shut down.
int verify_flag(uint32_t **out) {
  thread_local void *flag = nullptr;
  if (!flag) TRS_TRY(hipMalloc(&flag, sizeof(uint32_t)));
  *out = static_cast<uint32_t *>(flag);
  return TRS_OK;
}

// Single-coefficient matrices for the slice API: entry c holds c at byte 0.
// Allocated once and kept for the life of the process (see verify_flag).
int slice_coef(uint8_t c, const uint32_t **out) {
  static std::once_flag once;
  static void *buf = nullptr;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    std::vector<uint8_t> t(256 * kOutStride, 0);
    for (int c = 0; c < 256; ++c) t[size_t(c) * kOutStride] = uint8_t(c);
    err = hipMalloc(&buf, t.size());
    if (err == hipSuccess) err = hipMemcpy(buf, t.data(), t.size(), hipMemcpyHostToDevice);
  });
  if (err != hipSuccess) return hip_status(err);
  *out = reinterpret_cast<const uint32_t *>(static_cast<const uint8_t *>(buf) +
                                            size_t(c) * kOutStride);
  return TRS_OK;
}

// Device copy of a pointer table, stream-ordered (allocated, filled through
// pinned staging and freed on `stream`).
struct PtrTable {
  void *d = nullptr;
  hipStream_t s = nullptr;
  ~PtrTable() {
    if (d) (void)hipFreeAsync(d, s);
  }
  int upload(const std::vector<const uint8_t *> &v, hipStream_t stream) {
    s = stream;
    const size_t bytes = v.size() * sizeof(void *);
    TRS_TRY(hipMallocAsync(&d, bytes ? bytes : 8, stream));
    if (bytes) TRS_TRY(staged_upload(d, v.data(), bytes, stream));
    return TRS_OK;
  }
};

// One stripe of a device program: ptrs = its loaded then stored shards,
// `plane` columns.  Up to kFftInlinePtrs pointers travel in the kernel
// arguments; more go through a device table.
// flag non-null: verify (compare the stored registers with the shards).
int launch_fft16_one(const Fft16Device &d, const std::vector<const uint8_t *> &ptrs,
                     size_t nstore, uint64_t plane, hipStream_t stream, bool byte_planes = false,
                     uint32_t *flag = nullptr, uint64_t lds_target = 0) {
  bool al = true;
  for (auto q : ptrs) al = al && aligned16(q);
  Fft16Call c{};
  c.plane = plane;
  c.stripes = 1;
  c.nstore = uint32_t(nstore);
  c.vec_ok = al ? 1u : 0u;
  c.flag = flag;
  c.lds_target = lds_target;
  if (ptrs.size() <= kFftInlinePtrs) {
    Fft16Ptrs P{};
    std::copy(ptrs.begin(), ptrs.end(), P.p);
    return hip_status(launch_fft16(d, c, stream, &P, byte_planes));
  }
  PtrTable table;
  int st = table.upload(ptrs, stream);
  if (st) return st;
  c.ptrs = static_cast<const uint8_t *const *>(table.d);
  return hip_status(launch_fft16(d, c, stream, nullptr, byte_planes));
}

// ---- host-resident shards ----------------------------------------------------
// The reference's buffers are host slices.  Calls on host memory stream their
// shards through device memory in column chunks: the host-to-device copy of
// chunk c+1, the kernels of chunk c and the device-to-host copy of chunk c-1
// run on three streams at once.

// Library-owned staging of host-resident calls: S slots, each holding one
// chunk of every shard row of a call in device memory and, for rows packed
// through staging, in pinned host memory; one stream per pipeline stage.
// Kept per host thread and grown on demand.
struct HostPipe {
  static constexpr int S = 3;
  void *dev[S] = {};
  void *pin[S] = {};
  size_t dev_bytes = 0, pin_bytes = 0;
  hipStream_t h2d = nullptr, comp = nullptr, d2h = nullptr;
  hipEvent_t e_h2d[S] = {}, e_comp[S] = {}, e_d2h[S] = {};
  bool ready = false;
  ~HostPipe() {
    for (int s = 0; s < S; ++s) {
      if (dev[s]) (void)hipFree(dev[s]);
      if (pin[s]) (void)hipHostFree(pin[s]);
      if (e_h2d[s]) (void)hipEventDestroy(e_h2d[s]);
      if (e_comp[s]) (void)hipEventDestroy(e_comp[s]);
      if (e_d2h[s]) (void)hipEventDestroy(e_d2h[s]);
    }
    if (h2d) (void)hipStreamDestroy(h2d);
    if (comp) (void)hipStreamDestroy(comp);
    if (d2h) (void)hipStreamDestroy(d2h);
  }
  hipError_t setup(size_t bytes, bool staged) {
    hipError_t e;
    if (!ready) {
      for (hipStream_t *st : {&h2d, &comp, &d2h})
        if ((e = hipStreamCreateWithFlags(st, hipStreamNonBlocking)) != hipSuccess) return e;
      for (int s = 0; s < S; ++s)
        for (hipEvent_t *ev : {&e_h2d[s], &e_comp[s], &e_d2h[s]})
          if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) return e;
      ready = true;
    }
    if (bytes > dev_bytes) {
      for (int s = 0; s < S; ++s) {
        if (dev[s]) (void)hipFree(dev[s]);
        dev[s] = nullptr;
      }
      dev_bytes = 0;
      for (int s = 0; s < S; ++s)
        if ((e = hipMalloc(&dev[s], bytes)) != hipSuccess) return e;
      dev_bytes = bytes;
    }
    if (staged && bytes > pin_bytes) {
      for (int s = 0; s < S; ++s) {
        if (pin[s]) (void)hipHostFree(pin[s]);
        pin[s] = nullptr;
      }
      pin_bytes = 0;
      for (int s = 0; s < S; ++s)
        if ((e = hipHostMalloc(&pin[s], bytes, hipHostMallocDefault)) != hipSuccess) return e;
      pin_bytes = bytes;
    }
    return hipSuccess;
  }
};

HostPipe &host_pipe() {
  static thread_local HostPipe pipe;
  return pipe;
}

enum class HostMem { kPageable, kPinned, kDevice };

HostMem host_mem(const void *p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory is not known to HIP: clear the query's error
    return HostMem::kPageable;
  }
  if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeArray) return HostMem::kDevice;
  return a.type == hipMemoryTypeHost ? HostMem::kPinned : HostMem::kPageable;
}

constexpr size_t kDefaultHostChunk = size_t(1) << 20;  // bytes of each shard per chunk
constexpr size_t kHostSlotMax = size_t(64) << 20;      // bytes of all rows per chunk
constexpr size_t kDirectMinRow = size_t(256) << 10;
std::atomic<size_t> g_host_chunk{0};

// One host-resident call pipelined over column chunks.  Host shard i of
// `len` bytes holds `planes` planes of len / planes bytes (2 for the tower's
// split-plane shards); chunk c of every shard is the same column range of
// each plane, laid out on the device as a shard of its own ([plane 0 piece |
// plane 1 piece]).  in_host rows are read, out_host rows written; run(rows,
// chunk_bytes, stream) gets nin + nout device rows.  Pinned callers with long
// chunks are copied row by row; other rows are packed into pinned staging and
// cross the bus in one copy per direction and chunk.
template <typename Run>
int host_pipeline(const uint8_t *const *in_host, size_t nin, uint8_t *const *out_host,
                  size_t nout, size_t len, size_t planes, Run &&run) {
  const size_t rows = nin + nout;
  if (rows == 0 || len == 0) return TRS_OK;
  const size_t plane = len / planes;
  size_t chunk = g_host_chunk.load();
  if (chunk == 0) chunk = kDefaultHostChunk;
  size_t cols = std::max<size_t>(chunk / planes, 64);
  cols = std::min(cols, std::max<size_t>(kHostSlotMax / (rows * planes), 64));
  cols &= ~size_t(63);  // chunk pieces keep the caller's 64-byte alignment
  cols = std::min(cols, plane);
  const size_t rowb = cols * planes;
  bool pinned = true;
  for (size_t r = 0; r < rows; ++r) {
    const void *q = r < nin ? static_cast<const void *>(in_host[r]) : out_host[r - nin];
    const HostMem k = host_mem(q);
    if (k == HostMem::kDevice) return TRS_ERR_INVALID_ARGUMENT;  // a device shard in a host call
    pinned = pinned && k == HostMem::kPinned;
  }
  const bool direct = pinned && rowb >= kDirectMinRow;
  HostPipe &P = host_pipe();
  TRS_TRY(P.setup(rows * rowb, !direct));
  const size_t nchunks = (plane + cols - 1) / cols;
  auto drain = [&](size_t c) -> int {  // chunk c's outputs are home once this returns
    const int s = int(c % HostPipe::S);
    TRS_TRY(hipEventSynchronize(P.e_d2h[s]));
    if (!direct) {
      const size_t off = c * cols, cl = std::min(cols, plane - off);
      const uint8_t *pin = static_cast<const uint8_t *>(P.pin[s]) + nin * rowb;
      for (size_t j = 0; j < nout; ++j)
        for (size_t p = 0; p < planes; ++p)
          std::memcpy(out_host[j] + p * plane + off, pin + j * rowb + p * cl, cl);
    }
    return TRS_OK;
  };
  auto body = [&]() -> int {
    std::vector<uint8_t *> dev_rows(rows);
    for (size_t c = 0; c < nchunks; ++c) {
      const int s = int(c % HostPipe::S);
      const size_t off = c * cols, cl = std::min(cols, plane - off);
      if (c >= size_t(HostPipe::S)) {
        const int st = drain(c - HostPipe::S);  // frees slot s
        if (st) return st;
      }
      uint8_t *dev = static_cast<uint8_t *>(P.dev[s]);
      uint8_t *pin = static_cast<uint8_t *>(P.pin[s]);
      for (size_t r = 0; r < rows; ++r) dev_rows[r] = dev + r * rowb;
      if (direct) {
        for (size_t i = 0; i < nin; ++i)
          for (size_t p = 0; p < planes; ++p)
            TRS_TRY(hipMemcpyAsync(dev + i * rowb + p * cl, in_host[i] + p * plane + off, cl,
                                   hipMemcpyHostToDevice, P.h2d));
      } else if (nin) {
        for (size_t i = 0; i < nin; ++i)
          for (size_t p = 0; p < planes; ++p)
            std::memcpy(pin + i * rowb + p * cl, in_host[i] + p * plane + off, cl);
        TRS_TRY(hipMemcpyAsync(dev, pin, nin * rowb, hipMemcpyHostToDevice, P.h2d));
      }
      TRS_TRY(hipEventRecord(P.e_h2d[s], P.h2d));
      TRS_TRY(hipStreamWaitEvent(P.comp, P.e_h2d[s], 0));
      const int st = run(dev_rows.data(), cl * planes, P.comp);
      if (st) return st;
      TRS_TRY(hipEventRecord(P.e_comp[s], P.comp));
      TRS_TRY(hipStreamWaitEvent(P.d2h, P.e_comp[s], 0));
      if (direct) {
        for (size_t j = 0; j < nout; ++j)
          for (size_t p = 0; p < planes; ++p)
            TRS_TRY(hipMemcpyAsync(out_host[j] + p * plane + off, dev + (nin + j) * rowb + p * cl,
                                   cl, hipMemcpyDeviceToHost, P.d2h));
      } else if (nout) {
        TRS_TRY(hipMemcpyAsync(pin + nin * rowb, dev + nin * rowb, nout * rowb,
                               hipMemcpyDeviceToHost, P.d2h));
      }
      TRS_TRY(hipEventRecord(P.e_d2h[s], P.d2h));
    }
    const size_t first = nchunks > size_t(HostPipe::S) ? nchunks - HostPipe::S : 0;
    for (size_t c = first; c < nchunks; ++c) {
      const int st = drain(c);
      if (st) return st;
    }
    return TRS_OK;
  };
  const int st = body();
  if (st) {  // nothing of this call may still be in flight on return
    (void)hipStreamSynchronize(P.h2d);
    (void)hipStreamSynchronize(P.comp);
    (void)hipStreamSynchronize(P.d2h);
  }
  return st;
}

// Where a call's shards live.  Device pointers: the work is enqueued on
// `stream`.  Host pointers (`host`): the shards stream through device memory
// by host_pipeline, `planes` planes per shard.
struct Exec {
  hipStream_t stream = nullptr;
  bool host = false;
  size_t planes = 1;
  Exec(hipStream_t s) : stream(s) {}  // NOLINT: a stream means device shards
  static Exec on_host(size_t planes) {
    Exec e(nullptr);
    e.host = true;
    e.planes = planes;
    return e;
  }
};

// Runs launch(ptrs, len, stream) over a validated shard list: ptrs[0, nin)
// are read, the rest written.  Device calls launch once; host calls launch
// once per chunk on device copies of the rows, with the chunk's length.
template <typename Launch>
int dispatch(const Exec &ex, const std::vector<const uint8_t *> &ptrs, size_t nin, size_t len,
             Launch &&launch) {
  if (!ex.host) return launch(ptrs, len, ex.stream);
  std::vector<uint8_t *> out;
  for (size_t j = nin; j < ptrs.size(); ++j) out.push_back(const_cast<uint8_t *>(ptrs[j]));
  std::vector<const uint8_t *> dev(ptrs.size());
  return host_pipeline(ptrs.data(), nin, out.data(), out.size(), len, ex.planes,
                       [&](uint8_t **rows, size_t cl, hipStream_t s) {
                         dev.assign(rows, rows + ptrs.size());
                         return launch(dev, cl, s);
                       });
}

}  // namespace
}  // namespace trs

using namespace trs;

// ============================================================================
// codec objects
// ============================================================================

struct trs_rs8 {
  size_t k = 0, m = 0, n = 0;
  int route = kRouteTable;  // encode route: bit-sliced FFT > bit-sliced matrix > table kernel
  std::atomic<int> apply_route{0};  // matrix applies: 0 auto, 1 table kernel, 2 int8 MFMA
  Matrix<F8> G;
  DevCoef parity;
  DevBuf parity_frag;  // MFMA A fragments of the parity rows (k <= 32)
  PlanCache plans;
  // The GF(2^8) staged FFT program (src/fft.rs:393-592) on the LDS
  // interpreter with byte-plane registers (its constants all lie in the
  // subfield).  The automatic encode route takes it for lengths up to
  // `staged_max` (0: never).  The device form is built and uploaded on first
  // use.
  size_t staged_max = 0;
  // reconstruct: 0 auto, 1 matrix plans, 2 the FFT erasure decoder on the
  // byte-plane interpreter
  std::atomic<int> decode_route{0};
  mutable std::once_flag staged_once;
  mutable hipError_t staged_status = hipSuccess;
  mutable Fft16DevicePtr staged_dev;
};

struct trs_dec8 {
  size_t k = 0, n = 0;
  int apply_route = 0;   // the codec's apply route when the decoder was prepared
  int decode_route = 0;  // and its reconstruct route
  std::shared_ptr<Plan> plan;
};

struct trs_rs16 {
  size_t k = 0, m = 0, n = 0;
  Matrix<F16> G;
  DevCoef parity;
  DevBuf parity_frag;  // int8-MFMA fragments of the parity rows (k <= 16)
  PlanCache plans;
  std::atomic<int> decode_route{0};  // reconstruct: 0 auto, 1 matrix, 2 tower FFT
  // matrix applies: 0 auto, 1 the Karatsuba table kernel, 2 int8 MFMA (k <= 16)
  std::atomic<int> apply_route{0};
  // The staged tower-FFT program (src/reedsolomon16.rs:300-308), built for
  // 2 <= k <= 1024 and k + m <= 4096.  `staged`: the encode route takes it.
  // The reference routes by a multiply-count rule measured on x86 GFNI
  // (src/fft16.rs:85-106); on this GPU the interpreter beat the matrix kernel
  // on 1771 of 1774 sampled tower shapes (the cost-model study), so the route
  // is the FFT whenever the program's register file fits one workgroup's LDS.
  // Both routes are byte-identical.  The device form is uploaded on first use.
  bool has_program = false;
  bool staged = false;
  Fft16Program program;
  mutable std::once_flag fft_once;
  mutable hipError_t fft_status = hipSuccess;
  mutable Fft16DevicePtr fft;
};

struct trs_dec16 {
  size_t k = 0, n = 0;
  int decode_route = 0;  // the codec's decode and apply routes when the decoder was prepared
  int apply_route = 0;
  std::shared_ptr<Plan> plan;
};

namespace {

// ---- GF(2^8) plan build / lookup (src/reedsolomon.rs:537-613) --------------
int plan8_for(const trs_rs8 *rs, const Mask &m, std::shared_ptr<Plan> *out) {
  if (auto p = rs->plans.find(m)) {
    *out = p;
    return TRS_OK;
  }
  auto p = std::make_shared<Plan>();
  p->presence = m;
  split_indices(m, rs->k, rs->n, *p);
  if (!fused_rows(rs->G, rs->k, *p, p->rows8)) return TRS_SINGULAR_MATRIX;
  const size_t nrows = p->missing_data.size() + p->missing_parity.size();
  int st = upload_coef8(p->coef, p->rows8.data(), nrows, rs->k);
  if (st) return st;
  if ((st = upload_frags(p->frag, p->rows8.data(), nrows, rs->k))) return st;
  *out = rs->plans.insert(m, p);
  return TRS_OK;
}

int plan16_for(const trs_rs16 *rs, const Mask &m, std::shared_ptr<Plan> *out) {
  if (auto p = rs->plans.find(m)) {
    *out = p;
    return TRS_OK;
  }
  auto p = std::make_shared<Plan>();
  p->presence = m;
  split_indices(m, rs->k, rs->n, *p);
  if (p->valid.size() < rs->k) return TRS_TOO_FEW_SHARDS_PRESENT;
  *out = rs->plans.insert(m, p);
  return TRS_OK;
}

// The matrix rows of a tower plan, built and uploaded once.
int plan16_matrix(Plan &p, size_t k) {
  std::call_once(p.coef_once, [&p, k] {
    std::vector<uint16_t> rows;
    if (!fused_rows(Matrix<F16>(), k, p, rows)) {
      p.coef_status = TRS_SINGULAR_MATRIX;
      return;
    }
    const size_t nrows = p.missing_data.size() + p.missing_parity.size();
    p.coef_status = upload_coef16(p.coef, rows.data(), nrows, k);
    if (!p.coef_status) p.coef_status = upload_frags16(p.frag, rows.data(), nrows, k);
  });
  return p.coef_status;
}

// scan_shards (src/reedsolomon.rs:801-821; RS16 adds the even-length rule,
// src/reedsolomon16.rs:758-789).  Returns the status; fills mask/count/len.
int scan_shards(const size_t *lens, const uint8_t *present, size_t n, bool even, Mask &mask,
                size_t &count, size_t &len) {
  mask.assign((n + 63) / 64, 0);
  count = 0;
  len = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!present[i]) continue;
    const size_t l = lens[i];
    if (l == 0) return TRS_EMPTY_SHARD;
    if (even && (l & 1)) return TRS_INCORRECT_SHARD_SIZE;
    if (len && l != len) return TRS_INCORRECT_SHARD_SIZE;
    len = l;
    mask[i / 64] |= uint64_t(1) << (i % 64);
    ++count;
  }
  return TRS_OK;
}

// The gather of reconstruct_with_plan (src/reedsolomon.rs:829-864): survivors
// are the plan's valid indices; every missing shard that will be written must
// have a buffer of the shard length.  Fills ptrs as [inputs..., outputs...].
int gather(const Plan &p, uint8_t *const *shards, const size_t *lens, size_t k, size_t len,
           bool data_only, std::vector<const uint8_t *> &ptrs, size_t &nout) {
  // The reference checks missing buffers in index order while gathering.
  for (uint32_t i : p.missing_data)
    if (lens[i] != len) return TRS_INCORRECT_SHARD_SIZE;
  if (!data_only)
    for (uint32_t i : p.missing_parity)
      if (lens[i] != len) return TRS_INCORRECT_SHARD_SIZE;
  ptrs.clear();
  for (size_t r = 0; r < k; ++r) ptrs.push_back(shards[p.valid[r]]);
  for (uint32_t i : p.missing_data) ptrs.push_back(shards[i]);
  nout = p.missing_data.size();
  if (!data_only) {
    for (uint32_t i : p.missing_parity) ptrs.push_back(shards[i]);
    nout += p.missing_parity.size();
  }
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return TRS_OK;
}

// Single-call runtime-coefficient applies from this shard length up run on
// the bit-sliced planes tile (8 KiB chunks, five waves per SIMD); shorter ones
// on the 4 KiB table tile, whose smaller tiles give a short launch more
// workgroups.  Measured per call at 128 KiB - 4 MiB on (10,4), (20,10),
// (48,16), (100,30): planes 1.0-1.5x the table tile at 128-512 KiB, 1.3-1.4x at
// 4 MiB.  TRS_APPLY_PLANES_MIN (bytes) overrides the crossover.
size_t planes_min_len() {
  static const size_t v = [] {
    size_t d = size_t(128) << 10;
    if (const char *e = std::getenv("TRS_APPLY_PLANES_MIN")) d = size_t(std::strtoull(e, nullptr, 10));
    return d;
  }();
  return v;
}

// variant: the kernel (kVarTable, kVarPlanes4, kVarPlanes8); -1 = by length.
int run_gf8(const std::vector<const uint8_t *> &ptrs, size_t k, size_t nout, const DevCoef &coef,
            size_t len, uint32_t mode, uint32_t *flag, hipStream_t stream, int variant = -1) {
  if (nout == 0 || len == 0) return TRS_OK;
  if (k + nout > size_t(kInlinePtrs)) return TRS_ERR_INVALID_ARGUMENT;
  Gf8Call c{};
  bool al = true;
  for (size_t i = 0; i < k + nout; ++i) {
    c.ptr[i] = ptrs[i];
    al = al && aligned16(ptrs[i]);
  }
  c.coef = coef.ptr();
  c.len = len;
  c.k = uint32_t(k);
  c.nout = uint32_t(nout);
  c.ostride = coef.ostride;
  c.vec_ok = al ? 1u : 0u;
  c.flag = flag;
  const bool planes = variant == 1 || variant == 2 || (variant < 0 && len >= planes_min_len());
  if (planes && coef.planes) {  // the bit-sliced tile takes the scaled rows
    c.coef = coef.planes_ptr();
    return hip_status(
        launch_gf8_planes(c, mode, stream, variant == 1 ? 4 : variant == 2 ? kPlanesOut : 0));
  }
  return hip_status(launch_gf8_apply(c, mode, stream));
}

// Route of a reconstruct apply.  Measured on MI355X (per-call applies of
// 226 sampled GF(2^8) shapes with k <= 32): the int8-MFMA kernel is about 2x
// faster than the table kernel below 256 KiB shards and still ahead below
// 512 KiB -- its 2048-column workgroup tiles give a small launch more
// workgroups -- and slower above.  Batched work lists keep the bit-sliced
// planes tile, 3.4x faster than the batched MFMA kernel on the mixed C2
// reconstruct.  Route 2 forces MFMA, route 1 the table kernel.
constexpr size_t kMfmaAutoMaxLen = size_t(512) << 10;

bool use_mfma(int route, size_t k, size_t len, bool batched) {
  if (!mfma_available(uint32_t(k))) return false;
  if (route == 2) return true;
  return route == 0 && !batched && len < kMfmaAutoMaxLen;
}


int run_mfma(const std::vector<const uint8_t *> &ptrs, size_t k, size_t nout, const void *frag,
             size_t len, hipStream_t stream) {
  if (nout == 0 || len == 0) return TRS_OK;
  if (k + nout > size_t(kInlinePtrs)) return TRS_ERR_INVALID_ARGUMENT;
  Gf8MfmaCall c{};
  bool al = true;
  for (size_t i = 0; i < k + nout; ++i) {
    c.ptr[i] = ptrs[i];
    al = al && (reinterpret_cast<uintptr_t>(ptrs[i]) & 3) == 0;
  }
  c.frag = frag;
  c.len = len;
  c.k = uint32_t(k);
  c.nout = uint32_t(nout);
  c.vec_ok = al ? 1u : 0u;
  return hip_status(launch_gf8_mfma(c, stream));
}

// A tower apply on the int8-MFMA kernel as its byte problem: shard i's low
// and high planes are byte shards 2i and 2i + 1 of len / 2 bytes.
int run_mfma16(const std::vector<const uint8_t *> &ptrs, size_t k, size_t nout, const void *frag,
               size_t len, hipStream_t stream) {
  if (nout == 0 || len == 0) return TRS_OK;
  if (!frag || !mfma16_available(k, nout)) return TRS_ERR_UNSUPPORTED;
  const size_t plane = len / 2;
  std::vector<const uint8_t *> b(2 * (k + nout));
  for (size_t i = 0; i < k + nout; ++i) {
    b[2 * i] = ptrs[i];
    b[2 * i + 1] = ptrs[i] + plane;
  }
  return run_mfma(b, 2 * k, 2 * nout, frag, plane, stream);
}

// Split-plane shards of `len` bytes (plane = len / 2).
int run_gf16(const std::vector<const uint8_t *> &ptrs, size_t k, size_t nout,
             const DevCoef &coef, size_t len, uint32_t mode, uint32_t *flag,
             hipStream_t stream) {
  if (nout == 0 || len == 0) return TRS_OK;
  const size_t plane = len / 2;
  bool al = (plane % 16) == 0;
  for (auto q : ptrs) al = al && aligned16(q);
  auto fill = [&](Gf16Call &c) {
    c.coef = coef.ptr();
    c.plane = plane;
    c.col0 = 0;
    c.cols = plane;
    c.k = uint32_t(k);
    c.nout = uint32_t(nout);
    c.ostride = coef.ostride;
    c.vec_ok = al ? 1u : 0u;
    c.mode = mode;
    c.flag = flag;
  };
  if (k + nout <= size_t(kInlinePtrs)) {  // pointers in the kernel arguments
    Gf16InlineCall c{};
    fill(c.call);
    for (size_t i = 0; i < k + nout; ++i) c.ptr[i] = ptrs[i];
    return hip_status(launch_gf16_apply_inline(c, stream));
  }
  PtrTable table;
  int st = table.upload(ptrs, stream);
  if (st) return st;
  Gf16Call c{};
  c.ptrs = static_cast<const uint8_t *const *>(table.d);
  c.coef = coef.ptr();
  c.plane = plane;
  c.col0 = 0;
  c.cols = plane;
  c.k = uint32_t(k);
  c.nout = uint32_t(nout);
  c.ostride = coef.ostride;
  c.vec_ok = al ? 1u : 0u;
  c.mode = mode;
  c.flag = flag;
  return hip_status(launch_gf16_apply(c, stream));
}

int read_flag(uint32_t *flag, hipStream_t stream, int *ok) {
  // The flag lands in a pinned word per host thread (a copy into pageable
  // memory goes through the runtime's staging); kept for the life of the
  // process, like the device flag.
  thread_local void *h = nullptr;
  if (!h) TRS_TRY(hipHostMalloc(&h, sizeof(uint32_t), hipHostMallocDefault));
  TRS_TRY(hipMemcpyAsync(h, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  TRS_TRY(hipStreamSynchronize(stream));
  *ok = *static_cast<volatile uint32_t *>(h) ? 0 : 1;
  return TRS_OK;
}

// ---- kernel variants and the learned router ---------------------------------
//
// One apply of k inputs to nout outputs over `len` bytes per shard can run
// on several kernels, all byte-identical:
//   kVarTable       packed-byte table tile (4 KiB chunks; 4/8/16 rows by nout)
//   kVarPlanes4/8   bit-sliced planes tile, 4 or 8 output rows per workgroup
//   kVarMfma        int8 MFMA on the bit-expanded matrix (GF(2^8), k <= 32)
//   kVarProgram*    the shape's FFT program (encode only): the generated
//                   bit-sliced program where it has one (one variant), else
//                   the staged program on the LDS interpreter sized for 16,
//                   32 or 64 KiB of LDS per workgroup (launches that come out
//                   the same are one candidate).
// The codec's rules (measured thresholds) pick one.  With a trained cost
// model installed (trs_tuner_install), an automatic call whose (field, op,
// k, nout, length bucket) has no cached choice ranks its candidates with the
// GPU MLP, times the top few on the call's own buffers (each run writes the
// same bytes) and caches the fastest.
enum Variant : int {
  kVarTable = 0,
  kVarPlanes4 = 1,
  kVarPlanes8 = 2,
  kVarMfma = 3,
  kVarProgram16K = 4,
  kVarProgram32K = 5,
  kVarProgram64K = 6,
  kVariants = 7
};

// Where the tower's int8-MFMA byte problem beats the Karatsuba kernel on a
// single call (per-call sweep over k in {2, 4, 8, 10, 16}, outputs in {1, 2,
// 4, 8, 16, 40}, 4 KiB to 8 MiB shards).  Both carry their shard pointers in
// the kernel arguments; MFMA re-reads the inputs once per 4 outputs, the
// Karatsuba kernel streams them once per 4 or 8 outputs at a higher rate, so
// MFMA wins short shards (except the tiniest shapes, a tie) and, up to 2 MiB,
// few outputs over enough inputs.
bool mfma16_rule(size_t k, size_t nout, size_t len) {
  if (len <= (size_t(256) << 10)) return k >= 4 || nout >= 8;
  if (len <= (size_t(1) << 20)) return k >= 4 && nout <= 8;
  if (len < (size_t(2) << 20)) return k >= 8 && nout <= 4;
  return false;
}

// The automatic tower encode.  Past 16 data shards, the FFT program wherever
// it fits (it won 1767 of 1774 sampled tower shapes).  For k <= 16 the same
// sweep puts the program ahead only for power-of-two k with enough parity
// (short shards: k/4 <= m < 2k; from 512 KiB: k >= 8 and 8 <= m <= 2.5k;
// from 4 MiB: k = 16 and m > 8); elsewhere the matrix kernel of mfma16_rule
// wins, by up to 5x at long shards.  Mean regret against the sweep's best:
// 1 % for this rule, 114 % for the program everywhere.
int enc16_rule_variant(size_t k, size_t m, size_t len, bool program, bool frag) {
  const int mat = frag && mfma16_rule(k, m, len) ? kVarMfma : kVarTable;
  if (!program) return mat;
  if (k > 16) return kVarProgram64K;
  const bool pow2 = (k & (k - 1)) == 0;
  if (len >= (size_t(4) << 20)) return k <= 10 || m <= 8 ? kVarTable : kVarProgram64K;
  if (len >= (size_t(512) << 10))
    return pow2 && k >= 8 && m >= 8 && 2 * m <= 5 * k ? kVarProgram64K : mat;
  if (mat != kVarMfma || (pow2 && 4 * m >= k && m < 2 * k)) return kVarProgram64K;
  return mat;
}

// Tower matrix applies: kVarTable (the Karatsuba kernel) or kVarMfma (the
// byte problem on int8 MFMA, where fragments exist).  route as
// trs_rs16_set_apply_route: 1 and 2 force a kernel, 0 follows the rule.
int apply16_rule(int route, size_t k, size_t nout, size_t len, bool frag) {
  if (!frag || route == 1) return kVarTable;
  if (route == 2) return kVarMfma;
  return mfma16_rule(k, nout, len) ? kVarMfma : kVarTable;
}
constexpr size_t kTuneFeatures = 12;

uint64_t variant_lds(int v) {
  return v == kVarProgram16K ? 16384 : v == kVarProgram32K ? 32768 : v == kVarProgram64K ? 65536 : 0;
}
bool is_program(int v) { return v >= kVarProgram16K && v <= kVarProgram64K; }

// The model's inputs: shape and length (log2), which kernel, its tile
// parameter (log2 rows per workgroup, or LDS KiB) and the data fraction.
void tune_features(int field, size_t k, size_t nout, size_t len, int v, float *f) {
  const double L = double(std::max<size_t>(len, 1)), K = double(k), N = double(nout);
  double tile = 0;
  if (v == kVarTable) tile = field == 16 ? (nout <= 4 ? 2 : 3) : (nout <= 4 ? 2 : nout <= 8 ? 3 : 4);
  if (v == kVarPlanes4) tile = 2;
  if (v == kVarPlanes8 || v == kVarMfma) tile = 3;
  if (is_program(v)) tile = std::log2(double(variant_lds(v) >> 10));
  f[0] = field == 16 ? 1.f : 0.f;
  f[1] = float(std::log2(K));
  f[2] = float(std::log2(N));
  f[3] = float(std::log2(L));
  f[4] = float(std::log2(K * N));
  f[5] = float(std::log2((K + N) * L));
  f[6] = v == kVarTable ? 1.f : 0.f;
  f[7] = (v == kVarPlanes4 || v == kVarPlanes8) ? 1.f : 0.f;
  f[8] = v == kVarMfma ? 1.f : 0.f;
  f[9] = is_program(v) ? 1.f : 0.f;
  f[10] = float(tile);
  f[11] = float(K / (K + N));
}

// Rows per planes tile the launch rule picks (launch_gf8_planes with nt = 0).
int planes_rule_variant(size_t len, size_t nout) {
  const uint64_t gx = (len + kPlanesChunkBytes - 1) / kPlanesChunkBytes;
  const uint64_t wg8 = gx * ((nout + kPlanesOut - 1) / kPlanesOut);
  return nout <= 4 || wg8 <= 640 ? kVarPlanes4 : kVarPlanes8;
}

// Feasible runtime-coefficient variants of a GF(2^8) apply.
uint32_t matrix_variants8(size_t k, size_t nout) {
  uint32_t mask = (1u << kVarTable) | (1u << kVarPlanes4);
  if (nout > 4) mask |= 1u << kVarPlanes8;  // at most 4 rows: the same work as 4-row tiles
  if (mfma_available(uint32_t(k))) mask |= 1u << kVarMfma;
  return mask;
}

// The distinct LDS-target variants of an interpreter program over `cols`.
uint32_t program_variants(const Fft16Device &d, uint64_t cols, bool byte_planes) {
  uint32_t mask = 0;
  uint32_t seen[3] = {~0u, ~0u, ~0u};
  for (int v = kVarProgram16K, i = 0; v <= kVarProgram64K; ++v, ++i) {
    const uint32_t gs = fft16_launch_gshift(d, cols, byte_planes, variant_lds(v));
    bool dup = false;
    for (int j = 0; j < i; ++j) dup = dup || seen[j] == gs;
    seen[i] = gs;
    if (!dup) mask |= 1u << v;
  }
  return mask;
}

struct RouteKey {
  int field, op;  // op: 0 encode, 1 reconstruct apply
  size_t k, nout;
  int bucket;     // round(2 log2 len): half-octave length buckets
  bool operator<(const RouteKey &o) const {
    return std::tie(field, op, k, nout, bucket) < std::tie(o.field, o.op, o.k, o.nout, o.bucket);
  }
};

int length_bucket(size_t len) {
  return int(std::lround(2.0 * std::log2(double(std::max<size_t>(len, 1)))));
}

struct TunerModel {
  trs_mlp *mlp = nullptr;
  float mu[kTuneFeatures], sd[kTuneFeatures];
  int top_n = 3;
  ~TunerModel() {
    if (mlp) trs_mlp_free(mlp);
  }
};

struct Router {
  std::shared_mutex mu;
  std::shared_ptr<TunerModel> model;
  std::map<RouteKey, int> cache;
};

Router &router() {
  static Router *r = new Router();  // kept for the life of the process
  return *r;
}

// Runs one automatic apply: run(v) launches variant v on the call's stream.
// Without a model: the rule's variant.  With one: the cached choice for the
// key, or -- on a miss -- the model's top candidates timed on this call (a
// warm-up run, then a timed run each), the fastest cached.  Every run writes
// the same bytes, so the call's outputs are complete whichever ran last.
template <typename Feasible, typename Run>
int router_run(const RouteKey &key, size_t len, int rule, hipStream_t s, Feasible &&feasible_fn,
               Run &&run) {
  Router &R = router();
  std::shared_ptr<TunerModel> model;
  {
    std::shared_lock<std::shared_mutex> lk(R.mu);
    model = R.model;
    if (!model) return run(rule);
    auto it = R.cache.find(key);
    if (it != R.cache.end()) return run(it->second);
  }
  const uint32_t feasible = feasible_fn() | (1u << rule);
  std::vector<int> cand;
  for (int v = 0; v < kVariants; ++v)
    if ((feasible >> v) & 1u) cand.push_back(v);
  if (cand.size() > 1) {
    std::vector<float> X(cand.size() * kTuneFeatures), Y(cand.size());
    for (size_t c = 0; c < cand.size(); ++c) {
      float *f = X.data() + c * kTuneFeatures;
      tune_features(key.field, key.k, key.nout, len, cand[c], f);
      for (size_t i = 0; i < kTuneFeatures; ++i) f[i] = (f[i] - model->mu[i]) / model->sd[i];
    }
    if (trs_mlp_predict(model->mlp, X.data(), cand.size(), Y.data()) != TRS_OK) return run(rule);
    std::vector<size_t> order(cand.size());
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return Y[a] < Y[b]; });
    std::vector<int> ranked;
    for (size_t i : order) ranked.push_back(cand[i]);
    cand.swap(ranked);
    if (cand.size() > size_t(model->top_n)) cand.resize(size_t(model->top_n));
  }
  thread_local hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!e0) TRS_TRY(hipEventCreate(&e0));
  if (!e1) TRS_TRY(hipEventCreate(&e1));
  int best = -1;
  float best_ms = 0;
  for (int v : cand) {
    if (run(v) != TRS_OK) continue;  // warm-up (and first write of the outputs)
    TRS_TRY(hipEventRecord(e0, s));
    if (run(v) != TRS_OK) continue;
    TRS_TRY(hipEventRecord(e1, s));
    TRS_TRY(hipEventSynchronize(e1));
    float ms = 0;
    TRS_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (best < 0 || ms < best_ms) {
      best = v;
      best_ms = ms;
    }
  }
  if (best < 0) return run(rule);
  std::unique_lock<std::shared_mutex> lk(R.mu);
  if (R.model == model) R.cache[key] = best;
  return TRS_OK;
}

// verify over a validated shard list (every shard read): launch(ptrs, len,
// flag, stream) sets *flag on a mismatch; *ok = 1 when nothing differed.
template <typename Launch>
int verify_dispatch(const Exec &ex, const std::vector<const uint8_t *> &ptrs, size_t len, int *ok,
                    Launch &&launch) {
  uint32_t *flag = nullptr;
  int st = verify_flag(&flag);
  if (st) return st;
  bool first = true;
  hipStream_t last = ex.stream;
  st = dispatch(ex, ptrs, ptrs.size(), len,
                [&](const std::vector<const uint8_t *> &p, size_t l, hipStream_t s) -> int {
                  if (first) TRS_TRY(hipMemsetAsync(flag, 0, sizeof(uint32_t), s));
                  first = false;
                  last = s;
                  return launch(p, l, flag, s);
                });
  if (st) return st;
  return read_flag(flag, last, ok);
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================

extern "C" {

const char *trs_status_message(int s) {
  switch (s) {
    case TRS_OK: return "ok";
    case TRS_TOO_FEW_SHARDS: return "The number of provided shards is smaller than the one in codec";
    case TRS_TOO_MANY_SHARDS: return "The number of provided shards is greater than the one in codec";
    case TRS_TOO_FEW_DATA_SHARDS: return "The number of provided data shards is smaller than the one in codec";
    case TRS_TOO_MANY_DATA_SHARDS: return "The number of provided data shards is greater than the one in codec";
    case TRS_TOO_FEW_PARITY_SHARDS: return "The number of provided parity shards is smaller than the one in codec";
    case TRS_TOO_MANY_PARITY_SHARDS: return "The number of provided parity shards is greater than the one in codec";
    case TRS_TOO_FEW_BUFFER_SHARDS: return "The number of provided buffer shards is smaller than the number of parity shards in codec";
    case TRS_TOO_MANY_BUFFER_SHARDS: return "The number of provided buffer shards is greater than the number of parity shards in codec";
    case TRS_INCORRECT_SHARD_SIZE: return "At least one of the provided shards is not of the correct size";
    case TRS_TOO_FEW_SHARDS_PRESENT: return "The number of shards present is smaller than number of parity shards, cannot reconstruct missing shards";
    case TRS_EMPTY_SHARD: return "The first shard provided is of zero length";
    case TRS_INVALID_SHARD_FLAGS: return "The number of flags does not match the total number of shards";
    case TRS_INVALID_INDEX: return "The data shard index provided is greater or equal to the number of data shards in codec";
    case TRS_SINGULAR_MATRIX: return "The decode matrix for this shard pattern is singular and cannot be inverted";
    case TRS_ERR_DEVICE: return "HIP runtime error";
    case TRS_ERR_OUT_OF_MEMORY: return "device out of memory";
    case TRS_ERR_INVALID_ARGUMENT: return "invalid argument";
    case TRS_ERR_UNSUPPORTED: return "unsupported";
    case TRS_ERR_SCRATCH_LIMIT:
      return "request exceeds the allocation-free entry point's scratch (1 MiB of shard pointers)";
    default: return "unknown status";
  }
}

const char *trs_version(void) { return "trs_gpu 0.1 (gfx950)"; }

int trs_gf8_mul_slice(uint8_t *out, size_t out_len, const uint8_t *in, size_t in_len, uint8_t c,
                      int xor_into, void *stream) {
  if (out_len != in_len) return TRS_ERR_INVALID_ARGUMENT;
  if (out_len == 0) return TRS_OK;
  if (!out || !in) return TRS_ERR_INVALID_ARGUMENT;
  const uint32_t *coef = nullptr;
  int st = slice_coef(c, &coef);
  if (st) return st;
  Gf8Call call{};
  call.ptr[0] = in;
  call.ptr[1] = out;
  call.coef = coef;
  call.len = out_len;
  call.k = 1;
  call.nout = 1;
  call.ostride = kOutStride;
  call.vec_ok = aligned16(in) && aligned16(out);
  return hip_status(launch_gf8_apply(call, xor_into ? kXorInto : kWrite, as_stream(stream)));
}

int trs_fill_random(uint8_t *dst, size_t len, uint64_t seed, void *stream) {
  if (len && !dst) return TRS_ERR_INVALID_ARGUMENT;
  return hip_status(launch_fill(dst, len, seed, as_stream(stream)));
}

// ---------------------------------------------------------------- RS8

namespace {
// Longest shard length at which the automatic encode of (k, m) runs the
// staged FFT program (0: never).  The reference takes power-of-two k whose
// program has fewer than k*m/2 multiplies and fits 160 registers
// (src/reedsolomon.rs:204-234); its register limit does not apply here (the
// file lives in LDS).  Measured on this GPU at full size against the
// automatic matrix route (51 shapes, 16 KiB - 4 MiB): for k > 32 the matrix
// kernels leave most of the chip idle up to a few hundred KiB (the MFMA route
// covers k <= 32) and the program is 1.1-4x faster through 256 KiB.  Beyond,
// the ratio of program multiplies to k*m decides: the program stays ahead at
// any length below 0.13 for k = 64, 128 (few, wide steps) and below 0.078
// for other k (more glue, more steps); k = 64, 128 above that stay ahead up
// to 1 MiB.  Generated shapes keep their bit-sliced programs.  Env
// TRS_STAGED8=0 turns the route off.
size_t staged8_gpu_max(size_t k, size_t m) {
  static const bool off = [] {
    const char *e = std::getenv("TRS_STAGED8");
    return e && e[0] == '0';
  }();
  if (off || k <= 32 || k + m > 256) return 0;
  const Fft16Program p = build_staged_program8(k, m);
  if (!fft16_fits_device(p)) return 0;
  const bool pow2 = (k & (k - 1)) == 0;
  if ((pow2 && p.mult_count * 100 < 13 * k * m) || p.mult_count * 1000 < 78 * k * m)
    return SIZE_MAX;
  return pow2 ? size_t(1) << 20 : size_t(256) << 10;
}

// The encode route of an automatic call at this length.
int auto_route8(const trs_rs8 *rs, size_t len) {
  if (rs->route == kRouteTable && len <= rs->staged_max) return kRouteStaged8;
  return rs->route;
}

}  // namespace

int trs_rs8_new(size_t k, size_t m, trs_rs8 **out) {
  if (!out) return TRS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  // src/reedsolomon.rs:167-175
  if (k == 0) return TRS_TOO_FEW_DATA_SHARDS;
  if (m == 0) return TRS_TOO_FEW_PARITY_SHARDS;
  if (k + m > 256) return TRS_TOO_MANY_SHARDS;
  auto rs = std::make_unique<trs_rs8>();
  rs->k = k;
  rs->m = m;
  rs->n = k + m;
  if (!build_generator<F8>(k, k + m, rs->G)) return TRS_SINGULAR_MATRIX;
  int st = upload_coef8(rs->parity, rs->G.row(k), m, k);
  if (st) return st;
  if ((st = upload_frags(rs->parity_frag, rs->G.row(k), m, k))) return st;
  if (bs_encode_available(kRouteBsFft, uint32_t(k), uint32_t(m)))
    rs->route = kRouteBsFft;
  else if (bs_encode_available(kRouteBsMatrix, uint32_t(k), uint32_t(m)))
    rs->route = kRouteBsMatrix;
  if (rs->route == kRouteTable) rs->staged_max = staged8_gpu_max(k, m);
  *out = rs.release();
  return TRS_OK;
}

void trs_rs8_free(trs_rs8 *rs) { delete rs; }
size_t trs_rs8_data_shard_count(const trs_rs8 *rs) { return rs ? rs->k : 0; }
size_t trs_rs8_parity_shard_count(const trs_rs8 *rs) { return rs ? rs->m : 0; }
size_t trs_rs8_total_shard_count(const trs_rs8 *rs) { return rs ? rs->n : 0; }

const char *trs_rs8_encode_route(const trs_rs8 *rs, size_t shard_len) {
  if (!rs) return "none";
  switch (auto_route8(rs, shard_len)) {
    case kRouteStaged8: return "gpu-staged-fft";
    case kRouteBsFft: return "gpu-bitsliced-fft";
    case kRouteBsMatrix: return "gpu-bitsliced-matrix";
    default:
      if (use_mfma(rs->apply_route.load(), rs->k, shard_len, false) && rs->parity_frag.p)
        return "gpu-int8-mfma-matrix";
      return shard_len >= planes_min_len() ? "gpu-bitsliced-planes-matrix" : "gpu-fused-matrix";
  }
}

const char *trs_rs8_reference_route(const trs_rs8 *rs, size_t shard_len) {
  if (!rs) return "none";
  // src/reedsolomon.rs:239-253 with src/gf/fft_x86.rs:239-250 (GFNI host,
  // 32-byte strips)
  if (shard_len >= 32) {
    if (bs_encode_available(kRouteBsFft, uint32_t(rs->k), uint32_t(rs->m))) return "fft-generated";
    if (staged8_reference_route(rs->k, rs->m)) return "fft-staged";
  }
  return "fused-matrix";
}

int trs_rs8_generator(const trs_rs8 *rs, uint8_t *out, size_t out_len) {
  if (!rs || !out || out_len != rs->n * rs->k) return TRS_ERR_INVALID_ARGUMENT;
  std::copy(rs->G.a.begin(), rs->G.a.end(), out);
  return TRS_OK;
}

namespace {
// The codec's staged program on the device, built and uploaded once.
int staged8_device(const trs_rs8 *rs, const Fft16Device **out) {
  std::call_once(rs->staged_once, [rs] {
    rs->staged_status = fft16_device_new(build_staged_program8(rs->k, rs->m), &rs->staged_dev);
  });
  if (rs->staged_status == hipErrorInvalidValue) return TRS_ERR_UNSUPPORTED;
  if (rs->staged_status != hipSuccess) return hip_status(rs->staged_status);
  *out = rs->staged_dev.get();
  return TRS_OK;
}

int run_staged8(const trs_rs8 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                hipStream_t stream) {
  const Fft16Device *dp = nullptr;
  const int dst = staged8_device(rs, &dp);
  if (dst) return dst;
  const Fft16Device &d = *dp;
  if (fft16_device_loads(d) + fft16_device_stores(d) != ptrs.size()) return TRS_ERR_UNSUPPORTED;
  if (len == 0) return TRS_OK;
  return launch_fft16_one(d, ptrs, fft16_device_stores(d), len, stream, true);
}

// check_piece_count_all + check_equal_lengths (src/reedsolomon.rs:701-722).
int check_all8(size_t total, const size_t *lens, size_t n) {
  if (n < total) return TRS_TOO_FEW_SHARDS;
  if (n > total) return TRS_TOO_MANY_SHARDS;
  const size_t len = lens[0];
  if (len == 0) return TRS_EMPTY_SHARD;
  for (size_t i = 0; i < n; ++i)
    if (lens[i] != len) return TRS_INCORRECT_SHARD_SIZE;
  return TRS_OK;
}

int encode8_ptrs(const trs_rs8 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                 hipStream_t s, int route) {
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  if (route == kRouteStaged8) return run_staged8(rs, ptrs, len, s);
  if (route == kRouteTable || len == 0) {
    if (use_mfma(rs->apply_route.load(), rs->k, len, false) && rs->parity_frag.p)
      return run_mfma(ptrs, rs->k, rs->m, rs->parity_frag.p, len, s);
    return run_gf8(ptrs, rs->k, rs->m, rs->parity, len, kWrite, nullptr, s);
  }
  if (!bs_encode_available(route, uint32_t(rs->k), uint32_t(rs->m))) return TRS_ERR_UNSUPPORTED;
  Gf8Call c{};
  bool al = true;
  for (size_t i = 0; i < ptrs.size(); ++i) {
    c.ptr[i] = ptrs[i];
    al = al && aligned16(ptrs[i]);
  }
  c.len = len;
  c.k = uint32_t(rs->k);
  c.nout = uint32_t(rs->m);
  c.vec_ok = al ? 1u : 0u;
  return hip_status(launch_bs_encode(route, c, s));
}

// The kernel variant the codec's rules take for an encode of this length:
// the generated program, or the staged program at its default LDS target
// (32 KiB of byte planes), or the matrix kernels by length.
int enc8_rule_variant(const trs_rs8 *rs, size_t len) {
  if (auto_route8(rs, len) != kRouteTable) return kVarProgram32K;
  if (use_mfma(rs->apply_route.load(), rs->k, len, false) && rs->parity_frag.p) return kVarMfma;
  return len >= planes_min_len() ? planes_rule_variant(len, rs->m) : kVarTable;
}

// Every kernel an encode of this codec can run at this length.
uint32_t enc8_variants(const trs_rs8 *rs, size_t len) {
  uint32_t mask = matrix_variants8(rs->k, rs->m);
  if (!rs->parity_frag.p) mask &= ~(1u << kVarMfma);
  if (rs->route != kRouteTable) return mask | (1u << kVarProgram32K);
  const Fft16Device *d = nullptr;
  if (staged8_device(rs, &d) == TRS_OK &&
      fft16_device_loads(*d) + fft16_device_stores(*d) == rs->k + rs->m)
    mask |= program_variants(*d, len, true);
  return mask;
}

int run_enc8_variant(const trs_rs8 *rs, int v, const std::vector<const uint8_t *> &ptrs,
                     size_t len, hipStream_t s) {
  if (v == kVarMfma) {
    if (!rs->parity_frag.p) return TRS_ERR_UNSUPPORTED;
    return run_mfma(ptrs, rs->k, rs->m, rs->parity_frag.p, len, s);
  }
  if (!is_program(v)) return run_gf8(ptrs, rs->k, rs->m, rs->parity, len, kWrite, nullptr, s, v);
  if (rs->route != kRouteTable) return encode8_ptrs(rs, ptrs, len, s, rs->route);
  const Fft16Device *d = nullptr;
  const int st = staged8_device(rs, &d);
  if (st) return st;
  if (fft16_device_loads(*d) + fft16_device_stores(*d) != ptrs.size()) return TRS_ERR_UNSUPPORTED;
  return launch_fft16_one(*d, ptrs, fft16_device_stores(*d), len, s, true, nullptr,
                          variant_lds(v));
}

// The automatic encode of a validated shard list (k data then m parity).
int encode8_auto(const trs_rs8 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                 const Exec &ex) {
  return dispatch(ex, ptrs, rs->k, len,
                  [rs](const std::vector<const uint8_t *> &p, size_t l, hipStream_t s) {
                    for (auto q : p)
                      if (!q) return int(TRS_ERR_INVALID_ARGUMENT);
                    // a forced apply route (trs_rs8_set_apply_route) is not tuned
                    if (rs->apply_route.load() != 0)
                      return run_enc8_variant(rs, enc8_rule_variant(rs, l), p, l, s);
                    return router_run(
                        RouteKey{8, 0, rs->k, rs->m, length_bucket(l)}, l, enc8_rule_variant(rs, l),
                        s, [&] { return enc8_variants(rs, l); },
                        [&](int v) { return run_enc8_variant(rs, v, p, l, s); });
                  });
}

// verify of k data + m parity shards through the codec's encode route: the
// generated bit-sliced program or the staged FFT program compare their
// outputs with the parity shards instead of writing them; the matrix route
// runs the apply kernel's verify mode.  Reads every shard once.
int verify8_launch(const trs_rs8 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                   uint32_t *flag, hipStream_t s) {
  const int route = auto_route8(rs, len);
  if (route == kRouteStaged8) {
    const Fft16Device *d = nullptr;
    const int st = staged8_device(rs, &d);
    if (st) return st;
    if (fft16_device_loads(*d) + fft16_device_stores(*d) != ptrs.size())
      return TRS_ERR_UNSUPPORTED;
    return launch_fft16_one(*d, ptrs, fft16_device_stores(*d), len, s, true, flag);
  }
  if (route == kRouteBsFft || route == kRouteBsMatrix) {
    Gf8Call c{};
    bool al = true;
    for (size_t i = 0; i < ptrs.size(); ++i) {
      c.ptr[i] = ptrs[i];
      al = al && aligned16(ptrs[i]);
    }
    c.len = len;
    c.k = uint32_t(rs->k);
    c.nout = uint32_t(rs->m);
    c.vec_ok = al ? 1u : 0u;
    c.flag = flag;
    return hip_status(launch_bs_verify(route, c, s));
  }
  return run_gf8(ptrs, rs->k, rs->m, rs->parity, len, kVerify, flag, s);
}

// ReedSolomon::encode (src/reedsolomon.rs:271-279) over device or host shards.
int rs8_encode(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens, size_t n,
               const Exec &ex) {
  if (!rs || (n && (!shards || !lens))) return TRS_ERR_INVALID_ARGUMENT;
  if (n == 0) return TRS_TOO_FEW_SHARDS;
  int st = check_all8(rs->n, lens, n);
  if (st) return st;
  std::vector<const uint8_t *> ptrs(shards, shards + n);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return encode8_auto(rs, ptrs, lens[0], ex);
}

// The checks of encode_sep (src/reedsolomon.rs:289-308); fills ptrs and len.
int check_sep8(const trs_rs8 *rs, const uint8_t *const *data, const size_t *dl, size_t nd,
               uint8_t *const *parity, const size_t *pl, size_t np,
               std::vector<const uint8_t *> &ptrs, size_t &len) {
  if (!rs || (nd && (!data || !dl)) || (np && (!parity || !pl))) return TRS_ERR_INVALID_ARGUMENT;
  if (nd != rs->k) return nd < rs->k ? TRS_TOO_FEW_DATA_SHARDS : TRS_TOO_MANY_DATA_SHARDS;
  if (np != rs->m) return np < rs->m ? TRS_TOO_FEW_PARITY_SHARDS : TRS_TOO_MANY_PARITY_SHARDS;
  len = dl[0];
  for (size_t i = 0; i < nd; ++i)
    if (dl[i] != len) return TRS_INCORRECT_SHARD_SIZE;
  for (size_t i = 0; i < np; ++i)
    if (pl[i] != len) return TRS_INCORRECT_SHARD_SIZE;
  ptrs.assign(data, data + nd);
  ptrs.insert(ptrs.end(), parity, parity + np);
  if (len)
    for (auto q : ptrs)
      if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return TRS_OK;
}

int rs8_encode_sep(const trs_rs8 *rs, const uint8_t *const *data, const size_t *dl, size_t nd,
                   uint8_t *const *parity, const size_t *pl, size_t np, const Exec &ex) {
  std::vector<const uint8_t *> ptrs;
  size_t len = 0;
  int st = check_sep8(rs, data, dl, nd, parity, pl, np, ptrs, len);
  if (st || len == 0) return st;
  return encode8_auto(rs, ptrs, len, ex);
}

// ReedSolomon::encode_rows_into, src/reedsolomon.rs:667-682.
int rs8_encode_rows_into(const trs_rs8 *rs, const uint8_t *data, size_t data_len, size_t row_len,
                         uint8_t *parity, size_t parity_len, const Exec &ex) {
  if (!rs) return TRS_ERR_INVALID_ARGUMENT;
  // src/reedsolomon.rs:673-680
  if (data_len != rs->k * row_len) return TRS_INCORRECT_SHARD_SIZE;
  if (parity_len != rs->m * row_len) return TRS_INCORRECT_SHARD_SIZE;
  if (row_len == 0) return TRS_OK;
  if (!data || !parity) return TRS_ERR_INVALID_ARGUMENT;
  std::vector<const uint8_t *> ptrs;
  for (size_t i = 0; i < rs->k; ++i) ptrs.push_back(data + i * row_len);
  for (size_t i = 0; i < rs->m; ++i) ptrs.push_back(parity + i * row_len);
  // The reference's rows API always takes its matrix route
  // (src/reedsolomon.rs:684-697); every route here is byte-identical, so the
  // rows take the codec's fastest one.
  return encode8_auto(rs, ptrs, row_len, ex);
}

// ReedSolomon::verify, src/reedsolomon.rs:430-463.
int rs8_verify(const trs_rs8 *rs, const uint8_t *const *shards, const size_t *lens, size_t n,
               int *ok, const Exec &ex) {
  if (!rs || !ok || (n && (!shards || !lens))) return TRS_ERR_INVALID_ARGUMENT;
  if (n == 0) return TRS_TOO_FEW_SHARDS;
  int st = check_all8(rs->n, lens, n);
  if (st) return st;
  std::vector<const uint8_t *> ptrs(shards, shards + n);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return verify_dispatch(ex, ptrs, lens[0], ok,
                         [rs](const std::vector<const uint8_t *> &p, size_t l, uint32_t *flag,
                              hipStream_t s) { return verify8_launch(rs, p, l, flag, s); });
}

// The apply of a reconstruct: ptrs = the plan's k survivors, then its nout
// outputs.
int rec8_launch(const Plan &plan, size_t k, size_t total, size_t nout,
                const std::vector<const uint8_t *> &ptrs, size_t len, hipStream_t s,
                int apply_route, int decode_route) {
  if (nout && use_fft_decode8(decode_route, total, k, nout, len)) {
    const Fft16Device *d = nullptr;
    const int st = plan_fft16(const_cast<Plan &>(plan), total, &d, true);
    if (st) return st;
    return launch_fft16_one(*d, ptrs, nout, len, s, true);
  }
  // the decode plan's matrix on the runtime-coefficient kernels
  const bool mfma = use_mfma(apply_route, k, len, false) && plan.frag.p;
  const int rule = mfma ? kVarMfma : len >= planes_min_len() ? planes_rule_variant(len, nout)
                                                             : kVarTable;
  auto run = [&](int v) {
    if (v == kVarMfma) {
      if (!plan.frag.p) return int(TRS_ERR_UNSUPPORTED);
      return run_mfma(ptrs, k, nout, plan.frag.p, len, s);
    }
    return run_gf8(ptrs, k, nout, plan.coef, len, kWrite, nullptr, s, v);
  };
  // a forced apply route (trs_rs8_set_apply_route) is not tuned
  if (apply_route != 0) return run(rule);
  return router_run(RouteKey{8, 1, k, nout, length_bucket(len)}, len, rule, s,
                    [&] {
                      uint32_t mask = matrix_variants8(k, nout);
                      if (!plan.frag.p) mask &= ~(1u << kVarMfma);
                      return mask;
                    },
                    run);
}

// reconstruct_internal / PreparedDecoder::reconstruct_internal
// (src/reedsolomon.rs:615-641, :771-796).  `plan` non-null = prepared.
int reconstruct8(const trs_rs8 *rs, size_t k, size_t total, std::shared_ptr<Plan> plan,
                 uint8_t *const *shards, const size_t *lens, const uint8_t *present, size_t n,
                 int data_only, const Exec &ex, int apply_route, int dec_decode_route = 0) {
  if (n && (!shards || !lens || !present)) return TRS_ERR_INVALID_ARGUMENT;
  if (n != total) return n < total ? TRS_TOO_FEW_SHARDS : TRS_TOO_MANY_SHARDS;
  Mask mask;
  size_t count = 0, len = 0;
  int st = scan_shards(lens, present, n, false, mask, count, len);
  if (st) return st;
  if (plan) {
    if (mask != plan->presence) return TRS_INVALID_SHARD_FLAGS;
    if (count == total) return TRS_OK;
    if (len == 0) return TRS_TOO_FEW_SHARDS_PRESENT;
  } else {
    if (count == total) return TRS_OK;
    if (count < k) return TRS_TOO_FEW_SHARDS_PRESENT;
    if ((st = plan8_for(rs, mask, &plan))) return st;
  }
  std::vector<const uint8_t *> ptrs;
  size_t nout = 0;
  if ((st = gather(*plan, shards, lens, k, len, data_only != 0, ptrs, nout))) return st;
  if (nout == 0) return TRS_OK;
  const int decode_route = rs ? rs->decode_route.load() : dec_decode_route;
  const Plan &p = *plan;
  return dispatch(ex, ptrs, k, len,
                  [&](const std::vector<const uint8_t *> &q, size_t l, hipStream_t s) {
                    return rec8_launch(p, k, total, nout, q, l, s, apply_route, decode_route);
                  });
}

// ReedSolomon::reconstruct_rows, src/reedsolomon.rs:514-535.
int rs8_reconstruct_rows(const trs_rs8 *rs, uint8_t *rows, size_t rows_len, size_t row_len,
                         const uint8_t *present, size_t n_present, const Exec &ex) {
  if (!rs || (n_present && !present)) return TRS_ERR_INVALID_ARGUMENT;
  // src/reedsolomon.rs:520-528
  if (n_present != rs->n) return TRS_INVALID_SHARD_FLAGS;
  if (row_len == 0) return TRS_EMPTY_SHARD;
  if (rows_len != rs->n * row_len) return TRS_INCORRECT_SHARD_SIZE;
  if (!rows) return TRS_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t *> shards(rs->n);
  std::vector<size_t> lens(rs->n, row_len);
  for (size_t i = 0; i < rs->n; ++i) shards[i] = rows + i * row_len;
  return reconstruct8(rs, rs->k, rs->n, nullptr, shards.data(), lens.data(), present, rs->n, 0,
                      ex, rs->apply_route.load());
}
}  // namespace

int trs_rs8_encode(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens, size_t n,
                   void *stream) {
  return rs8_encode(rs, shards, lens, n, as_stream(stream));
}

int trs_rs8_encode_host(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens, size_t n) {
  return rs8_encode(rs, shards, lens, n, Exec::on_host(1));
}

int trs_rs8_encode_sep(const trs_rs8 *rs, const uint8_t *const *data, const size_t *dl,
                       size_t nd, uint8_t *const *parity, const size_t *pl, size_t np,
                       void *stream) {
  return rs8_encode_sep(rs, data, dl, nd, parity, pl, np, as_stream(stream));
}

int trs_rs8_encode_sep_host(const trs_rs8 *rs, const uint8_t *const *data, const size_t *dl,
                            size_t nd, uint8_t *const *parity, const size_t *pl, size_t np) {
  return rs8_encode_sep(rs, data, dl, nd, parity, pl, np, Exec::on_host(1));
}

int trs_rs8_encode_sep_route(const trs_rs8 *rs, int route, const uint8_t *const *data,
                             const size_t *dl, size_t nd, uint8_t *const *parity, const size_t *pl,
                             size_t np, void *stream) {
  if (!rs || (nd && (!data || !dl)) || (np && (!parity || !pl))) return TRS_ERR_INVALID_ARGUMENT;
  if (route < kRouteTable || route > kRouteStaged8) return TRS_ERR_INVALID_ARGUMENT;
  std::vector<const uint8_t *> ptrs;
  size_t len = 0;
  int st = check_sep8(rs, data, dl, nd, parity, pl, np, ptrs, len);
  if (st || len == 0) return st;
  return encode8_ptrs(rs, ptrs, len, as_stream(stream), route);
}

int trs_rs8_encode_rows_into(const trs_rs8 *rs, const uint8_t *data, size_t data_len,
                             size_t row_len, uint8_t *parity, size_t parity_len, void *stream) {
  return rs8_encode_rows_into(rs, data, data_len, row_len, parity, parity_len, as_stream(stream));
}

int trs_rs8_encode_rows_into_host(const trs_rs8 *rs, const uint8_t *data, size_t data_len,
                                  size_t row_len, uint8_t *parity, size_t parity_len) {
  return rs8_encode_rows_into(rs, data, data_len, row_len, parity, parity_len, Exec::on_host(1));
}

int trs_rs8_verify(const trs_rs8 *rs, const uint8_t *const *shards, const size_t *lens, size_t n,
                   int *ok, void *stream) {
  return rs8_verify(rs, shards, lens, n, ok, as_stream(stream));
}

int trs_rs8_verify_host(const trs_rs8 *rs, const uint8_t *const *shards, const size_t *lens,
                        size_t n, int *ok) {
  return rs8_verify(rs, shards, lens, n, ok, Exec::on_host(1));
}

int trs_rs8_set_decode_route(trs_rs8 *rs, int route) {
  if (!rs || route < 0 || route > 2) return TRS_ERR_INVALID_ARGUMENT;
  rs->decode_route.store(route);
  return TRS_OK;
}

int trs_rs8_set_apply_route(trs_rs8 *rs, int route) {
  if (!rs || route < 0 || route > 2) return TRS_ERR_INVALID_ARGUMENT;
  if (route == 2 && !mfma_available(uint32_t(rs->k))) return TRS_ERR_UNSUPPORTED;
  rs->apply_route.store(route);
  return TRS_OK;
}

int trs_rs8_reconstruct(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens,
                        const uint8_t *present, size_t n, int data_only, void *stream) {
  if (!rs) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct8(rs, rs->k, rs->n, nullptr, shards, lens, present, n, data_only,
                      as_stream(stream), rs->apply_route.load());
}

int trs_rs8_reconstruct_host(const trs_rs8 *rs, uint8_t *const *shards, const size_t *lens,
                             const uint8_t *present, size_t n, int data_only) {
  if (!rs) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct8(rs, rs->k, rs->n, nullptr, shards, lens, present, n, data_only,
                      Exec::on_host(1), rs->apply_route.load());
}

int trs_rs8_reconstruct_rows(const trs_rs8 *rs, uint8_t *rows, size_t rows_len, size_t row_len,
                             const uint8_t *present, size_t n_present, void *stream) {
  return rs8_reconstruct_rows(rs, rows, rows_len, row_len, present, n_present, as_stream(stream));
}

int trs_rs8_reconstruct_rows_host(const trs_rs8 *rs, uint8_t *rows, size_t rows_len,
                                  size_t row_len, const uint8_t *present, size_t n_present) {
  return rs8_reconstruct_rows(rs, rows, rows_len, row_len, present, n_present, Exec::on_host(1));
}

int trs_rs8_prepare_decode(const trs_rs8 *rs, const uint8_t *present, size_t n, trs_dec8 **out) {
  if (!rs || !out || (n && !present)) return TRS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  // src/reedsolomon.rs:485-500
  if (n != rs->n) return TRS_INVALID_SHARD_FLAGS;
  Mask m = make_mask(present, n);
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) count += present[i] != 0;
  if (count < rs->k) return TRS_TOO_FEW_SHARDS_PRESENT;
  std::shared_ptr<Plan> plan;
  int st = plan8_for(rs, m, &plan);
  if (st) return st;
  auto d = std::make_unique<trs_dec8>();
  d->k = rs->k;
  d->n = rs->n;
  d->apply_route = rs->apply_route.load();
  d->decode_route = rs->decode_route.load();
  d->plan = plan;
  *out = d.release();
  return TRS_OK;
}

int trs_dec8_reconstruct(const trs_dec8 *dec, uint8_t *const *shards, const size_t *lens,
                         const uint8_t *present, size_t n, int data_only, void *stream) {
  if (!dec) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct8(nullptr, dec->k, dec->n, dec->plan, shards, lens, present, n, data_only,
                      as_stream(stream), dec->apply_route, dec->decode_route);
}

int trs_dec8_reconstruct_host(const trs_dec8 *dec, uint8_t *const *shards, const size_t *lens,
                              const uint8_t *present, size_t n, int data_only) {
  if (!dec) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct8(nullptr, dec->k, dec->n, dec->plan, shards, lens, present, n, data_only,
                      Exec::on_host(1), dec->apply_route, dec->decode_route);
}

void trs_dec8_free(trs_dec8 *dec) { delete dec; }

// ---------------------------------------------------------------- RS16

namespace {
// The device form of the codec's program, uploaded once.
int fft16_device_for(const trs_rs16 *rs, const Fft16Device **out) {
  if (!rs->has_program) return TRS_ERR_UNSUPPORTED;
  std::call_once(rs->fft_once, [rs] { rs->fft_status = fft16_device_new(rs->program, &rs->fft); });
  if (rs->fft_status == hipErrorInvalidValue) return TRS_ERR_UNSUPPORTED;
  if (rs->fft_status != hipSuccess) return hip_status(rs->fft_status);
  *out = rs->fft.get();
  return TRS_OK;
}
}  // namespace

int trs_rs16_new(size_t k, size_t m, trs_rs16 **out) {
  if (!out) return TRS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  // src/reedsolomon16.rs:282-290
  if (k == 0) return TRS_TOO_FEW_DATA_SHARDS;
  if (m == 0) return TRS_TOO_FEW_PARITY_SHARDS;
  if (k + m > 65536) return TRS_TOO_MANY_SHARDS;
  auto rs = std::make_unique<trs_rs16>();
  rs->k = k;
  rs->m = m;
  rs->n = k + m;
  if (!build_generator<F16>(k, k + m, rs->G)) return TRS_SINGULAR_MATRIX;
  int st = upload_coef16(rs->parity, rs->G.row(k), m, k);
  if (st) return st;
  if ((st = upload_frags16(rs->parity_frag, rs->G.row(k), m, k))) return st;
  if (k >= 2 && k <= kMaxFftDataShards && k + m <= 4096) {
    rs->program = build_staged_program16(k, m);
    rs->has_program = true;
    rs->staged = fft16_fits_device(rs->program);
  }
  if (rs->staged) {  // the encode route's device program, uploaded with the codec
    const Fft16Device *d = nullptr;
    const int ust = fft16_device_for(rs.get(), &d);
    if (ust) return ust;
  }
  *out = rs.release();
  return TRS_OK;
}

void trs_rs16_free(trs_rs16 *rs) { delete rs; }
size_t trs_rs16_data_shard_count(const trs_rs16 *rs) { return rs ? rs->k : 0; }
size_t trs_rs16_parity_shard_count(const trs_rs16 *rs) { return rs ? rs->m : 0; }
size_t trs_rs16_total_shard_count(const trs_rs16 *rs) { return rs ? rs->n : 0; }
const char *trs_rs16_encode_route(const trs_rs16 *rs) {
  return rs && rs->staged ? "fft" : "matrix";
}

int trs_rs16_generator(const trs_rs16 *rs, uint16_t *out, size_t out_len) {
  if (!rs || !out || out_len != rs->n * rs->k) return TRS_ERR_INVALID_ARGUMENT;
  std::copy(rs->G.a.begin(), rs->G.a.end(), out);
  return TRS_OK;
}

namespace {

// Tower-FFT encode of `stripes` stripes; ptrs holds stripes * (k + m) shards.
int run_fft16(const trs_rs16 *rs, const std::vector<const uint8_t *> &ptrs, size_t stripes,
              size_t len, hipStream_t stream) {
  const Fft16Device *d = nullptr;
  int st = fft16_device_for(rs, &d);
  if (st) return st;
  if (len == 0 || stripes == 0) return TRS_OK;
  if (stripes == 1) return launch_fft16_one(*d, ptrs, fft16_device_stores(*d), len / 2, stream);
  bool al = true;
  for (auto q : ptrs) al = al && aligned16(q);
  PtrTable table;
  if ((st = table.upload(ptrs, stream))) return st;
  Fft16Call c{};
  c.ptrs = static_cast<const uint8_t *const *>(table.d);
  c.plane = len / 2;
  c.stripes = uint32_t(stripes);
  c.nstore = uint32_t(fft16_device_stores(*d));
  c.vec_ok = al ? 1u : 0u;
  return hip_status(launch_fft16(*d, c, stream));
}

// One launch of a device program over one stripe: ptrs = nload loaded shards
// then nstore stored shards.
int run_fft16_program(const Fft16Device &d, const std::vector<const uint8_t *> &ptrs,
                      size_t nstore, size_t len, hipStream_t stream) {
  if (len == 0 || nstore == 0) return TRS_OK;
  if (ptrs.size() != fft16_device_loads(d) + nstore) return TRS_ERR_INVALID_ARGUMENT;
  return launch_fft16_one(d, ptrs, nstore, len / 2, stream);
}

// The matrix encode of one stripe on the apply route's kernel (the
// Karatsuba table kernel or int8 MFMA), without routing by the tuner.
int matrix16_encode(const trs_rs16 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                    hipStream_t stream) {
  const int v = apply16_rule(rs->apply_route.load(), rs->k, rs->m, len, rs->parity_frag.p != nullptr);
  return v == kVarMfma ? run_mfma16(ptrs, rs->k, rs->m, rs->parity_frag.p, len, stream)
                       : run_gf16(ptrs, rs->k, rs->m, rs->parity, len, kWrite, nullptr, stream);
}

// RS16 encode through the rules (the FFT program at its default 64 KiB LDS
// target whenever the shape has one) or, with a model installed, the router
// over the matrix kernel and the program's LDS targets.
int encode16_ptrs(const trs_rs16 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                  hipStream_t stream) {
  const Fft16Device *d = nullptr;
  if (rs->staged) {
    const int st = fft16_device_for(rs, &d);
    if (st) return st;
  }
  auto run = [&](int v) {
    if (v == kVarTable) return run_gf16(ptrs, rs->k, rs->m, rs->parity, len, kWrite, nullptr, stream);
    if (v == kVarMfma) return run_mfma16(ptrs, rs->k, rs->m, rs->parity_frag.p, len, stream);
    if (!d) return int(TRS_ERR_UNSUPPORTED);
    return launch_fft16_one(*d, ptrs, fft16_device_stores(*d), len / 2, stream, false, nullptr,
                            variant_lds(v));
  };
  const int apply_route = rs->apply_route.load();
  const bool frag = rs->parity_frag.p != nullptr;
  // a forced apply route keeps the program and forces the matrix kernel
  const int rule = apply_route != 0 && !d ? apply16_rule(apply_route, rs->k, rs->m, len, frag)
                   : apply_route != 0     ? kVarProgram64K
                                          : enc16_rule_variant(rs->k, rs->m, len, d != nullptr, frag);
  // a forced apply route (trs_rs16_set_apply_route) is not tuned
  if (apply_route != 0) return run(rule);
  return router_run(RouteKey{16, 0, rs->k, rs->m, length_bucket(len)}, len, rule, stream,
                    [&] {
                      return (1u << kVarTable) | (rs->parity_frag.p ? 1u << kVarMfma : 0u) |
                             (d ? program_variants(*d, len / 2, false) : 0u);
                    },
                    run);
}

// note_shard_len over a shard list (src/reedsolomon16.rs:758-772, :704-711).
int check_shards16(const size_t *lens, size_t n, size_t *len_out) {
  size_t len = 0;
  for (size_t i = 0; i < n; ++i) {
    const size_t l = lens[i];
    if (l == 0) return TRS_EMPTY_SHARD;
    if (l & 1) return TRS_INCORRECT_SHARD_SIZE;
    if (len && l != len) return TRS_INCORRECT_SHARD_SIZE;
    len = l;
  }
  if (!len) return TRS_EMPTY_SHARD;
  *len_out = len;
  return TRS_OK;
}

// The automatic encode of a validated shard list, device or host.  Host
// shards keep their split-plane layout per column chunk (two planes).
int encode16_auto(const trs_rs16 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                  const Exec &ex) {
  return dispatch(ex, ptrs, rs->k, len,
                  [rs](const std::vector<const uint8_t *> &p, size_t l, hipStream_t s) {
                    return encode16_ptrs(rs, p, l, s);
                  });
}

// verify through the encode route: the tower FFT program compares its parity
// registers with the parity shards; the matrix route runs the apply kernel's
// verify mode.
int verify16_launch(const trs_rs16 *rs, const std::vector<const uint8_t *> &ptrs, size_t len,
                    uint32_t *flag, hipStream_t s) {
  if (rs->staged) {
    const Fft16Device *d = nullptr;
    const int st = fft16_device_for(rs, &d);
    if (st) return st;
    return launch_fft16_one(*d, ptrs, fft16_device_stores(*d), len / 2, s, false, flag);
  }
  return run_gf16(ptrs, rs->k, rs->m, rs->parity, len, kVerify, flag, s);
}

// ReedSolomon16::encode, src/reedsolomon16.rs:412-418.
int rs16_encode(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens, size_t n,
                const Exec &ex) {
  if (!rs || (n && (!shards || !lens))) return TRS_ERR_INVALID_ARGUMENT;
  if (n < rs->n) return TRS_TOO_FEW_SHARDS;
  if (n > rs->n) return TRS_TOO_MANY_SHARDS;
  size_t len = 0;
  int st = check_shards16(lens, n, &len);
  if (st) return st;
  std::vector<const uint8_t *> ptrs(shards, shards + n);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return encode16_auto(rs, ptrs, len, ex);
}

// ReedSolomon16::encode_sep, src/reedsolomon16.rs:422-447.
int rs16_encode_sep(const trs_rs16 *rs, const uint8_t *const *data, const size_t *dl, size_t nd,
                    uint8_t *const *parity, const size_t *pl, size_t np, const Exec &ex) {
  if (!rs || (nd && (!data || !dl)) || (np && (!parity || !pl))) return TRS_ERR_INVALID_ARGUMENT;
  // src/reedsolomon16.rs:427-446
  if (nd != rs->k) return nd < rs->k ? TRS_TOO_FEW_DATA_SHARDS : TRS_TOO_MANY_DATA_SHARDS;
  if (np != rs->m) return np < rs->m ? TRS_TOO_FEW_PARITY_SHARDS : TRS_TOO_MANY_PARITY_SHARDS;
  std::vector<size_t> lens(dl, dl + nd);
  lens.insert(lens.end(), pl, pl + np);
  size_t len = 0;
  int st = check_shards16(lens.data(), lens.size(), &len);
  if (st) return st;
  std::vector<const uint8_t *> ptrs(data, data + nd);
  ptrs.insert(ptrs.end(), parity, parity + np);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return encode16_auto(rs, ptrs, len, ex);
}
}  // namespace

int trs_rs16_encode(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens, size_t n,
                    void *stream) {
  return rs16_encode(rs, shards, lens, n, as_stream(stream));
}

int trs_rs16_encode_host(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens,
                         size_t n) {
  return rs16_encode(rs, shards, lens, n, Exec::on_host(2));
}

int trs_rs16_encode_sep(const trs_rs16 *rs, const uint8_t *const *data, const size_t *dl,
                        size_t nd, uint8_t *const *parity, const size_t *pl, size_t np,
                        void *stream) {
  return rs16_encode_sep(rs, data, dl, nd, parity, pl, np, as_stream(stream));
}

int trs_rs16_encode_sep_host(const trs_rs16 *rs, const uint8_t *const *data, const size_t *dl,
                             size_t nd, uint8_t *const *parity, const size_t *pl, size_t np) {
  return rs16_encode_sep(rs, data, dl, nd, parity, pl, np, Exec::on_host(2));
}

int trs_rs16_encode_sep_matrix(const trs_rs16 *rs, const uint8_t *const *data,
                               const size_t *dl, size_t nd, uint8_t *const *parity,
                               const size_t *pl, size_t np, void *stream) {
  if (!rs || (nd && (!data || !dl)) || (np && (!parity || !pl))) return TRS_ERR_INVALID_ARGUMENT;
  // src/reedsolomon16.rs:329-335 checks lengths only (counts via the apply).
  std::vector<size_t> lens(dl, dl + nd);
  lens.insert(lens.end(), pl, pl + np);
  size_t len = 0;
  int st = check_shards16(lens.data(), lens.size(), &len);
  if (st) return st;
  if (nd != rs->k || np != rs->m) return TRS_ERR_INVALID_ARGUMENT;
  std::vector<const uint8_t *> ptrs(data, data + nd);
  ptrs.insert(ptrs.end(), parity, parity + np);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return matrix16_encode(rs, ptrs, len, as_stream(stream));
}

int trs_rs16_encode_sep_fft(const trs_rs16 *rs, const uint8_t *const *data, const size_t *dl,
                            size_t nd, uint8_t *const *parity, const size_t *pl, size_t np,
                            void *stream) {
  if (!rs || (nd && (!data || !dl)) || (np && (!parity || !pl))) return TRS_ERR_INVALID_ARGUMENT;
  std::vector<size_t> lens(dl, dl + nd);
  lens.insert(lens.end(), pl, pl + np);
  size_t len = 0;
  int st = check_shards16(lens.data(), lens.size(), &len);
  if (st) return st;
  if (nd != rs->k || np != rs->m) return TRS_ERR_INVALID_ARGUMENT;
  std::vector<const uint8_t *> ptrs(data, data + nd);
  ptrs.insert(ptrs.end(), parity, parity + np);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return run_fft16(rs, ptrs, 1, len, as_stream(stream));
}

int trs_fft16_program_info(size_t k, size_t m, size_t *out, size_t n) {
  if (n && !out) return TRS_ERR_INVALID_ARGUMENT;
  if (k < 2 || k > kMaxFftDataShards || m == 0 || k + m > 4096) return TRS_ERR_UNSUPPORTED;
  const Fft16Program p = build_staged_program16(k, m);
  size_t v[9] = {p.register_count, p.mult_count, p.glue.size(), p.stages.size(),
                 p.zero_regs.size(), fft16_profitable(p, k, m) ? 1u : 0u, 0, 0, 0};
  const Fft16Lowered L = fft16_lower(p);
  v[6] = L.regs;
  v[7] = L.steps.size();
  v[8] = L.regs * 64;
  for (size_t i = 0; i < n && i < 9; ++i) out[i] = v[i];
  return TRS_OK;
}

namespace {
// counts[0..5) and, when the buffers hold them, the words of L.
int dump_lowered(const Fft16Lowered &L, uint32_t *steps, size_t steps_cap, uint32_t *consts,
                 size_t consts_cap, uint32_t *ops, size_t ops_cap, uint32_t *parity,
                 size_t parity_cap, size_t *counts) {
  counts[0] = L.steps.size() * 8;
  counts[1] = L.consts.size();
  counts[2] = L.ops.size();
  counts[3] = L.store.size();
  counts[4] = L.regs;
  if (steps_cap < counts[0] || consts_cap < counts[1] || ops_cap < counts[2] ||
      parity_cap < counts[3])
    return TRS_OK;  // sizes only
  if (!steps || !consts || !ops || !parity) return TRS_ERR_INVALID_ARGUMENT;
  std::memcpy(steps, L.steps.data(), counts[0] * sizeof(uint32_t));
  std::copy(L.consts.begin(), L.consts.end(), consts);
  std::copy(L.ops.begin(), L.ops.end(), ops);
  std::copy(L.store.begin(), L.store.end(), parity);
  return TRS_OK;
}

bool fft8_shape_ok(size_t k, size_t m) { return k >= 1 && m >= 1 && k + m <= 256; }
}  // namespace

int trs_fft16_lowered_program(size_t k, size_t m, uint32_t *steps, size_t steps_cap,
                              uint32_t *consts, size_t consts_cap, uint32_t *ops, size_t ops_cap,
                              uint32_t *parity, size_t parity_cap, size_t *counts) {
  if (!counts) return TRS_ERR_INVALID_ARGUMENT;
  if (k < 2 || k > kMaxFftDataShards || m == 0 || k + m > 4096) return TRS_ERR_UNSUPPORTED;
  return dump_lowered(fft16_lower(build_staged_program16(k, m)), steps, steps_cap, consts,
                      consts_cap, ops, ops_cap, parity, parity_cap, counts);
}

int trs_fft8_program_info(size_t k, size_t m, size_t *out, size_t n) {
  if (n && !out) return TRS_ERR_INVALID_ARGUMENT;
  if (!fft8_shape_ok(k, m)) return TRS_ERR_UNSUPPORTED;
  const Fft16Program r = build_staged_program8(k, m, false);
  const Fft16Program p = build_staged_program8(k, m, true);
  const Fft16Lowered L = fft16_lower(p);
  const size_t v[10] = {r.register_count, r.mult_count, r.glue.size(),
                        r.stages.size(), r.zero_regs.size(),
                        staged8_reference_route(k, m) ? 1u : 0u,
                        L.regs, L.steps.size(), L.regs * 32, p.mult_count};
  for (size_t i = 0; i < n && i < 10; ++i) out[i] = v[i];
  return TRS_OK;
}

int trs_fft8_lowered_program(size_t k, size_t m, uint32_t *steps, size_t steps_cap,
                             uint32_t *consts, size_t consts_cap, uint32_t *ops, size_t ops_cap,
                             uint32_t *parity, size_t parity_cap, size_t *counts) {
  if (!counts) return TRS_ERR_INVALID_ARGUMENT;
  if (!fft8_shape_ok(k, m)) return TRS_ERR_UNSUPPORTED;
  return dump_lowered(fft16_lower(build_staged_program8(k, m)), steps, steps_cap, consts,
                      consts_cap, ops, ops_cap, parity, parity_cap, counts);
}

int trs_decode_matrix(int field, size_t k, size_t n, const uint8_t *present, void *out,
                      size_t out_len, size_t *rows_out) {
  if (!present || !rows_out || (field != 8 && field != 16) || k == 0 || n <= k)
    return TRS_ERR_INVALID_ARGUMENT;
  if (n > (field == 8 ? 256u : 65536u)) return TRS_TOO_MANY_SHARDS;
  Mask m = make_mask(present, n);
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) count += present[i] != 0;
  if (count < k) return TRS_TOO_FEW_SHARDS_PRESENT;
  Plan p;
  split_indices(m, k, n, p);
  const size_t nrows = p.missing_data.size() + p.missing_parity.size();
  *rows_out = nrows;
  if (!out || out_len < nrows * k) return out ? TRS_ERR_INVALID_ARGUMENT : TRS_OK;
  if (field == 8) {
    std::vector<uint8_t> rows;
    fused_rows(Matrix<F8>(), k, p, rows);
    std::copy(rows.begin(), rows.end(), static_cast<uint8_t *>(out));
  } else {
    std::vector<uint16_t> rows;
    fused_rows(Matrix<F16>(), k, p, rows);
    std::copy(rows.begin(), rows.end(), static_cast<uint16_t *>(out));
  }
  return TRS_OK;
}

namespace {
int decode_program_dump(int field, size_t k, size_t n, const uint8_t *present, uint32_t *steps,
                        size_t steps_cap, uint32_t *consts, size_t consts_cap, uint32_t *ops,
                        size_t ops_cap, uint32_t *load, size_t load_cap, uint32_t *store,
                        size_t store_cap, size_t *counts) {
  if (!counts || !present || k == 0 || n <= k) return TRS_ERR_INVALID_ARGUMENT;
  if (n > (field == 8 ? 256 : kMaxDecodePoints)) return TRS_ERR_UNSUPPORTED;
  Plan p;
  split_indices(make_mask(present, n), k, n, p);
  if (p.valid.size() < k) return TRS_TOO_FEW_SHARDS_PRESENT;
  std::vector<uint32_t> outputs(p.missing_data);
  outputs.insert(outputs.end(), p.missing_parity.begin(), p.missing_parity.end());
  const Fft16Lowered L = field == 8 ? fft8_decode_lowered(n, p.valid, outputs)
                                    : fft16_decode_lowered(n, p.valid, outputs);
  const size_t c[6] = {L.steps.size() * 8, L.consts.size(), L.ops.size(), L.load.size(),
                       L.store.size(), L.regs};
  std::copy(c, c + 6, counts);
  if (steps_cap < c[0] || consts_cap < c[1] || ops_cap < c[2] || load_cap < c[3] ||
      store_cap < c[4])
    return TRS_OK;  // sizes only
  if (!steps || !consts || !ops || !load || !store) return TRS_ERR_INVALID_ARGUMENT;
  std::memcpy(steps, L.steps.data(), c[0] * sizeof(uint32_t));
  std::copy(L.consts.begin(), L.consts.end(), consts);
  std::copy(L.ops.begin(), L.ops.end(), ops);
  std::copy(L.load.begin(), L.load.end(), load);
  std::copy(L.store.begin(), L.store.end(), store);
  return TRS_OK;
}
}  // namespace

int trs_fft16_decode_program(size_t k, size_t n, const uint8_t *present, uint32_t *steps,
                             size_t steps_cap, uint32_t *consts, size_t consts_cap, uint32_t *ops,
                             size_t ops_cap, uint32_t *load, size_t load_cap, uint32_t *store,
                             size_t store_cap, size_t *counts) {
  return decode_program_dump(16, k, n, present, steps, steps_cap, consts, consts_cap, ops,
                             ops_cap, load, load_cap, store, store_cap, counts);
}

int trs_fft8_decode_program(size_t k, size_t n, const uint8_t *present, uint32_t *steps,
                            size_t steps_cap, uint32_t *consts, size_t consts_cap, uint32_t *ops,
                            size_t ops_cap, uint32_t *load, size_t load_cap, uint32_t *store,
                            size_t store_cap, size_t *counts) {
  return decode_program_dump(8, k, n, present, steps, steps_cap, consts, consts_cap, ops,
                             ops_cap, load, load_cap, store, store_cap, counts);
}

int trs_rs16_reconstruct_stripes(const trs_rs16 *rs, uint8_t *const *shards,
                                 const uint8_t *present, size_t stripes, size_t len, int data_only,
                                 void *stream) {
  if (!rs || (stripes && (!shards || !present))) return TRS_ERR_INVALID_ARGUMENT;
  if (stripes == 0) return TRS_OK;
  if (len == 0) return TRS_EMPTY_SHARD;
  if (len & 1) return TRS_INCORRECT_SHARD_SIZE;
  const size_t n = rs->n, k = rs->k;
  const int route = rs->decode_route.load();
  const std::vector<size_t> lens(n, len);
  // Stripes whose shards sit one fixed stride apart (the layout of a
  // (stripes, n, len) tensor) are addressed on the FFT route as base + index *
  // stride through the plan's device index list: no per-call pointer table.
  const uintptr_t stride = n > 1 ? reinterpret_cast<uintptr_t>(shards[1]) -
                                       reinterpret_cast<uintptr_t>(shards[0])
                                 : 0;
  bool regular = n > 1;
  for (size_t i = 0; i < stripes && regular; ++i) {
    const uintptr_t base = reinterpret_cast<uintptr_t>(shards[i * n]);
    regular = base != 0;
    for (size_t j = 1; j < n && regular; ++j)
      regular = reinterpret_cast<uintptr_t>(shards[i * n + j]) == base + j * stride;
  }
  struct Work {
    std::shared_ptr<Plan> plan;
    std::vector<const uint8_t *> ptrs;
    size_t nout = 0;
    const Fft16Device *fft = nullptr;
    const uint8_t *base = nullptr;  // indexed addressing (ptrs empty)
  };
  // Validate and plan every stripe before launching anything.
  std::vector<Work> work;
  for (size_t i = 0; i < stripes; ++i) {
    // every buffer has the call's (nonzero, even) length, so of scan_shards'
    // checks only the presence count remains
    Mask mask;
    size_t count = 0;
    flags_mask(present + i * n, n, mask, count);
    if (count == n) continue;
    if (count < k) return TRS_TOO_FEW_SHARDS_PRESENT;
    Work w;
    int st = plan16_for(rs, mask, &w.plan);
    if (st) return st;
    const Plan &p = *w.plan;
    const size_t nout = p.missing_data.size() + (data_only ? 0 : p.missing_parity.size());
    if (nout == 0) continue;
    if (use_fft_decode(route, n, k, nout)) {
      if ((st = plan_fft16(*w.plan, n, &w.fft))) return st;
      if (regular) {  // every buffer has the call's length and a non-null address
        w.nout = nout;
        w.base = shards[i * n];
        work.push_back(std::move(w));
        continue;
      }
    } else if ((st = plan16_matrix(*w.plan, k))) {
      return st;
    }
    if ((st = gather(*w.plan, shards + i * n, lens.data(), k, len, data_only != 0, w.ptrs, w.nout)))
      return st;
    work.push_back(std::move(w));
  }
  hipStream_t s = as_stream(stream);
  std::vector<const uint8_t *> all;
  std::vector<size_t> offset;
  for (Work &w : work) {
    if (!w.fft) {
      const int v = apply16_rule(rs->apply_route.load(), k, w.nout, len, w.plan->frag.p != nullptr);
      const int st = v == kVarMfma ? run_mfma16(w.ptrs, k, w.nout, w.plan->frag.p, len, s)
                                   : run_gf16(w.ptrs, k, w.nout, w.plan->coef, len, kWrite, nullptr, s);
      if (st) return st;
      continue;
    }
    if (w.base) continue;
    offset.push_back(all.size());
    all.insert(all.end(), w.ptrs.begin(), w.ptrs.end());
  }
  PtrTable table;
  if (!all.empty()) {
    const int st = table.upload(all, s);
    if (st) return st;
  }
  std::vector<Fft16JobSpec> jobs;
  size_t j = 0;
  for (Work &w : work) {
    if (!w.fft) continue;
    Fft16JobSpec spec{};
    spec.prog = w.fft;
    spec.plane = len / 2;
    spec.nstore = uint32_t(w.nout);
    if (w.base) {
      spec.idx = static_cast<const uint32_t *>(w.plan->fft_idx.p);
      spec.base = w.base;
      spec.stride = stride;
      spec.vec_ok = (aligned16(w.base) && stride % 16 == 0) ? 1u : 0u;
    } else {
      bool al = true;
      for (auto q : w.ptrs) al = al && aligned16(q);
      spec.ptrs = static_cast<const uint8_t *const *>(table.d) + offset[j++];
      spec.vec_ok = al ? 1u : 0u;
    }
    jobs.push_back(spec);
  }
  if (jobs.empty()) return TRS_OK;
  return hip_status(launch_fft16_jobs(jobs, s));
}

int trs_rs16_encode_stripes(const trs_rs16 *rs, uint8_t *const *shards, size_t stripes,
                            size_t len, int route, void *stream) {
  if (!rs || route < 0 || route > 2 || (stripes && !shards)) return TRS_ERR_INVALID_ARGUMENT;
  if (stripes == 0) return TRS_OK;
  if (len == 0) return TRS_EMPTY_SHARD;
  if (len & 1) return TRS_INCORRECT_SHARD_SIZE;
  std::vector<const uint8_t *> ptrs(shards, shards + stripes * rs->n);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  hipStream_t s = as_stream(stream);
  const bool fft = route == 2 || (route == 0 && rs->staged);
  if (fft) return run_fft16(rs, ptrs, stripes, len, s);
  for (size_t i = 0; i < stripes; ++i) {
    std::vector<const uint8_t *> one(ptrs.begin() + i * rs->n, ptrs.begin() + (i + 1) * rs->n);
    const int st = matrix16_encode(rs, one, len, s);
    if (st) return st;
  }
  return TRS_OK;
}

namespace {
// ReedSolomon16::verify, src/reedsolomon16.rs:453-505.
int rs16_verify(const trs_rs16 *rs, const uint8_t *const *shards, const size_t *lens, size_t n,
                int *ok, const Exec &ex) {
  if (!rs || !ok || (n && (!shards || !lens))) return TRS_ERR_INVALID_ARGUMENT;
  if (n < rs->n) return TRS_TOO_FEW_SHARDS;
  if (n > rs->n) return TRS_TOO_MANY_SHARDS;
  size_t len = 0;
  int st = check_shards16(lens, n, &len);
  if (st) return st;
  std::vector<const uint8_t *> ptrs(shards, shards + n);
  for (auto q : ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  return verify_dispatch(ex, ptrs, len, ok,
                         [rs](const std::vector<const uint8_t *> &p, size_t l, uint32_t *flag,
                              hipStream_t s) { return verify16_launch(rs, p, l, flag, s); });
}

// The apply of a tower reconstruct over the plan's survivors and outputs.
int rec16_launch(Plan &plan, size_t k, size_t total, size_t nout,
                 const std::vector<const uint8_t *> &ptrs, size_t len, hipStream_t s, int route,
                 int apply_route) {
  if (use_fft_decode(route, total, k, nout)) {
    const Fft16Device *d = nullptr;
    const int st = plan_fft16(plan, total, &d);
    if (st) return st;
    return run_fft16_program(*d, ptrs, nout, len, s);
  }
  const int st = plan16_matrix(plan, k);
  if (st) return st;
  // the decode plan's rows on the Karatsuba kernel or on int8 MFMA
  auto run = [&](int v) {
    return v == kVarMfma ? run_mfma16(ptrs, k, nout, plan.frag.p, len, s)
                         : run_gf16(ptrs, k, nout, plan.coef, len, kWrite, nullptr, s);
  };
  const int rule = apply16_rule(apply_route, k, nout, len, plan.frag.p != nullptr);
  if (apply_route != 0) return run(rule);
  return router_run(RouteKey{16, 1, k, nout, length_bucket(len)}, len, rule, s,
                    [&] { return (1u << kVarTable) | (plan.frag.p ? 1u << kVarMfma : 0u); },
                    run);
}

int reconstruct16(const trs_rs16 *rs, size_t k, size_t total, std::shared_ptr<Plan> plan,
                  uint8_t *const *shards, const size_t *lens, const uint8_t *present, size_t n,
                  int data_only, const Exec &ex, int route, int apply_route) {
  if (n && (!shards || !lens || !present)) return TRS_ERR_INVALID_ARGUMENT;
  if (n != total) return n < total ? TRS_TOO_FEW_SHARDS : TRS_TOO_MANY_SHARDS;
  Mask mask;
  size_t count = 0, len = 0;
  int st = scan_shards(lens, present, n, true, mask, count, len);
  if (st) return st;
  if (plan) {
    if (mask != plan->presence) return TRS_INVALID_SHARD_FLAGS;
    if (count == total) return TRS_OK;
    if (len == 0) return TRS_TOO_FEW_SHARDS_PRESENT;
  } else {
    if (count == total) return TRS_OK;
    if (count < k) return TRS_TOO_FEW_SHARDS_PRESENT;
    if ((st = plan16_for(rs, mask, &plan))) return st;
  }
  std::vector<const uint8_t *> ptrs;
  size_t nout = 0;
  if ((st = gather(*plan, shards, lens, k, len, data_only != 0, ptrs, nout))) return st;
  if (nout == 0) return TRS_OK;
  Plan &p = *plan;
  return dispatch(ex, ptrs, k, len,
                  [&](const std::vector<const uint8_t *> &q, size_t l, hipStream_t s) {
                    return rec16_launch(p, k, total, nout, q, l, s, route, apply_route);
                  });
}

// ReedSolomon16::reconstruct_rows, src/reedsolomon16.rs:546-569.
int rs16_reconstruct_rows(const trs_rs16 *rs, uint8_t *rows, size_t rows_len, size_t row_len,
                          const uint8_t *present, size_t n_present, const Exec &ex) {
  if (!rs || (n_present && !present)) return TRS_ERR_INVALID_ARGUMENT;
  // src/reedsolomon16.rs:552-560
  if (n_present != rs->n) return TRS_INVALID_SHARD_FLAGS;
  if (row_len == 0 || (row_len & 1)) return TRS_INCORRECT_SHARD_SIZE;
  if (rows_len != rs->n * row_len) return TRS_INCORRECT_SHARD_SIZE;
  if (!rows) return TRS_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t *> shards(rs->n);
  std::vector<size_t> lens(rs->n, row_len);
  for (size_t i = 0; i < rs->n; ++i) shards[i] = rows + i * row_len;
  return reconstruct16(rs, rs->k, rs->n, nullptr, shards.data(), lens.data(), present, rs->n, 0,
                       ex, rs->decode_route.load(), rs->apply_route.load());
}
}  // namespace

int trs_rs16_encode_stripes_noalloc(const trs_rs16 *rs, uint8_t *const *shards, size_t stripes,
                                    size_t len, void *stream) {
  if (!rs || (stripes && !shards)) return TRS_ERR_INVALID_ARGUMENT;
  if (stripes == 0) return TRS_OK;
  if (len == 0) return TRS_EMPTY_SHARD;
  if (len & 1) return TRS_INCORRECT_SHARD_SIZE;
  if (!rs->staged) return TRS_ERR_UNSUPPORTED;  // the FFT route only
  const size_t nptr = stripes * rs->n;
  if (nptr > kNoAllocScratchBytes / sizeof(void *)) return TRS_ERR_SCRATCH_LIMIT;
  for (size_t i = 0; i < nptr; ++i)
    if (!shards[i]) return TRS_ERR_INVALID_ARGUMENT;
  // uploaded by trs_rs16_new, so this only reads the codec
  const Fft16Device *d = nullptr;
  int st = fft16_device_for(rs, &d);
  if (st) return st;
  void *table = nullptr;  // the device address of this GPU's scratch array
  TRS_TRY(noalloc_scratch(&table));
  hipStream_t s = as_stream(stream);
  // pageable source: the runtime has read it when the call returns
  TRS_TRY(hipMemcpyAsync(table, shards, nptr * sizeof(void *), hipMemcpyHostToDevice, s));
  bool al = true;
  for (size_t i = 0; i < nptr; ++i) al = al && aligned16(shards[i]);
  Fft16Call c{};
  c.ptrs = static_cast<const uint8_t *const *>(table);
  c.plane = len / 2;
  c.stripes = uint32_t(stripes);
  c.nstore = uint32_t(fft16_device_stores(*d));
  c.vec_ok = al ? 1u : 0u;
  return hip_status(launch_fft16(*d, c, s));
}

int trs_rs16_verify(const trs_rs16 *rs, const uint8_t *const *shards, const size_t *lens,
                    size_t n, int *ok, void *stream) {
  return rs16_verify(rs, shards, lens, n, ok, as_stream(stream));
}

int trs_rs16_verify_host(const trs_rs16 *rs, const uint8_t *const *shards, const size_t *lens,
                         size_t n, int *ok) {
  return rs16_verify(rs, shards, lens, n, ok, Exec::on_host(2));
}

int trs_rs16_reconstruct(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens,
                         const uint8_t *present, size_t n, int data_only, void *stream) {
  if (!rs) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct16(rs, rs->k, rs->n, nullptr, shards, lens, present, n, data_only,
                       as_stream(stream), rs->decode_route.load(), rs->apply_route.load());
}

int trs_rs16_reconstruct_host(const trs_rs16 *rs, uint8_t *const *shards, const size_t *lens,
                              const uint8_t *present, size_t n, int data_only) {
  if (!rs) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct16(rs, rs->k, rs->n, nullptr, shards, lens, present, n, data_only,
                       Exec::on_host(2), rs->decode_route.load(), rs->apply_route.load());
}

int trs_rs16_set_apply_route(trs_rs16 *rs, int route) {
  if (!rs || route < 0 || route > 2) return TRS_ERR_INVALID_ARGUMENT;
  if (route == 2 && !rs->parity_frag.p) return TRS_ERR_UNSUPPORTED;
  rs->apply_route.store(route);
  return TRS_OK;
}

int trs_rs16_set_decode_route(trs_rs16 *rs, int route) {
  if (!rs || route < 0 || route > 2) return TRS_ERR_INVALID_ARGUMENT;
  if (route == 2 && rs->n > kMaxDecodePoints) return TRS_ERR_UNSUPPORTED;
  rs->decode_route.store(route);
  return TRS_OK;
}

int trs_rs16_reconstruct_rows(const trs_rs16 *rs, uint8_t *rows, size_t rows_len, size_t row_len,
                              const uint8_t *present, size_t n_present, void *stream) {
  return rs16_reconstruct_rows(rs, rows, rows_len, row_len, present, n_present,
                               as_stream(stream));
}

int trs_rs16_reconstruct_rows_host(const trs_rs16 *rs, uint8_t *rows, size_t rows_len,
                                   size_t row_len, const uint8_t *present, size_t n_present) {
  return rs16_reconstruct_rows(rs, rows, rows_len, row_len, present, n_present,
                               Exec::on_host(2));
}

int trs_rs16_prepare_decode(const trs_rs16 *rs, const uint8_t *present, size_t n,
                            trs_dec16 **out) {
  if (!rs || !out || (n && !present)) return TRS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (n != rs->n) return TRS_INVALID_SHARD_FLAGS;
  Mask m = make_mask(present, n);
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) count += present[i] != 0;
  if (count < rs->k) return TRS_TOO_FEW_SHARDS_PRESENT;
  std::shared_ptr<Plan> plan;
  int st = plan16_for(rs, m, &plan);
  if (st) return st;
  auto d = std::make_unique<trs_dec16>();
  d->k = rs->k;
  d->n = rs->n;
  d->decode_route = rs->decode_route.load();
  d->apply_route = rs->apply_route.load();
  d->plan = plan;
  *out = d.release();
  return TRS_OK;
}

int trs_dec16_reconstruct(const trs_dec16 *dec, uint8_t *const *shards, const size_t *lens,
                          const uint8_t *present, size_t n, int data_only, void *stream) {
  if (!dec) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct16(nullptr, dec->k, dec->n, dec->plan, shards, lens, present, n, data_only,
                       as_stream(stream), dec->decode_route, dec->apply_route);
}

int trs_dec16_reconstruct_host(const trs_dec16 *dec, uint8_t *const *shards, const size_t *lens,
                               const uint8_t *present, size_t n, int data_only) {
  if (!dec) return TRS_ERR_INVALID_ARGUMENT;
  return reconstruct16(nullptr, dec->k, dec->n, dec->plan, shards, lens, present, n, data_only,
                       Exec::on_host(2), dec->decode_route, dec->apply_route);
}

void trs_dec16_free(trs_dec16 *dec) { delete dec; }

}  // extern "C"

// ============================================================================
// batched GF(2^8) work list
// ============================================================================

struct trs_batch8 {
  struct Req {
    std::shared_ptr<Plan> plan;  // keeps a reconstruct plan's coefficients alive
    const DevCoef *coef = nullptr;
    std::vector<const uint8_t *> ptrs;
    uint64_t len = 0;
    uint32_t k = 0, nout = 0;
    uint32_t prog = 0;  // generated bit-sliced program (encode of generated shapes), 0 = table
    const void *frag = nullptr;  // MFMA A fragments: the job runs on the int8-MFMA kernel
    // staged FFT program or FFT erasure decoder: the job runs on the byte-plane
    // interpreter (one launch for all such jobs)
    const Fft16Device *fft = nullptr;
  };
  // One device work list per kernel: table-kernel jobs and generated-program jobs.
  struct List {
    const void *jobs = nullptr, *tiles = nullptr;  // inside `blob`
    uint32_t njobs = 0;
    uint64_t total_tiles = 0;
    uint32_t grid = 0;
  };
  std::vector<Req> reqs;
  // One device buffer for everything prepare lays out (pointer tables, each
  // list's jobs and tile map, the interpreter job table).  Prepare builds its
  // image in `image`; the next run copies it up on the run's stream through
  // the pinned staging ring (one asynchronous copy, ordered before the
  // launches), so neither call blocks on a pageable copy.
  DevBuf blob;
  std::vector<uint8_t> image;
  size_t image_bytes = 0;
  bool upload_pending = false;
  // Largest growth of `blob` that prepare may allocate (the server lowers it
  // to the device's free memory less its reserve); past it prepare fails
  // with TRS_ERR_OUT_OF_MEMORY before allocating.
  size_t grow_limit = SIZE_MAX;
  List table, mfma;
  const void *fft_table = nullptr;  // interpreter job table, inside `blob`
  Fft16JobsPlan fft_plan;
  // Runtime-coefficient jobs run on the bit-sliced tile (k_batch8_planes)
  // unless TRS_BATCH8_TILE=table selects the packed-byte table tile.
  bool planes = true;
  static constexpr size_t kMaxPrograms = 32;
  List bs[kMaxPrograms];  // one work list per generated program id (index 0 unused)
  bool prepared = false;
  // hipGraph of the prepared launches (captured on first run when enabled).
  bool use_graph = false;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Recorded after every run on the caller's stream: prepare rewrites `blob`
  // (pointer tables, job lists, tile maps) with a host copy, so it first waits
  // for the kernels of earlier runs, which read those tables, to finish.
  hipEvent_t done = nullptr;
  bool done_pending = false;
  hipError_t wait_done() {
    if (!done_pending) return hipSuccess;
    const hipError_t e = hipEventSynchronize(done);
    if (e == hipSuccess) done_pending = false;
    return e;
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~trs_batch8() {
    (void)wait_done();
    drop_graph();
    if (capture) (void)hipStreamDestroy(capture);
    if (done) (void)hipEventDestroy(done);
  }
};
extern "C" {

int trs_batch8_new(trs_batch8 **out) {
  if (!out) return TRS_ERR_INVALID_ARGUMENT;
  *out = new trs_batch8();
  return TRS_OK;
}

void trs_batch8_free(trs_batch8 *b) { delete b; }

void trs_batch8_clear(trs_batch8 *b) {
  if (!b) return;
  (void)b->wait_done();
  b->reqs.clear();
  b->prepared = false;
  b->upload_pending = false;
  b->drop_graph();
  b->table.njobs = b->mfma.njobs = 0;
  b->table.total_tiles = b->mfma.total_tiles = 0;
  b->fft_plan = Fft16JobsPlan{};
  for (auto &L : b->bs) {
    L.njobs = 0;
    L.total_tiles = 0;
  }
}

size_t trs_batch8_size(const trs_batch8 *b) { return b ? b->reqs.size() : 0; }

uint64_t trs_batch8_bytes(const trs_batch8 *b) {
  if (!b) return 0;
  uint64_t t = 0;
  for (auto &r : b->reqs) t += uint64_t(r.k + r.nout) * r.len;
  return t;
}

int trs_batch8_add_encode(trs_batch8 *b, const trs_rs8 *rs, uint8_t *const *shards,
                          const size_t *lens, size_t n) {
  if (!b || !rs || (n && (!shards || !lens))) return TRS_ERR_INVALID_ARGUMENT;
  if (n == 0) return TRS_TOO_FEW_SHARDS;
  int st = check_all8(rs->n, lens, n);
  if (st) return st;
  trs_batch8::Req r;
  r.coef = &rs->parity;
  r.ptrs.assign(shards, shards + n);
  for (auto q : r.ptrs)
    if (!q) return TRS_ERR_INVALID_ARGUMENT;
  r.len = lens[0];
  r.k = uint32_t(rs->k);
  r.nout = uint32_t(rs->m);
  if (rs->route != kRouteTable) r.prog = bs_program_id(rs->route, r.k, r.nout);
  // the staged FFT where it stays ahead at any length (its full-chip rule)
  if (rs->route == kRouteTable && rs->staged_max == SIZE_MAX &&
      (st = staged8_device(rs, &r.fft)))
    return st;
  b->reqs.push_back(std::move(r));
  b->prepared = false;
  return TRS_OK;
}

int trs_batch8_add_reconstruct(trs_batch8 *b, const trs_rs8 *rs, uint8_t *const *shards,
                               const size_t *lens, const uint8_t *present, size_t n,
                               int data_only) {
  if (!b || !rs || (n && (!shards || !lens || !present))) return TRS_ERR_INVALID_ARGUMENT;
  if (n != rs->n) return n < rs->n ? TRS_TOO_FEW_SHARDS : TRS_TOO_MANY_SHARDS;
  Mask mask;
  size_t count = 0, len = 0;
  int st = scan_shards(lens, present, n, false, mask, count, len);
  if (st) return st;
  if (count == rs->n) return TRS_OK;
  if (count < rs->k) return TRS_TOO_FEW_SHARDS_PRESENT;
  std::shared_ptr<Plan> plan;
  if ((st = plan8_for(rs, mask, &plan))) return st;
  trs_batch8::Req r;
  size_t nout = 0;
  if ((st = gather(*plan, shards, lens, rs->k, len, data_only != 0, r.ptrs, nout))) return st;
  if (nout == 0) return TRS_OK;
  r.plan = plan;
  r.coef = &plan->coef;
  r.len = len;
  r.k = uint32_t(rs->k);
  r.nout = uint32_t(nout);
  if (use_fft_decode8_batched(rs->decode_route.load(), rs->n, rs->k, nout)) {
    if ((st = plan_fft16(*plan, rs->n, &r.fft, true))) return st;
  } else if (use_mfma(rs->apply_route.load(), rs->k, len, true) && plan->frag.p) {
    bool al = true;
    for (auto q : r.ptrs) al = al && (reinterpret_cast<uintptr_t>(q) & 3) == 0;
    if (al) r.frag = plan->frag.p;
  }
  b->reqs.push_back(std::move(r));
  b->prepared = false;
  return TRS_OK;
}

int trs_batch8_prepare(trs_batch8 *b) {
  if (!b) return TRS_ERR_INVALID_ARGUMENT;
  TRS_TRY(b->wait_done());  // earlier runs still read the tables in `blob`
  b->drop_graph();
  size_t nptr = 0;
  for (auto &r : b->reqs) nptr += r.ptrs.size();
  std::vector<const uint8_t *> hp;
  hp.reserve(nptr);
  const size_t nprog = std::min<size_t>(bs_program_count() + 1, trs_batch8::kMaxPrograms);
  std::vector<BatchJob> jt, jm;
  std::vector<uint64_t> tt, tm;
  std::vector<std::vector<BatchJob>> jb(nprog);
  std::vector<std::vector<uint64_t>> tb(nprog);
  std::vector<uint64_t> tile_b(nprog, 0);
  uint64_t tile_t = 0, tile_m = 0;
  b->planes = true;
  if (const char *e = std::getenv("TRS_BATCH8_TILE")) b->planes = std::strcmp(e, "table") != 0;
  const uint64_t table_chunk = b->planes ? kPlanesChunkBytes : kChunkBytes;
  // Pointer tables sit at the start of `blob`; until it is allocated, job
  // pointers hold byte offsets into it (patched below).
  auto at_offset = [](size_t off) {
    return reinterpret_cast<const uint8_t *const *>(static_cast<uintptr_t>(off));
  };
  std::vector<Fft16JobSpec> fjobs;
  for (const auto &r : b->reqs) {
    BatchJob J{};
    J.ptrs = at_offset(hp.size() * sizeof(void *));
    bool al = true;
    for (auto q : r.ptrs) {
      hp.push_back(q);
      al = al && aligned16(q);
    }
    if (r.fft) {
      fjobs.push_back({r.fft, J.ptrs, r.len, r.nout, al ? 1u : 0u});
      continue;
    }
    const bool planes_job = b->planes && !r.frag && !r.prog;
    if (planes_job && !r.coef->planes) return TRS_ERR_INVALID_ARGUMENT;
    // the bit-sliced tile takes the scaled rows (and their exponents)
    J.coef = planes_job ? r.coef->planes_ptr() : r.coef->ptr();
    J.len = r.len;
    J.k = r.k;
    J.nout = r.nout;
    J.ostride = r.coef->ostride;
    J.vec_ok = al ? 1u : 0u;
    const uint32_t rows_per_tile = planes_job ? kBatchPlanesOut : kOutStride;
    J.out_tiles = (r.nout + rows_per_tile - 1) / rows_per_tile;
    J.prog = r.prog;
    if (r.frag) {
      J.coef = static_cast<const uint32_t *>(r.frag);
      J.tile_begin = tile_m;
      tm.push_back(tile_m);
      tile_m += ((r.len + mfma_group_cols() - 1) / mfma_group_cols()) * ((r.nout + 7) / 8);
      jm.push_back(J);
    } else if (r.prog && r.prog < nprog) {
      J.tile_begin = tile_b[r.prog];
      tb[r.prog].push_back(tile_b[r.prog]);
      tile_b[r.prog] += (r.len + kBsChunkBytes - 1) / kBsChunkBytes;
      jb[r.prog].push_back(J);
    } else {
      J.tile_begin = tile_t;
      tt.push_back(tile_t);
      tile_t += ((r.len + table_chunk - 1) / table_chunk) * J.out_tiles;
      jt.push_back(J);
    }
  }
  // Layout of `blob`: pointers, then per list its jobs and tile -> job map,
  // then the interpreter job table; 256-byte aligned pieces.
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t size = align(nptr * sizeof(void *) + 8);
  struct Piece {
    trs_batch8::List *list;
    std::vector<BatchJob> *jobs;
    std::vector<uint32_t> map;
    uint64_t total;
    uint64_t per_cu;
    size_t jobs_off = 0, map_off = 0;
  };
  // Workgroups per CU of each list's grid (capped at its tile count; each
  // workgroup strides over the tiles).  128 makes the table and generated-
  // program lists effectively one workgroup per tile, so the hardware
  // balances tiles of uneven cost: measured on the mixed C2 batch,
  // reconstruct 3100 -> 3670-3870 GB/s against 8 per CU and encode
  // 4467 -> 4675 GB/s against 4.
  uint64_t per_cu_table = 128, per_cu_bs = 128, per_cu_mfma = 2;
  if (const char *e = std::getenv("TRS_BATCH_GRID_PER_CU")) per_cu_table = std::atoi(e);
  if (const char *e = std::getenv("TRS_BATCH_BS_GRID_PER_CU")) per_cu_bs = std::atoi(e);
  if (const char *e = std::getenv("TRS_BATCH_MFMA_GRID_PER_CU")) per_cu_mfma = std::atoi(e);
  std::vector<Piece> pieces;
  auto add_piece = [&](trs_batch8::List &L, std::vector<BatchJob> &jobs,
                       std::vector<uint64_t> &tiles, uint64_t total, uint64_t per_cu) {
    L.njobs = 0;
    L.total_tiles = 0;
    L.grid = 0;
    L.jobs = L.tiles = nullptr;
    if (total == 0) return;  // nothing for this kernel
    Piece pc{&L, &jobs, std::vector<uint32_t>(total, 0), total, per_cu};
    for (size_t j = 0; j < jobs.size(); ++j) {  // tile -> job map, one uint32 per tile
      const uint64_t end = j + 1 < jobs.size() ? tiles[j + 1] : total;
      for (uint64_t t = tiles[j]; t < end; ++t) pc.map[t] = uint32_t(j);
    }
    pc.jobs_off = size;
    size = align(size + jobs.size() * sizeof(BatchJob) + 8);
    pc.map_off = size;
    size = align(size + pc.map.size() * sizeof(uint32_t));
    pieces.push_back(std::move(pc));
  };
  add_piece(b->table, jt, tt, tile_t, per_cu_table);
  for (size_t p = 1; p < nprog; ++p) add_piece(b->bs[p], jb[p], tb[p], tile_b[p], per_cu_bs);
  add_piece(b->mfma, jm, tm, tile_m, per_cu_mfma);
  // the interpreter table's size does not depend on the pointer values
  TRS_TRY(fft16_jobs_plan(fjobs, true, &b->fft_plan));
  const size_t fft_off = size;
  size = align(size + b->fft_plan.blob.size());
  if (size > b->blob.n && size - b->blob.n > b->grow_limit) return TRS_ERR_OUT_OF_MEMORY;
  TRS_TRY(b->blob.ensure(size));
  const uintptr_t base = reinterpret_cast<uintptr_t>(b->blob.p);
  auto patch = [base](const uint8_t *const *q) {
    return reinterpret_cast<const uint8_t *const *>(base + reinterpret_cast<uintptr_t>(q));
  };
  std::vector<uint8_t> &host = b->image;
  host.assign(size, 0);
  if (!hp.empty()) std::memcpy(host.data(), hp.data(), hp.size() * sizeof(void *));
  for (Piece &pc : pieces) {
    for (BatchJob &J : *pc.jobs) J.ptrs = patch(J.ptrs);
    if (!pc.jobs->empty())
      std::memcpy(host.data() + pc.jobs_off, pc.jobs->data(), pc.jobs->size() * sizeof(BatchJob));
    std::memcpy(host.data() + pc.map_off, pc.map.data(), pc.map.size() * sizeof(uint32_t));
  }
  b->fft_table = nullptr;
  if (!fjobs.empty()) {
    for (Fft16JobSpec &f : fjobs) f.ptrs = patch(f.ptrs);
    TRS_TRY(fft16_jobs_plan(fjobs, true, &b->fft_plan));
    if (b->fft_plan.total) {
      std::memcpy(host.data() + fft_off, b->fft_plan.blob.data(), b->fft_plan.blob.size());
      b->fft_table = reinterpret_cast<const void *>(base + fft_off);
    }
  }
  b->image_bytes = size;
  b->upload_pending = true;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  for (Piece &pc : pieces) {
    trs_batch8::List &L = *pc.list;
    L.jobs = reinterpret_cast<const void *>(base + pc.jobs_off);
    L.tiles = reinterpret_cast<const void *>(base + pc.map_off);
    L.njobs = uint32_t(pc.jobs->size());
    L.total_tiles = pc.total;
    uint64_t grid = uint64_t(cus) * pc.per_cu;
    if (grid > pc.total) grid = pc.total;
    L.grid = uint32_t(grid ? grid : 1);
  }
  b->prepared = true;
  return TRS_OK;
}

namespace {
int batch8_launch(trs_batch8 *b, hipStream_t s) {
  for (size_t p = 1; p < trs_batch8::kMaxPrograms; ++p) {
    const trs_batch8::List &L = b->bs[p];
    if (!L.total_tiles) continue;
    const int st = hip_status(launch_batch_bs(uint32_t(p), static_cast<const BatchJob *>(L.jobs),
                                              static_cast<const uint32_t *>(L.tiles),
                                              L.total_tiles, L.grid, s));
    if (st) return st;
  }
  if (b->mfma.total_tiles) {
    const int st = hip_status(launch_batch_mfma(static_cast<const BatchJob *>(b->mfma.jobs),
                                                static_cast<const uint32_t *>(b->mfma.tiles),
                                                b->mfma.total_tiles, b->mfma.grid, s));
    if (st) return st;
  }
  if (b->fft_plan.total) {
    const int st = hip_status(fft16_jobs_launch(b->fft_plan, b->fft_table, s));
    if (st) return st;
  }
  if (b->table.total_tiles)
    return hip_status((b->planes ? launch_batch8_planes : launch_batch8)(
        static_cast<const BatchJob *>(b->table.jobs),
        static_cast<const uint32_t *>(b->table.tiles), b->table.total_tiles, b->table.grid, s));
  return TRS_OK;
}
}  // namespace

int trs_batch8_set_graph(trs_batch8 *b, int enable) {
  if (!b) return TRS_ERR_INVALID_ARGUMENT;
  b->use_graph = enable != 0;
  if (!b->use_graph) {
    TRS_TRY(b->wait_done());
    b->drop_graph();
  }
  return TRS_OK;
}

int trs_batch8_run(trs_batch8 *b, void *stream) {
  if (!b) return TRS_ERR_INVALID_ARGUMENT;
  if (!b->prepared) {
    int st = trs_batch8_prepare(b);
    if (st) return st;
  }
  hipStream_t s = as_stream(stream);
  if (!b->done) TRS_TRY(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
  if (b->upload_pending) {
    // prepare waited for earlier runs, so nothing reads `blob` any more
    TRS_TRY(staged_upload(b->blob.p, b->image.data(), b->image_bytes, s));
    b->upload_pending = false;
  }
  if (!b->use_graph) {
    const int st = batch8_launch(b, s);
    if (st) return st;
    TRS_TRY(hipEventRecord(b->done, s));
    b->done_pending = true;
    return TRS_OK;
  }
  if (!b->exec) {
    // The launches read only buffers uploaded by prepare, so the captured
    // graph replays them unchanged; it holds kernel nodes only.
    if (!b->capture) TRS_TRY(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
    TRS_TRY(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
    const int st = batch8_launch(b, b->capture);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->capture, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    TRS_TRY(e);
    b->graph = g;
    TRS_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  TRS_TRY(hipGraphLaunch(b->exec, s));
  TRS_TRY(hipEventRecord(b->done, s));
  b->done_pending = true;
  return TRS_OK;
}

}  // extern "C"

// ============================================================================
// autotuning support: time one apply candidate on synthetic device data
// ============================================================================

namespace {

struct TuneSpace {
  std::mutex mu;
  DevBuf data;
  DevCoef coef;
  DevBuf frag;
  // (field << 48 | k << 24 | nout) -> staged FFT program on the interpreter
  std::vector<std::pair<uint64_t, Fft16DevicePtr>> fft16;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

TuneSpace &tune_space() {
  static TuneSpace *t = new TuneSpace();  // kept for the life of the process
  return *t;
}

uint64_t tune_rng(uint64_t &s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

namespace {
// A tuning shape's interpreter program (the tower's, or a GF(2^8) shape's
// staged program when it has no generated one), built once per shape.
int tune_program(TuneSpace &T, int field, uint32_t k, uint32_t nout, const Fft16Device **out) {
  *out = nullptr;
  if (field == 16 && (k < 2 || k > kMaxFftDataShards || k + nout > 4096)) return TRS_ERR_UNSUPPORTED;
  if (field == 8 && k + nout > 256) return TRS_ERR_UNSUPPORTED;
  const uint64_t key = (uint64_t(field) << 48) | (uint64_t(k) << 24) | nout;
  for (auto &e : T.fft16)
    if (e.first == key) {
      *out = e.second.get();
      return *out ? TRS_OK : TRS_ERR_UNSUPPORTED;
    }
  Fft16DevicePtr d;
  const hipError_t e = fft16_device_new(
      field == 16 ? build_staged_program16(k, nout) : build_staged_program8(k, nout), &d);
  if (e != hipSuccess && e != hipErrorInvalidValue) return hip_status(e);
  if (T.fft16.size() >= 64) T.fft16.clear();
  *out = d.get();
  T.fft16.emplace_back(key, std::move(d));  // an entry without a device program: does not fit
  return *out ? TRS_OK : TRS_ERR_UNSUPPORTED;
}

// The encode variants of shape (k, nout) at this length (the codec's own set).
int tune_feasible(TuneSpace &T, int field, uint32_t k, uint32_t nout, uint64_t len,
                  uint32_t *mask) {
  if (field == 16) {
    *mask = 1u << kVarTable;
    if (mfma16_available(k, nout)) *mask |= 1u << kVarMfma;
  } else {
    *mask = matrix_variants8(k, nout);
    if (bs_encode_available(kRouteBsFft, k, nout) || bs_encode_available(kRouteBsMatrix, k, nout)) {
      *mask |= 1u << kVarProgram32K;
      return TRS_OK;
    }
  }
  const Fft16Device *d = nullptr;
  const int st = tune_program(T, field, k, nout, &d);
  if (st == TRS_ERR_UNSUPPORTED) return TRS_OK;
  if (st) return st;
  *mask |= program_variants(*d, field == 16 ? len / 2 : len, field == 8);
  return TRS_OK;
}
}  // namespace

extern "C" int trs_tune_features(int field, uint32_t k, uint32_t nout, uint64_t len, int variant,
                                 float *out, size_t n) {
  if (!out || n != kTuneFeatures || (field != 8 && field != 16) || k == 0 || nout == 0 ||
      variant < 0 || variant >= kVariants)
    return TRS_ERR_INVALID_ARGUMENT;
  tune_features(field, k, nout, len, variant, out);
  return TRS_OK;
}

extern "C" int trs_tune_feasible(int field, uint32_t k, uint32_t nout, uint64_t len,
                                 uint32_t *mask) {
  if (!mask || (field != 8 && field != 16) || k == 0 || nout == 0 || len == 0)
    return TRS_ERR_INVALID_ARGUMENT;
  if (field == 16 && (len & 1)) return TRS_ERR_INVALID_ARGUMENT;
  if (field == 8 && k + nout > 256) return TRS_ERR_INVALID_ARGUMENT;
  TuneSpace &T = tune_space();
  std::lock_guard<std::mutex> lk(T.mu);
  return tune_feasible(T, field, k, nout, len, mask);
}

extern "C" int trs_tune_rule(int field, uint32_t k, uint32_t nout, uint64_t len, int *variant) {
  if (!variant || (field != 8 && field != 16) || k == 0 || nout == 0 || len == 0)
    return TRS_ERR_INVALID_ARGUMENT;
  if (field == 16 && (len & 1)) return TRS_ERR_INVALID_ARGUMENT;
  if (field == 8 && k + nout > 256) return TRS_ERR_INVALID_ARGUMENT;
  TuneSpace &T = tune_space();
  std::lock_guard<std::mutex> lk(T.mu);
  uint32_t mask = 0;
  int st = tune_feasible(T, field, k, nout, len, &mask);
  if (st) return st;
  int v;
  if (field == 16) {  // the codec's automatic tower encode
    const bool program = (mask >> kVarProgram16K) != 0;
    v = enc16_rule_variant(k, nout, len, program, ((mask >> kVarMfma) & 1u) != 0);
  } else if (bs_encode_available(kRouteBsFft, k, nout) ||
             bs_encode_available(kRouteBsMatrix, k, nout) || len <= staged8_gpu_max(k, nout)) {
    v = kVarProgram32K;
  } else if (use_mfma(0, k, len, false)) {
    v = kVarMfma;
  } else {
    v = len >= planes_min_len() ? planes_rule_variant(len, nout) : kVarTable;
  }
  // a default LDS target whose launch equals a smaller target's is listed as that one
  if (is_program(v) && !((mask >> v) & 1u)) {
    const Fft16Device *d = nullptr;
    if ((st = tune_program(T, field, k, nout, &d))) return st;
    const uint64_t cols = field == 16 ? len / 2 : len;
    const uint32_t gs = fft16_launch_gshift(*d, cols, field == 8, variant_lds(v));
    for (int w = kVarProgram16K; w <= kVarProgram64K; ++w)
      if (((mask >> w) & 1u) && fft16_launch_gshift(*d, cols, field == 8, variant_lds(w)) == gs) {
        v = w;
        break;
      }
  }
  *variant = v;
  return TRS_OK;
}

extern "C" int trs_tune_apply(int field, uint32_t k, uint32_t nout, uint64_t len, int variant,
                              int reps, double limit_ms, double *ms_out) {
  if (!ms_out || k == 0 || nout == 0 || len == 0 || reps <= 0) return TRS_ERR_INVALID_ARGUMENT;
  if (field != 8 && field != 16) return TRS_ERR_INVALID_ARGUMENT;
  if (field == 16 && (len & 1)) return TRS_ERR_INVALID_ARGUMENT;
  if (field == 8 && k + nout > 256) return TRS_ERR_INVALID_ARGUMENT;
  if (variant < 0 || variant >= kVariants) return TRS_ERR_INVALID_ARGUMENT;
  TuneSpace &T = tune_space();
  std::lock_guard<std::mutex> lk(T.mu);
  uint32_t mask = 0;
  int st = tune_feasible(T, field, k, nout, len, &mask);
  if (st) return st;
  if (!((mask >> variant) & 1u)) return TRS_ERR_UNSUPPORTED;
  const bool prog = is_program(variant);
  const int bs_route = field == 8 && prog
                           ? (bs_encode_available(kRouteBsFft, k, nout)      ? kRouteBsFft
                              : bs_encode_available(kRouteBsMatrix, k, nout) ? kRouteBsMatrix
                                                                               : 0)
                           : 0;
  const Fft16Device *prog16 = nullptr;
  if (prog && !bs_route && (st = tune_program(T, field, k, nout, &prog16))) return st;
  if (!T.stream) {
    TRS_TRY(hipStreamCreateWithFlags(&T.stream, hipStreamNonBlocking));
    TRS_TRY(hipEventCreate(&T.ev0));
    TRS_TRY(hipEventCreate(&T.ev1));
  }
  const uint64_t bytes = uint64_t(k + nout) * len;
  if (bytes > T.data.n) {
    TRS_TRY(T.data.alloc(bytes));
    TRS_TRY(launch_fill(static_cast<uint8_t *>(T.data.p), bytes, 0x7A9E5EEDull, T.stream));
  }
  // random nonzero coefficients (kernel time does not depend on their values)
  uint64_t seed = (uint64_t(field) << 40) ^ (uint64_t(k) << 20) ^ nout;
  if (field == 8) {
    std::vector<uint8_t> rows8(size_t(k) * nout);
    for (auto &c : rows8) c = uint8_t(1 + tune_rng(seed) % 255);
    if (!prog && variant != kVarMfma) st = upload_coef8(T.coef, rows8.data(), nout, k);
    if (!st && variant == kVarMfma) st = upload_frags(T.frag, rows8.data(), nout, k);
  } else if (!prog) {
    std::vector<uint16_t> rows16(size_t(k) * nout);
    for (auto &c : rows16) c = uint16_t(1 + tune_rng(seed) % 65535);
    st = variant == kVarMfma ? upload_frags16(T.frag, rows16.data(), nout, k)
                             : upload_coef16(T.coef, rows16.data(), nout, k);
  }
  if (st) return st;
  uint8_t *base = static_cast<uint8_t *>(T.data.p);
  std::vector<const uint8_t *> ptrs;
  for (uint32_t i = 0; i < k + nout; ++i) ptrs.push_back(base + uint64_t(i) * len);
  auto run_once = [&]() -> int {
    if (prog16) {
      return launch_fft16_one(*prog16, ptrs, nout, field == 16 ? len / 2 : len, T.stream,
                              field == 8, nullptr, variant_lds(variant));
    }
    if (field == 16)
      return variant == kVarMfma ? run_mfma16(ptrs, k, nout, T.frag.p, len, T.stream)
                                 : run_gf16(ptrs, k, nout, T.coef, len, kWrite, nullptr, T.stream);
    if (variant == kVarMfma) return run_mfma(ptrs, k, nout, T.frag.p, len, T.stream);
    if (bs_route) {
      Gf8Call c{};
      for (uint32_t i = 0; i < k + nout; ++i) c.ptr[i] = ptrs[i];
      c.len = len;
      c.k = k;
      c.nout = nout;
      c.vec_ok = len % 16 == 0 ? 1u : 0u;
      return hip_status(launch_bs_encode(bs_route, c, T.stream));
    }
    return run_gf8(ptrs, k, nout, T.coef, len, kWrite, nullptr, T.stream, variant);
  };
  float ms = 0;
  // Warm-up.  A first launch can carry one-time costs (loading the kernel's
  // code), so a warm-up over limit_ms runs once more before that time stands
  // as the measurement of a candidate already far behind.
  for (int w = 0; w < 2; ++w) {
    TRS_TRY(hipEventRecord(T.ev0, T.stream));
    if ((st = run_once())) return st;
    TRS_TRY(hipEventRecord(T.ev1, T.stream));
    TRS_TRY(hipEventSynchronize(T.ev1));
    TRS_TRY(hipEventElapsedTime(&ms, T.ev0, T.ev1));
    if (!(limit_ms > 0 && ms > limit_ms)) break;
    if (w == 1) {
      *ms_out = double(ms);
      return TRS_OK;
    }
  }
  TRS_TRY(hipEventRecord(T.ev0, T.stream));
  for (int r = 0; r < reps; ++r)
    if ((st = run_once())) return st;
  TRS_TRY(hipEventRecord(T.ev1, T.stream));
  TRS_TRY(hipEventSynchronize(T.ev1));
  TRS_TRY(hipEventElapsedTime(&ms, T.ev0, T.ev1));
  *ms_out = double(ms) / reps;
  return TRS_OK;
}

extern "C" int trs_tuner_install(const float *params, size_t n_params, const float *norm,
                                 size_t n_norm, int top_n) {
  if (!params || !norm || n_params != size_t(trs_mlp_param_count()) ||
      n_norm != 2 * kTuneFeatures || top_n < 1)
    return TRS_ERR_INVALID_ARGUMENT;
  auto m = std::make_shared<TunerModel>();
  int st = trs_mlp_new(0, &m->mlp);
  if (st) return st;
  if ((st = trs_mlp_set_params(m->mlp, params, n_params))) return st;
  for (size_t i = 0; i < kTuneFeatures; ++i) {
    m->mu[i] = norm[i];
    m->sd[i] = norm[kTuneFeatures + i] > 0 ? norm[kTuneFeatures + i] : 1.f;
  }
  m->top_n = top_n;
  Router &R = router();
  std::unique_lock<std::shared_mutex> lk(R.mu);
  R.model = std::move(m);
  R.cache.clear();
  return TRS_OK;
}

extern "C" int trs_tuner_remove(void) {
  Router &R = router();
  std::unique_lock<std::shared_mutex> lk(R.mu);
  R.model.reset();
  R.cache.clear();
  return TRS_OK;
}

extern "C" int trs_tuner_lookup(int field, int op, size_t k, size_t nout, size_t len,
                                int *variant) {
  if (!variant) return TRS_ERR_INVALID_ARGUMENT;
  Router &R = router();
  std::shared_lock<std::shared_mutex> lk(R.mu);
  auto it = R.cache.find(RouteKey{field, op, k, nout, length_bucket(len)});
  *variant = it == R.cache.end() ? -1 : it->second;
  return TRS_OK;
}

// ============================================================================
// serving loop: continuous batching of encode / reconstruct requests
// ============================================================================
//
// Callers submit requests from any thread; each is validated on submit
// exactly like the single call (same error codes) and gets a ticket.  A
// dispatcher thread packs pending requests into the work list of a free slot
// (one slot = one stream + one batch), admitting requests in arrival order
// while the bytes of in-flight work stay within the HBM budget, and launches
// it.  Slots run concurrently, so requests that arrive while others are on
// the GPU start as soon as a slot and budget free up.  A completion thread per
// slot waits for its batch, stamps every request's completion time and frees
// the slot.  GF(2^8) requests run in the slot's batched kernels; tower
// requests run as per-call applies on the slot's stream.

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

struct SrvReq {
  uint64_t ticket = 0;
  bool gf16 = false;
  trs_batch8::Req r8;                  // GF(2^8) request (batched)
  // GF(2^8) request run alone: the codec and the kind select the plain
  // call's own launch (kernel arguments carry the shard pointers, so nothing
  // is copied to the device first)
  const trs_rs8 *rs8 = nullptr;
  bool rec8 = false;
  std::shared_ptr<Plan> plan16;        // tower reconstruct plan (keeps coef alive)
  const DevCoef *coef16 = nullptr;     // tower coefficients
  std::vector<const uint8_t *> ptrs16;
  uint64_t len16 = 0;
  uint32_t k16 = 0, nout16 = 0;
  const Fft16Device *fft16 = nullptr;  // tower-FFT program (encode: the codec's; decode: the plan's)
  uint64_t bytes = 0;                  // algorithmic bytes (k + nout) * len
};

struct TicketState {
  int status = 0;
  bool done = false;
  // submitted; taken into a slot's batch; on the stream; seen complete
  Clock::time_point submit, dispatch, launched, finish;
  uint64_t bytes = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  trs_batch8 batch;
  std::vector<SrvReq> reqs;
  uint64_t bytes = 0;
  bool busy = false;      // owned by a batch (set when the dispatcher takes the slot)
  bool launched = false;  // the batch is on the stream; the completer may wait for it
  int launch_status = 0;
};

}  // namespace

struct trs_server {
  uint64_t budget = 0;
  size_t max_batch = 0;
  uint32_t window_us = 0;
  // Device memory the server's own tables may not take: a batch whose tables
  // need more than free HBM less this reserve is split, and a lone request
  // that still does not fit fails with TRS_ERR_OUT_OF_MEMORY.
  std::atomic<uint64_t> hbm_reserve{uint64_t(256) << 20};
  // How long the dispatcher and completion threads poll before they block.
  std::atomic<uint32_t> spin_us{0};
  std::atomic<uint64_t> npending{0};  // mirrors pending.size() for lock-free polling
  // counters (under mu): requests completed, batches launched, batches split
  // for memory, requests failed for memory
  uint64_t n_done = 0, n_batches = 0, n_splits = 0, n_oom = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SrvReq> pending;
  std::vector<std::unique_ptr<Slot>> slots;
  std::unordered_map<uint64_t, TicketState> tickets;
  uint64_t next_ticket = 1;
  uint64_t inflight = 0;
  bool stop = false;
  std::thread dispatcher;
  std::vector<std::thread> completers;
  std::vector<std::condition_variable *> slot_cv;
};

namespace {

// Poll `ready` for up to `spin_us` before the caller blocks: a thread that
// sleeps in the kernel takes tens of microseconds to be woken, which at low
// load is most of a small request's latency.
template <class F>
void spin_until(uint32_t spin_us, F ready) {
  if (!spin_us) return;
  const auto until = Clock::now() + std::chrono::microseconds(spin_us);
  while (!ready() && Clock::now() < until) std::this_thread::yield();
}

// Launches the requests placed in `slot->reqs` (called with the lock held,
// returns with it held): marks the slot busy, builds and enqueues its work
// outside the lock, records the slot's completion event and hands the slot
// to its completion thread.
void launch_slot(trs_server *S, Slot *slot, size_t si, std::unique_lock<std::mutex> &lk) {
  slot->busy = true;
  S->inflight += slot->bytes;
  lk.unlock();
  // build and launch the slot's batch outside the lock
  int st = TRS_OK;
  bool alone = false;
  while (true) {
    alone = slot->reqs.size() == 1 && !slot->reqs[0].gf16 && slot->reqs[0].rs8;
    if (alone) {
      // one GF(2^8) request: its plain per-call launch, as trs_rs8_encode /
      // trs_rs8_reconstruct make it, with no batch tables to upload
      const SrvReq &q = slot->reqs[0];
      const trs_rs8 *rs = q.rs8;
      st = q.rec8 ? rec8_launch(*q.r8.plan, q.r8.k, rs->n, q.r8.nout, q.r8.ptrs, q.r8.len,
                                slot->stream, rs->apply_route.load(), rs->decode_route.load())
                  : encode8_auto(rs, q.r8.ptrs, q.r8.len, Exec(slot->stream));
      break;
    }
    trs_batch8_clear(&slot->batch);
    for (auto &q : slot->reqs)
      if (!q.gf16) slot->batch.reqs.push_back(q.r8);
    if (slot->batch.reqs.empty()) break;
    // Tables that fit the slot's buffer need no memory; growing it is
    // admitted against free HBM less the reserve.
    slot->batch.grow_limit = 0;
    st = trs_batch8_prepare(&slot->batch);
    if (st != TRS_ERR_OUT_OF_MEMORY) break;
    size_t free_b = 0, total_b = 0;
    st = hip_status(hipMemGetInfo(&free_b, &total_b));
    if (st) break;
    const uint64_t reserve = S->hbm_reserve.load();
    slot->batch.grow_limit = free_b > reserve ? size_t(free_b - reserve) : 0;
    st = trs_batch8_prepare(&slot->batch);
    if (st != TRS_ERR_OUT_OF_MEMORY || slot->reqs.size() == 1) break;
    // split: the later half goes back to the head of the queue, in order
    // (a lone GF(2^8) request needs no tables at all)
    const size_t keep = slot->reqs.size() / 2;
    uint64_t back = 0;
    lk.lock();
    for (size_t i = slot->reqs.size(); i-- > keep;) {
      back += slot->reqs[i].bytes;
      S->pending.push_front(std::move(slot->reqs[i]));
    }
    slot->reqs.resize(keep);
    slot->bytes -= back;
    S->inflight -= back;
    S->npending.store(S->pending.size());
    ++S->n_splits;
    lk.unlock();
    st = TRS_OK;
  }
  if (!alone && !st && !slot->batch.reqs.empty())
    st = trs_batch8_run(&slot->batch, slot->stream);
  for (auto &q : slot->reqs) {
    if (!q.gf16 || st) continue;
    st = q.fft16 ? run_fft16_program(*q.fft16, q.ptrs16, q.nout16, q.len16, slot->stream)
                 : run_gf16(q.ptrs16, q.k16, q.nout16, *q.coef16, q.len16, kWrite, nullptr,
                            slot->stream);
  }
  if (!st) st = hip_status(hipEventRecord(slot->done, slot->stream));
  const auto launched = Clock::now();
  lk.lock();
  for (auto &q : slot->reqs) S->tickets[q.ticket].launched = launched;
  ++S->n_batches;
  if (st == TRS_ERR_OUT_OF_MEMORY) S->n_oom += slot->reqs.size();
  slot->launch_status = st;
  slot->launched = true;
  S->slot_cv[si]->notify_one();
}

void server_dispatch_loop(trs_server *S);

void server_dispatch(trs_server *S) {
  server_dispatch_loop(S);
  staged_upload_release();  // the batches' table uploads used this thread's ring
}

void server_dispatch_loop(trs_server *S) {
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    if (S->pending.empty() && !S->stop) {
      lk.unlock();
      spin_until(S->spin_us.load(), [&] { return S->npending.load() != 0; });
      lk.lock();
    }
    S->cv_work.wait(lk, [&] {
      if (S->stop) return true;
      if (S->pending.empty()) return false;
      for (auto &sl : S->slots)
        if (!sl->busy) return true;
      return false;
    });
    if (S->stop) return;
    // Short batching window, only while other work is in flight: then the
    // GPU is busy anyway and a burst of arrivals can join this batch.  With
    // every slot idle the request goes out at once (the window would only
    // add its length to the latency).
    bool busy = false;
    for (auto &sl : S->slots) busy = busy || sl->busy;
    if (S->window_us && busy && S->pending.size() < S->max_batch) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(S->window_us));
      lk.lock();
      if (S->stop) return;
    }
    Slot *slot = nullptr;
    size_t si = 0;
    for (size_t i = 0; i < S->slots.size(); ++i)
      if (!S->slots[i]->busy) {
        slot = S->slots[i].get();
        si = i;
        break;
      }
    if (!slot) continue;
    // admission by HBM budget, in arrival order
    slot->reqs.clear();
    slot->bytes = 0;
    const auto taken = Clock::now();
    while (!S->pending.empty() && slot->reqs.size() < S->max_batch) {
      const uint64_t b = S->pending.front().bytes;
      if (S->inflight + slot->bytes + b > S->budget && !(S->inflight == 0 && slot->reqs.empty()))
        break;
      slot->bytes += b;
      slot->reqs.push_back(std::move(S->pending.front()));
      S->pending.pop_front();
      S->tickets[slot->reqs.back().ticket].dispatch = taken;
    }
    S->npending.store(S->pending.size());
    if (slot->reqs.empty()) {
      // the next request does not fit beside in-flight work: wait for completions
      S->cv_work.wait(lk, [&] {
        return S->stop || S->pending.empty() || S->inflight == 0 ||
               S->inflight + S->pending.front().bytes <= S->budget;
      });
      continue;
    }
    launch_slot(S, slot, si, lk);
  }
}

void server_complete(trs_server *S, size_t si) {
  Slot *slot = S->slots[si].get();
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    S->slot_cv[si]->wait(lk, [&] { return S->stop || slot->launched; });
    if (!slot->launched) return;  // stopping with nothing in flight
    const int launch = slot->launch_status;
    lk.unlock();
    int st = launch;
    if (!st) {
      spin_until(S->spin_us.load(), [&] { return hipEventQuery(slot->done) != hipErrorNotReady; });
      st = hip_status(hipEventSynchronize(slot->done));
    }
    const auto now = Clock::now();
    lk.lock();
    for (auto &q : slot->reqs) {
      auto &t = S->tickets[q.ticket];
      t.done = true;
      t.status = st;
      t.finish = now;
    }
    S->n_done += slot->reqs.size();
    slot->reqs.clear();
    S->inflight -= slot->bytes;
    slot->bytes = 0;
    slot->launched = false;
    slot->busy = false;
    S->cv_done.notify_all();
    S->cv_work.notify_all();
    if (S->stop) return;
  }
}

int server_enqueue(trs_server *S, SrvReq &&q, uint64_t *ticket) {
  std::unique_lock<std::mutex> lk(S->mu);
  q.ticket = S->next_ticket++;
  TicketState t;
  t.submit = Clock::now();
  t.bytes = q.bytes;
  if (q.bytes == 0) {  // nothing to do (all present / data-only with no missing data)
    t.done = true;
    t.finish = t.submit;
  }
  S->tickets[q.ticket] = t;
  *ticket = q.ticket;
  if (!q.bytes) return TRS_OK;
  // With nothing queued and every slot idle (the dispatcher would launch at
  // once), the submitting thread launches the request itself: no wake-up of
  // the dispatcher thread on the way.  A budget too small for the request
  // still admits it, as the dispatcher does when nothing is in flight.
  bool idle = S->pending.empty() && !S->stop;
  for (auto &sl : S->slots) idle = idle && !sl->busy;
  if (idle) {
    Slot *slot = S->slots[0].get();
    slot->reqs.clear();
    slot->bytes = q.bytes;
    S->tickets[q.ticket].dispatch = Clock::now();
    slot->reqs.push_back(std::move(q));
    launch_slot(S, slot, 0, lk);
    return TRS_OK;
  }
  S->pending.push_back(std::move(q));
  S->npending.store(S->pending.size());
  S->cv_work.notify_one();
  return TRS_OK;
}

}  // namespace

extern "C" {

int trs_server_new(uint32_t slots, uint64_t hbm_budget_bytes, uint32_t max_batch,
                   uint32_t window_us, trs_server **out) {
  if (!out || slots == 0 || slots > 64 || max_batch == 0) return TRS_ERR_INVALID_ARGUMENT;
  auto S = std::make_unique<trs_server>();
  S->budget = hbm_budget_bytes ? hbm_budget_bytes : ~uint64_t(0);
  S->max_batch = max_batch;
  S->window_us = window_us;
  for (uint32_t i = 0; i < slots; ++i) {
    auto sl = std::make_unique<Slot>();
    TRS_TRY(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
    TRS_TRY(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
    S->slots.push_back(std::move(sl));
    S->slot_cv.push_back(new std::condition_variable());
  }
  trs_server *raw = S.get();
  raw->dispatcher = std::thread(server_dispatch, raw);
  for (uint32_t i = 0; i < slots; ++i) raw->completers.emplace_back(server_complete, raw, i);
  *out = S.release();
  return TRS_OK;
}

void trs_server_free(trs_server *S) {
  if (!S) return;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    // drain: wait for queued and in-flight work
    S->cv_done.wait(lk, [&] {
      if (!S->pending.empty()) return false;
      for (auto &sl : S->slots)
        if (sl->busy) return false;
      return true;
    });
    S->stop = true;
    S->cv_work.notify_all();
    for (auto *cv : S->slot_cv) cv->notify_all();
  }
  S->dispatcher.join();
  for (auto &t : S->completers) t.join();
  for (auto &sl : S->slots) {
    (void)hipEventDestroy(sl->done);
    (void)hipStreamDestroy(sl->stream);
  }
  for (auto *cv : S->slot_cv) delete cv;
  delete S;
}

int trs_server_submit_encode8(trs_server *S, const trs_rs8 *rs, uint8_t *const *shards,
                              const size_t *lens, size_t n, uint64_t *ticket) {
  if (!S || !ticket) return TRS_ERR_INVALID_ARGUMENT;
  trs_batch8 tmp;
  int st = trs_batch8_add_encode(&tmp, rs, shards, lens, n);
  if (st) return st;
  SrvReq q;
  q.r8 = std::move(tmp.reqs.front());
  q.rs8 = rs;
  q.bytes = uint64_t(q.r8.k + q.r8.nout) * q.r8.len;
  return server_enqueue(S, std::move(q), ticket);
}

int trs_server_submit_reconstruct8(trs_server *S, const trs_rs8 *rs, uint8_t *const *shards,
                                   const size_t *lens, const uint8_t *present, size_t n,
                                   int data_only, uint64_t *ticket) {
  if (!S || !ticket) return TRS_ERR_INVALID_ARGUMENT;
  trs_batch8 tmp;
  int st = trs_batch8_add_reconstruct(&tmp, rs, shards, lens, present, n, data_only);
  if (st) return st;
  SrvReq q;
  if (!tmp.reqs.empty()) {
    q.r8 = std::move(tmp.reqs.front());
    q.rs8 = rs;
    q.rec8 = true;
    q.bytes = uint64_t(q.r8.k + q.r8.nout) * q.r8.len;
  }
  return server_enqueue(S, std::move(q), ticket);
}

int trs_server_submit_encode16(trs_server *S, const trs_rs16 *rs, uint8_t *const *shards,
                               const size_t *lens, size_t n, uint64_t *ticket) {
  if (!S || !ticket || !rs || (n && (!shards || !lens))) return TRS_ERR_INVALID_ARGUMENT;
  if (n < rs->n) return TRS_TOO_FEW_SHARDS;
  if (n > rs->n) return TRS_TOO_MANY_SHARDS;
  size_t len = 0;
  int st = check_shards16(lens, n, &len);
  if (st) return st;
  SrvReq q;
  q.gf16 = true;
  q.ptrs16.assign(shards, shards + n);
  for (auto p : q.ptrs16)
    if (!p) return TRS_ERR_INVALID_ARGUMENT;
  q.coef16 = &rs->parity;
  q.len16 = len;
  q.k16 = uint32_t(rs->k);
  q.nout16 = uint32_t(rs->m);
  q.bytes = uint64_t(rs->n) * len;
  if (rs->staged && (st = fft16_device_for(rs, &q.fft16))) return st;
  return server_enqueue(S, std::move(q), ticket);
}

int trs_server_submit_reconstruct16(trs_server *S, const trs_rs16 *rs, uint8_t *const *shards,
                                    const size_t *lens, const uint8_t *present, size_t n,
                                    int data_only, uint64_t *ticket) {
  if (!S || !ticket || !rs || (n && (!shards || !lens || !present)))
    return TRS_ERR_INVALID_ARGUMENT;
  if (n != rs->n) return n < rs->n ? TRS_TOO_FEW_SHARDS : TRS_TOO_MANY_SHARDS;
  Mask mask;
  size_t count = 0, len = 0;
  int st = scan_shards(lens, present, n, true, mask, count, len);
  if (st) return st;
  SrvReq q;
  q.gf16 = true;
  if (count != rs->n) {
    if (count < rs->k) return TRS_TOO_FEW_SHARDS_PRESENT;
    std::shared_ptr<Plan> plan;
    if ((st = plan16_for(rs, mask, &plan))) return st;
    size_t nout = 0;
    if ((st = gather(*plan, shards, lens, rs->k, len, data_only != 0, q.ptrs16, nout))) return st;
    if (nout) {
      q.plan16 = plan;
      q.coef16 = &plan->coef;
      q.len16 = len;
      q.k16 = uint32_t(rs->k);
      q.nout16 = uint32_t(nout);
      q.bytes = uint64_t(rs->k + nout) * len;
      if (use_fft_decode(rs->decode_route.load(), rs->n, rs->k, nout)) {
        if ((st = plan_fft16(*plan, rs->n, &q.fft16))) return st;
      } else if ((st = plan16_matrix(*plan, rs->k))) {
        return st;
      }
    }
  }
  return server_enqueue(S, std::move(q), ticket);
}

int trs_server_wait(trs_server *S, uint64_t ticket, int *status) {
  if (!S) return TRS_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) return TRS_ERR_INVALID_ARGUMENT;
  S->cv_done.wait(lk, [&] { return S->tickets[ticket].done; });
  if (status) *status = S->tickets[ticket].status;
  return TRS_OK;
}

int trs_server_poll(trs_server *S, uint64_t ticket, int *done, int *status) {
  if (!S || !done) return TRS_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) return TRS_ERR_INVALID_ARGUMENT;
  *done = it->second.done ? 1 : 0;
  if (status) *status = it->second.status;
  return TRS_OK;
}

int trs_server_drain(trs_server *S) {
  if (!S) return TRS_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] {
    if (!S->pending.empty()) return false;
    for (auto &sl : S->slots)
      if (sl->busy) return false;
    return true;
  });
  return TRS_OK;
}

// Latency (microseconds, submit to completion) and bytes of a finished ticket;
// the entry is forgotten afterwards.
int trs_server_take(trs_server *S, uint64_t ticket, double *latency_us, uint64_t *bytes,
                    int *status) {
  return trs_server_take_detail(S, ticket, latency_us, nullptr, bytes, status);
}

int trs_server_take_detail(trs_server *S, uint64_t ticket, double *latency_us,
                           double *phases_us, uint64_t *bytes, int *status) {
  if (!S) return TRS_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end() || !it->second.done) return TRS_ERR_INVALID_ARGUMENT;
  const TicketState &t = it->second;
  auto us = [](Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::micro>(b - a).count();
  };
  if (latency_us) *latency_us = us(t.submit, t.finish);
  if (phases_us) {
    // a request with nothing to do never reaches a slot
    const bool ran = t.dispatch != Clock::time_point{};
    phases_us[0] = ran ? us(t.submit, t.dispatch) : 0.0;
    phases_us[1] = ran ? us(t.dispatch, t.launched) : 0.0;
    phases_us[2] = ran ? us(t.launched, t.finish) : 0.0;
  }
  if (bytes) *bytes = t.bytes;
  if (status) *status = t.status;
  S->tickets.erase(it);
  return TRS_OK;
}

int trs_server_set_hbm_reserve(trs_server *S, uint64_t reserve_bytes) {
  if (!S) return TRS_ERR_INVALID_ARGUMENT;
  S->hbm_reserve.store(reserve_bytes);
  return TRS_OK;
}

int trs_server_stats(trs_server *S, uint64_t *out4) {
  if (!S || !out4) return TRS_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(S->mu);
  out4[0] = S->n_done;
  out4[1] = S->n_batches;
  out4[2] = S->n_splits;
  out4[3] = S->n_oom;
  return TRS_OK;
}

int trs_server_set_spin(trs_server *S, uint32_t spin_us) {
  if (!S || spin_us > 100000) return TRS_ERR_INVALID_ARGUMENT;
  S->spin_us.store(spin_us);
  return TRS_OK;
}

}  // extern "C"

// ============================================================================
// host-side data path and wire formats (no reference counterpart: the
// reference's buffers are host slices)
// ============================================================================


extern "C" {

int trs_u16_to_split(const uint16_t *symbols, size_t n_symbols, uint8_t *shard, void *stream) {
  if (n_symbols && (!symbols || !shard)) return TRS_ERR_INVALID_ARGUMENT;
  return hip_status(launch_u16_split(reinterpret_cast<const uint8_t *>(symbols), shard, n_symbols,
                                     true, as_stream(stream)));
}

int trs_split_to_u16(const uint8_t *shard, size_t n_symbols, uint16_t *symbols, void *stream) {
  if (n_symbols && (!symbols || !shard)) return TRS_ERR_INVALID_ARGUMENT;
  return hip_status(launch_u16_split(reinterpret_cast<const uint8_t *>(symbols),
                                     const_cast<uint8_t *>(shard), n_symbols, false,
                                     as_stream(stream)));
}

int trs_rs8_reconstruct_alloc(const trs_rs8 *rs, uint8_t **shards, size_t *lens, size_t n,
                              int data_only, void *stream) {
  if (!rs || (n && (!shards || !lens))) return TRS_ERR_INVALID_ARGUMENT;
  if (n != rs->n) return n < rs->n ? TRS_TOO_FEW_SHARDS : TRS_TOO_MANY_SHARDS;
  std::vector<uint8_t> present(n);
  size_t count = 0, len = 0;
  for (size_t i = 0; i < n; ++i) {
    present[i] = shards[i] != nullptr;
    if (!present[i]) continue;
    if (lens[i] == 0) return TRS_EMPTY_SHARD;
    if (len && lens[i] != len) return TRS_INCORRECT_SHARD_SIZE;
    len = lens[i];
    ++count;
  }
  if (count == n) return TRS_OK;
  if (count < rs->k) return TRS_TOO_FEW_SHARDS_PRESENT;
  // src/reedsolomon.rs:88-101: a missing Option shard is allocated (zeroed,
  // then rebuilt); with data_only the missing parity stays None (:842-847)
  std::vector<size_t> made;
  std::vector<size_t> all_lens(lens, lens + n);
  hipStream_t s = as_stream(stream);
  for (size_t i = 0; i < n; ++i) {
    if (present[i] || (data_only && i >= rs->k)) continue;
    void *p = nullptr;  // every byte is rebuilt, so the zero fill is implied
    hipError_t e = hipMalloc(&p, len);
    if (e != hipSuccess) {
      if (p) (void)hipFree(p);
      for (size_t j : made) {
        (void)hipFree(shards[j]);
        shards[j] = nullptr;
      }
      return hip_status(e);
    }
    shards[i] = static_cast<uint8_t *>(p);
    all_lens[i] = len;
    made.push_back(i);
  }
  const int st = reconstruct8(rs, rs->k, rs->n, nullptr, shards, all_lens.data(), present.data(),
                              n, data_only, s, rs->apply_route.load());
  if (st) {
    (void)hipStreamSynchronize(s);
    for (size_t j : made) {
      (void)hipFree(shards[j]);
      shards[j] = nullptr;
    }
    return st;
  }
  for (size_t j : made) lens[j] = len;
  return TRS_OK;
}

int trs_device_free(void *p) { return p ? hip_status(hipFree(p)) : TRS_OK; }

int trs_device_alloc(size_t bytes, void **out) {
  if (!out) return TRS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  return hip_status(hipMalloc(out, bytes ? bytes : 1));
}

int trs_memcpy(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes && (!dst || !src)) return TRS_ERR_INVALID_ARGUMENT;
  if (!bytes) return TRS_OK;
  return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, as_stream(stream)));
}

int trs_stream_synchronize(void *stream) {
  return hip_status(hipStreamSynchronize(as_stream(stream)));
}

int trs_set_host_chunk(size_t bytes) {
  g_host_chunk.store(bytes);
  return TRS_OK;
}

// Clone (src/reedsolomon.rs:725-739): the same codec with a copy of the plan
// cache; the cached plans themselves are immutable and shared.
int trs_rs8_clone(const trs_rs8 *rs, trs_rs8 **out) {
  if (!rs || !out) return TRS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  trs_rs8 *c = nullptr;
  const int st = trs_rs8_new(rs->k, rs->m, &c);
  if (st) return st;
  c->apply_route.store(rs->apply_route.load());
  c->decode_route.store(rs->decode_route.load());
  {
    std::shared_lock<std::shared_mutex> lk(rs->plans.mu);
    c->plans.entries = rs->plans.entries;
  }
  *out = c;
  return TRS_OK;
}

int trs_rs16_clone(const trs_rs16 *rs, trs_rs16 **out) {
  if (!rs || !out) return TRS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  trs_rs16 *c = nullptr;
  const int st = trs_rs16_new(rs->k, rs->m, &c);
  if (st) return st;
  c->apply_route.store(rs->apply_route.load());
  c->decode_route.store(rs->decode_route.load());
  {
    std::shared_lock<std::shared_mutex> lk(rs->plans.mu);
    c->plans.entries = rs->plans.entries;
  }
  *out = c;
  return TRS_OK;
}

}  // extern "C"
