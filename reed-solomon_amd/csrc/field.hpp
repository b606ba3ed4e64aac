// Host-side field arithmetic and dense matrices for the cold path
// (generator construction, decode-plan inversion, table building).
//
// GF(2^8): poly x^8+x^4+x^3+x^2+1 (0x11d), generator 2, log/antilog tables
// with the antilog doubled so log a + log b needs no reduction
// (src/galois.rs:3-111).  GF((2^8)^2): high*y + low modulo y^2 + y + 32,
// packed (high << 8) | low (src/galois16.rs:16-59); here the product runs
// through log/antilog tables of the tower built from that definition.
// Matrices: row-major, Vandermonde rows use evaluation point = row index
// (src/field.rs:56-61, :84-87; src/matrix.rs:183-196).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace trs {

struct Gf8 {
  uint8_t log[256];
  uint8_t exp[510];
  uint8_t mul[256][256];
  Gf8() {
    std::memset(log, 0, sizeof log);
    unsigned b = 1;
    for (unsigned l = 0; l < 255; ++l) {
      log[b] = uint8_t(l);
      exp[l] = uint8_t(b);
      exp[l + 255] = uint8_t(b);
      b <<= 1;
      if (b & 0x100) b ^= 0x11d;
    }
    for (unsigned a = 0; a < 256; ++a)
      for (unsigned c = 0; c < 256; ++c)
        mul[a][c] = (a && c) ? exp[log[a] + log[c]] : 0;
  }
  uint8_t inv(uint8_t a) const { return a ? exp[255 - log[a]] : 0; }
};

const Gf8 &gf8();

constexpr uint8_t kTowerL = 32;

// Tower product from the definition (used to build the tables below).
inline uint16_t tower_mul_def(const Gf8 &f, uint16_t a, uint16_t b) {
  const uint8_t ah = a >> 8, al = a & 0xff, bh = b >> 8, bl = b & 0xff;
  const uint8_t p2 = f.mul[ah][bh];
  const uint8_t p1 = f.mul[ah][bl] ^ f.mul[al][bh];
  const uint8_t p0 = f.mul[al][bl];
  return uint16_t(((p1 ^ p2) << 8) | (p0 ^ f.mul[p2][kTowerL]));
}

struct Gf16 {
  std::vector<uint16_t> log;  // 65536
  std::vector<uint16_t> exp;  // 2 * 65535
  Gf16();
  uint16_t mul(uint16_t a, uint16_t b) const {
    return (a && b) ? exp[unsigned(log[a]) + log[b]] : 0;
  }
  uint16_t inv(uint16_t a) const { return a ? exp[(65535u - log[a]) % 65535u] : 0; }
};

const Gf16 &gf16();

// Field adapters for the generic matrix code.
struct F8 {
  using T = uint8_t;
  static T mul(T a, T b) { return gf8().mul[a][b]; }
  static T inv(T a) { return gf8().inv(a); }
};
struct F16 {
  using T = uint16_t;
  static T mul(T a, T b) { return gf16().mul(a, b); }
  static T inv(T a) { return gf16().inv(a); }
};

template <class F>
struct Matrix {
  using T = typename F::T;
  size_t rows = 0, cols = 0;
  std::vector<T> a;
  Matrix() = default;
  Matrix(size_t r, size_t c) : rows(r), cols(c), a(r * c, T(0)) {}
  T &at(size_t r, size_t c) { return a[r * cols + c]; }
  T at(size_t r, size_t c) const { return a[r * cols + c]; }
  const T *row(size_t r) const { return a.data() + r * cols; }
  T *row(size_t r) { return a.data() + r * cols; }
};

// Runs body(r0, r1) over [0, n) split across hardware threads when the work
// is large enough to pay for them.
template <class Body>
void parallel_rows(size_t n, size_t work_per_row, Body body) {
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  if (hw > 16) hw = 16;
  if (n * work_per_row < (1u << 22) || hw == 1 || n < 2) {
    body(size_t(0), n);
    return;
  }
  const size_t nt = std::min<size_t>(hw, n);
  std::vector<std::thread> ts;
  for (size_t t = 0; t < nt; ++t) {
    const size_t r0 = n * t / nt, r1 = n * (t + 1) / nt;
    ts.emplace_back([=] { body(r0, r1); });
  }
  for (auto &th : ts) th.join();
}

// out = a * b (src/matrix.rs:51-77, row-at-a-time accumulation).
template <class F>
Matrix<F> multiply(const Matrix<F> &a, const Matrix<F> &b) {
  Matrix<F> out(a.rows, b.cols);
  parallel_rows(a.rows, a.cols * b.cols, [&](size_t r0, size_t r1) {
    for (size_t r = r0; r < r1; ++r) {
      auto *o = out.row(r);
      for (size_t i = 0; i < a.cols; ++i) {
        const auto s = a.at(r, i);
        if (!s) continue;
        const auto *br = b.row(i);
        for (size_t c = 0; c < b.cols; ++c) o[c] ^= F::mul(s, br[c]);
      }
    }
  });
  return out;
}

// Inverse by Gauss-Jordan elimination on [M | I]; false when singular
// (src/matrix.rs:123-181).  The inverse of a nonsingular matrix is unique,
// so the elimination order does not affect the result.
template <class F>
bool invert(const Matrix<F> &m, Matrix<F> &out) {
  const size_t n = m.rows;
  Matrix<F> w(n, 2 * n);
  for (size_t r = 0; r < n; ++r) {
    std::copy(m.row(r), m.row(r) + n, w.row(r));
    w.at(r, n + r) = 1;
  }
  for (size_t c = 0; c < n; ++c) {
    size_t p = c;
    while (p < n && !w.at(p, c)) ++p;
    if (p == n) return false;
    if (p != c)
      for (size_t j = 0; j < 2 * n; ++j) std::swap(w.at(p, j), w.at(c, j));
    const auto s = F::inv(w.at(c, c));
    for (size_t j = 0; j < 2 * n; ++j) w.at(c, j) = F::mul(s, w.at(c, j));
    const auto *pr = w.row(c);
    parallel_rows(n, 2 * n, [&](size_t r0, size_t r1) {
      for (size_t r = r0; r < r1; ++r) {
        if (r == c) continue;
        auto *rr = w.row(r);
        const auto f = rr[c];
        if (!f) continue;
        for (size_t j = c; j < 2 * n; ++j) rr[j] ^= F::mul(f, pr[j]);
      }
    });
  }
  out = Matrix<F>(n, n);
  for (size_t r = 0; r < n; ++r) std::copy(w.row(r) + n, w.row(r) + 2 * n, out.row(r));
  return true;
}

// Systematic generator G = V(n,k) * inv(V[0..k]) (src/reedsolomon.rs:154-159).
template <class F>
bool build_generator(size_t k, size_t n, Matrix<F> &g) {
  Matrix<F> v(n, k);
  for (size_t r = 0; r < n; ++r) {
    typename F::T p = 1;
    const auto point = typename F::T(r);
    for (size_t c = 0; c < k; ++c) {
      v.at(r, c) = p;
      p = F::mul(p, point);
    }
  }
  Matrix<F> top(k, k), top_inv;
  std::copy(v.a.begin(), v.a.begin() + k * k, top.a.begin());
  if (!invert(top, top_inv)) return false;
  g = multiply(v, top_inv);
  return true;
}

}  // namespace trs
