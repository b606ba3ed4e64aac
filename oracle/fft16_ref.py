"""Oracle (test infrastructure only): the reference's tower-FFT program
builder and scalar replay, restated in Python.

Follows src/fft16.rs: what() tables (:42-67), Builder.interpolate (:225-265),
fft_with_zero_tail (:272-293), folded_chunk (:300-343), the two base
strategies chosen by multiply count (:345-376, :378-436), lower_staged
(:438-621), kara_triple (:176-180), profitable (:85-106) and
run_staged_symbols (:627-690).  The data polynomial is interpolated through
points 0..k-1 and evaluated at the parity points, so for systematic
Reed-Solomon with evaluation point i = i the parity equals the matrix encode
byte for byte (src/fft16.rs:4-8).

Pins: tests/test_fft16.py replays these programs against the symbol-wise
matrix oracle (the reference's staged_matches_matrix shapes, with stale
registers so the zero list is exercised, src/fft16.rs:958-995) and checks the
(86,170) multiply bound (src/fft16.rs:1077-1083).  Only tests may import this.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

MAX_FFT_DATA_SHARDS = 1024
LEVELS = 11  # block <= 2^LEVELS covers MAX_FFT_DATA_SHARDS (src/fft16.rs:30-34)
MAX_STAGED_REGISTERS16 = 2048
TOWER_L = 32


def _gf8():
    log = np.zeros(256, np.int64)
    exp = np.zeros(510, np.int64)
    b = 1
    for i in range(255):
        log[b] = i
        exp[i] = exp[i + 255] = b
        b <<= 1
        if b & 0x100:
            b ^= 0x11D
    return log, exp


_LOG8, _EXP8 = _gf8()


def gf8_mul(a, b):
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    r = _EXP8[(_LOG8[a] + _LOG8[b]) % 255]
    return np.where((a == 0) | (b == 0), 0, r)


def tower_mul(a, b):
    """(ah*y + al)(bh*y + bl) mod y^2 + y + 32 (src/galois16.rs:45-59)."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    ah, al, bh, bl = a >> 8, a & 0xFF, b >> 8, b & 0xFF
    p2 = gf8_mul(ah, bh)
    p1 = gf8_mul(ah, bl) ^ gf8_mul(al, bh)
    p0 = gf8_mul(al, bl)
    return ((p1 ^ p2) << 8) | (p0 ^ gf8_mul(p2, TOWER_L))


def tower_inv(a: int) -> int:
    # a^(65534) by square-and-multiply
    r, f, e = 1, a, 65534
    while e:
        if e & 1:
            r = int(tower_mul(r, f))
        f = int(tower_mul(f, f))
        e >>= 1
    return r


def _what_tables():
    x = np.arange(65536, dtype=np.int64)
    w = [x]
    for j in range(LEVELS - 1):
        piv = int(w[j][1 << j])
        w.append(tower_mul(w[j], w[j] ^ piv))
    out = []
    for j in range(LEVELS):
        inv = tower_inv(int(w[j][1 << j]))
        out.append(tower_mul(w[j], inv).astype(np.int64))
    return out


_WHAT: Optional[List[np.ndarray]] = None


def what() -> List[np.ndarray]:
    """WHAT[j][x]: normalised subspace polynomial of level j (src/fft16.rs:42-67)."""
    global _WHAT
    if _WHAT is None:
        _WHAT = _what_tables()
    return _WHAT


def kara_triple(c: int) -> Tuple[int, int, int]:
    """[c_lo, c_hi ^ c_lo, c_hi * L] (src/fft16.rs:176-180)."""
    hi, lo = c >> 8, c & 0xFF
    return lo, hi ^ lo, int(gf8_mul(hi, TOWER_L))


# op kinds (src/fft16.rs:112-123)
MX, XO, MC, CP = "mx", "xo", "mc", "cp"


@dataclass
class Program16:
    k: int
    m: int
    stages: List[Tuple]             # ("T", inverse, start, size, consts_start) | ("G", start, len)
    consts: List[int] = field(default_factory=list)
    consts_skip: List[bool] = field(default_factory=list)
    glue: List[Tuple] = field(default_factory=list)   # (kind, dst, src, c)
    zero_regs: List[int] = field(default_factory=list)
    parity_regs: List[int] = field(default_factory=list)
    register_count: int = 0
    mult_count: int = 0


class _Builder:
    def __init__(self):
        self.items: List[Tuple] = []
        self.next_reg = 0

    def alloc(self, n):
        r = list(range(self.next_reg, self.next_reg + n))
        self.next_reg += n
        return r

    def transform(self, inverse, regs, beta):
        if len(regs) > 1:
            self.items.append(("T", inverse, regs[0], len(regs), beta))

    def glue(self, op):
        self.items.append(("G", op))

    def interpolate(self, regs, beta, k):
        n = len(regs)
        if k == n:
            self.transform(True, regs, beta)
            return
        half = n // 2
        if k <= half:
            self.interpolate(regs[:half], beta, k)
            return
        self.transform(True, regs[:half], beta)
        c = int(what()[n.bit_length() - 2][beta])
        scratch = self.alloc(half)
        for j in range(half):
            self.glue((XO, scratch[j], regs[j], 0))
        self.transform(False, scratch, beta ^ half)
        for i in range(k - half):
            self.glue((XO, regs[half + i], scratch[i], 0))
        self.interpolate(regs[half:], beta ^ half, k - half)
        for j in range(k - half):
            if c == 1:
                self.glue((XO, regs[j], regs[half + j], 0))
            elif c:
                self.glue((MX, regs[j], regs[half + j], c))

    def fft_with_zero_tail(self, regs, beta, defined):
        size = len(regs)
        if defined >= size or size == 1:
            self.transform(False, regs, beta)
            return
        half = size // 2
        c = int(what()[size.bit_length() - 2][beta])
        for i in range(half):
            if c == 1:
                self.glue((XO, regs[i], regs[half + i], 0))
            elif c:
                self.glue((MX, regs[i], regs[half + i], c))
            self.glue((XO, regs[half + i], regs[i], 0))
        self.transform(False, regs[:half], beta)
        self.transform(False, regs[half:], beta ^ half)

    def folded_chunk(self, base, k, position, size):
        transform = len(base)
        tlevels = transform.bit_length() - 1
        chunk = self.alloc(size)
        shift = size.bit_length() - 1
        W = what()
        for j, reg in enumerate(chunk):
            for h in range(transform // size):
                index = h * size + j
                if index >= k:
                    continue
                factor = 1
                for level in range(shift, tlevels):
                    if index & (1 << level):
                        factor = int(tower_mul(factor, int(W[level][position])))
                if factor == 0:
                    continue
                if factor == 1:
                    self.glue((XO, reg, base[index], 0))
                else:
                    self.glue((MX, reg, base[index], factor))
        self.fft_with_zero_tail(chunk, position, min(k, size))
        return chunk


def _lower(b: _Builder, parity_regs: List[int], k: int, m: int) -> Program16:
    nreg = b.next_reg
    live = [False] * nreg
    for r in parity_regs:
        live[r] = True
    keep = [False] * len(b.items)
    live_out: List[Optional[List[bool]]] = [None] * len(b.items)
    for idx in range(len(b.items) - 1, -1, -1):
        it = b.items[idx]
        if it[0] == "G":
            _, dst, src, _c = it[1]
            if live[dst]:
                keep[idx] = True
                live[src] = True
        else:
            _, inverse, start, size, beta = it
            live_out[idx] = live[start:start + size]
            for r in range(start, start + size):
                live[r] = True
    is_zero = [True] * nreg
    for r in range(k):
        is_zero[r] = False
    needs_zero = [False] * nreg
    p = Program16(k=k, m=m, stages=[], register_count=nreg, parity_regs=list(parity_regs))
    open_glue = None
    W = what()
    for idx, it in enumerate(b.items):
        if it[0] == "G":
            if not keep[idx]:
                continue
            kind, dst, src, c = it[1]
            if is_zero[src]:
                continue
            if is_zero[dst]:
                is_zero[dst] = False
                kind = MC if kind == MX else CP
            if kind in (MX, MC):
                p.mult_count += 1
            if open_glue is None:
                open_glue = len(p.glue)
            p.glue.append((kind, dst, src, c))
        else:
            if open_glue is not None:
                p.stages.append(("G", open_glue, len(p.glue) - open_glue))
                open_glue = None
            _, inverse, start, size, beta = it
            for r in range(start, start + size):
                if is_zero[r]:
                    needs_zero[r] = True
                is_zero[r] = False
            alive = live_out[idx]
            cstart = len(p.consts)
            levels = size.bit_length() - 1
            for step in range(levels):
                level = step if inverse else levels - 1 - step
                stride = 1 << level
                for block in range(0, size, 2 * stride):
                    c = int(W[level][beta ^ block])
                    skip = not any(alive[block:block + 2 * stride])
                    p.consts.append(c)
                    p.consts_skip.append(skip)
                    if c and not skip:
                        p.mult_count += stride
            p.stages.append(("T", inverse, start, size, cstart))
    if open_glue is not None:
        p.stages.append(("G", open_glue, len(p.glue) - open_glue))
    p.zero_regs = [r for r in range(nreg) if needs_zero[r]]
    return p


def _build(k: int, m: int, strategy: str) -> Program16:
    n = k + m
    transform = 1
    while transform < k:
        transform <<= 1
    b = _Builder()
    base = b.alloc(transform)
    b.interpolate(base, 0, k)
    parity = [0] * m
    base_end = min(n, transform)
    if base_end > k and strategy == "window":
        size = 1
        while (k & ~(size - 1)) + size < base_end:
            size *= 2
        position = k & ~(size - 1)
        chunk = b.folded_chunk(base, k, position, size)
        for off, reg in enumerate(chunk):
            point = position + off
            if k <= point < n:
                parity[point - k] = reg
    position = transform
    while position < n:
        alignment = position & -position
        needed = n - position
        np2 = 1
        while np2 < needed:
            np2 <<= 1
        size = min(alignment, np2, transform)
        chunk = b.folded_chunk(base, k, position, size)
        for off, reg in enumerate(chunk):
            point = position + off
            if k <= point < n:
                parity[point - k] = reg
        position += size
    if base_end > k and strategy == "inplace":
        b.fft_with_zero_tail(base, 0, k)
        parity[:base_end - k] = base[k:base_end]
    return _lower(b, parity, k, m)


def build_staged_program16(k: int, m: int) -> Program16:
    """src/fft16.rs:386-406: the cheaper of the two base strategies."""
    n = k + m
    transform = 1
    while transform < k:
        transform <<= 1
    in_base = max(0, min(n, transform) - k)
    if in_base == 0:
        return _build(k, m, "window")
    w = _build(k, m, "window")
    ip = _build(k, m, "inplace")
    return w if w.mult_count < ip.mult_count else ip


def profitable(p: Program16, k: int, m: int) -> bool:
    """The reference's x86/GFNI routing rule (src/fft16.rs:85-106)."""
    if p.register_count > MAX_STAGED_REGISTERS16:
        return False
    return 9 * p.mult_count < 2 * k * m


def run_symbols(p: Program16, data: np.ndarray) -> np.ndarray:
    """Replays the program over symbol columns: data is (k, N) uint16 symbols;
    returns (m, N) parity symbols (src/fft16.rs:627-690 order)."""
    N = data.shape[1]
    R = np.full((p.register_count, N), 0xA5A5, np.int64)  # stale on purpose
    R[:p.k] = data
    for r in p.zero_regs:
        R[r] = 0
    for st in p.stages:
        if st[0] == "G":
            _, start, ln = st
            for kind, dst, src, c in p.glue[start:start + ln]:
                if kind == XO:
                    R[dst] ^= R[src]
                elif kind == CP:
                    R[dst] = R[src]
                elif kind == MX:
                    R[dst] ^= tower_mul(c, R[src])
                else:
                    R[dst] = tower_mul(c, R[src])
        else:
            _, inverse, start, size, cstart = st
            levels = size.bit_length() - 1
            ci = cstart
            for step in range(levels):
                level = step if inverse else levels - 1 - step
                stride = 1 << level
                for block in range(0, size, 2 * stride):
                    c, skip = p.consts[ci], p.consts_skip[ci]
                    ci += 1
                    if skip:
                        continue
                    lo = slice(start + block, start + block + stride)
                    hi = slice(start + block + stride, start + block + 2 * stride)
                    if inverse:
                        R[hi] ^= R[lo]
                        R[lo] ^= tower_mul(c, R[hi])
                    else:
                        R[lo] ^= tower_mul(c, R[hi])
                        R[hi] ^= R[lo]
    return R[p.parity_regs].astype(np.uint16)
