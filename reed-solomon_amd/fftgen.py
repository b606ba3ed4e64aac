"""Lin-Chung-Han additive-FFT encode programs for GF(2^8) (build-time generator).

Restates the reference's program builder (src/fft.rs:150-716): interpolate the
data polynomial through the points 0..k-1 in the LCH novel basis over the
standard basis {1,2,..,128}, then evaluate it at the parity points k..n-1.
Because the evaluation point of index i is the byte i, the parity equals the
systematic generator's parity byte for byte (src/fft.rs:1-8).

The builder emits transform items (whole FFT / IFFT blocks over contiguous
register ranges) and glue xors; flattening follows the reference's order
contract (forward transforms walk levels high to low, inverse low to high;
blocks in ascending base order; butterflies in ascending index order,
src/fft.rs:17-21).  Zero-state normalisation, backward dead-code elimination
and register reuse follow src/fft.rs:304-385, :594-716, so the op lists equal
the reference's generated programs (src/fft_programs.rs) -- tests compare
their SHA-256 with hashes taken from the reference.

The op list is data for the GPU code generator (gf8_codegen.py), which turns
every op into straight-line bit-sliced XORs.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Tuple

LEVELS = 8


def _gf8_mul_table() -> List[List[int]]:
    log = [0] * 256
    exp = [0] * 510
    b = 1
    for i in range(255):
        log[b] = i
        exp[i] = exp[i + 255] = b
        b <<= 1
        if b & 0x100:
            b ^= 0x11D
    return [[exp[log[a] + log[c]] if a and c else 0 for c in range(256)] for a in range(256)]


MUL = _gf8_mul_table()


def _what_tables() -> List[List[int]]:
    """Normalised subspace polynomials: W_0(x) = x, W_{j+1}(x) = W_j(x)(W_j(x) + W_j(2^j)),
    each level scaled so its value at 2^j is 1 (src/fft.rs:33-79)."""
    w = [list(range(256))]
    for j in range(LEVELS - 1):
        piv = w[j][1 << j]
        w.append([MUL[v][v ^ piv] for v in w[j]])
    out = []
    for j in range(LEVELS):
        piv = w[j][1 << j]
        inv = next(c for c in range(1, 256) if MUL[piv][c] == 1)
        out.append([MUL[v][inv] for v in w[j]])
    return out


WHAT = _what_tables()

# op kinds
MX, XO, MC, CP = "mx", "xo", "mc", "cp"


@dataclass
class Program:
    k: int
    m: int
    ops: List[Tuple]            # (kind, dst, src[, c])
    parity_regs: List[int]
    register_count: int
    mult_count: int

    def text_lines(self) -> List[str]:
        """The program in the reference's token form (src/macros.rs:16-46)."""
        lines = [f"ld {s} {s}" for s in range(self.k)]
        for op in self.ops:
            lines.append(" ".join(str(x) for x in op))
        lines += [f"st {p} {r}" for p, r in enumerate(self.parity_regs)]
        return lines


@dataclass
class _Builder:
    items: List[Tuple] = field(default_factory=list)  # ("T", inverse, start, size, beta) | ("G", op)
    next_reg: int = 0

    def alloc(self, n: int) -> List[int]:
        r = list(range(self.next_reg, self.next_reg + n))
        self.next_reg += n
        return r

    def transform(self, inverse: bool, regs: List[int], beta: int) -> None:
        if len(regs) > 1:
            self.items.append(("T", inverse, regs[0], len(regs), beta))

    def glue(self, op: Tuple) -> None:
        self.items.append(("G", op))

    def interpolate(self, regs: List[int], beta: int, k: int) -> None:
        """Coefficients of the degree<k polynomial through regs[0..k] on
        beta+{0..k-1}; registers above k must be zero (src/fft.rs:194-239)."""
        n = len(regs)
        if k == n:
            self.transform(True, regs, beta)
            return
        half = n // 2
        if k <= half:
            self.interpolate(regs[:half], beta, k)
            return
        self.transform(True, regs[:half], beta)
        c = WHAT[(n.bit_length() - 1) - 1][beta]
        scratch = self.alloc(half)
        for j in range(half):
            self.glue((XO, scratch[j], regs[j]))
        self.transform(False, scratch, beta ^ half)
        for i in range(k - half):
            self.glue((XO, regs[half + i], scratch[i]))
        self.interpolate(regs[half:], beta ^ half, k - half)
        for j in range(k - half):
            if c == 1:
                self.glue((XO, regs[j], regs[half + j]))
            elif c != 0:
                self.glue((MX, regs[j], regs[half + j], c))


def _flatten(items) -> List[Tuple]:
    ops = []
    for it in items:
        if it[0] == "G":
            ops.append(it[1])
            continue
        _, inverse, start, size, beta = it
        levels = size.bit_length() - 1
        order = range(levels) if inverse else range(levels - 1, -1, -1)
        for level in order:
            stride = 1 << level
            for base in range(0, size, 2 * stride):
                c = WHAT[level][beta ^ base]
                for i in range(base, base + stride):
                    lo, hi = start + i, start + i + stride
                    if inverse:
                        ops.append((XO, hi, lo))
                        if c == 1:
                            ops.append((XO, lo, hi))
                        elif c:
                            ops.append((MX, lo, hi, c))
                    else:
                        if c == 1:
                            ops.append((XO, lo, hi))
                        elif c:
                            ops.append((MX, lo, hi, c))
                        ops.append((XO, hi, lo))
    return ops


def _normalise(ops, nreg: int, k: int) -> List[Tuple]:
    zero = [True] * nreg
    for i in range(k):
        zero[i] = False
    out = []
    for op in ops:
        kind, dst, src = op[0], op[1], op[2]
        if zero[src]:
            continue
        if zero[dst]:
            zero[dst] = False
            out.append((MC, dst, src, op[3]) if kind == MX else (CP, dst, src))
        else:
            out.append(op)
    return out


def _dce(ops, parity_regs, nreg: int) -> List[Tuple]:
    need = [False] * nreg
    for r in parity_regs:
        need[r] = True
    kept = []
    for op in reversed(ops):
        kind, dst, src = op[0], op[1], op[2]
        if not need[dst]:
            continue
        if kind in (MC, CP):
            need[dst] = False
        need[src] = True
        kept.append(op)
    kept.reverse()
    return kept


def _remap(ops, parity_regs, nreg: int, k: int):
    last = [0] * nreg
    for pos, op in enumerate(ops):
        last[op[1]] = pos
        last[op[2]] = pos
    for r in parity_regs:
        last[r] = len(ops)
    mapping = [None] * nreg
    for r in range(k):
        mapping[r] = r
    expires = [[] for _ in range(len(ops) + 1)]
    for v, pos in enumerate(last):
        if mapping[v] is None:
            expires[pos].append(v)
    for r in range(k):
        if last[r] < len(ops):
            expires[last[r]].append(r)
    free: List[int] = []
    fresh = k

    def assign(v):
        nonlocal fresh
        if mapping[v] is None:
            if free:
                mapping[v] = free.pop()
            else:
                mapping[v] = fresh
                fresh += 1

    out = []
    for pos, op in enumerate(ops):
        assign(op[2])
        assign(op[1])
        out.append((op[0], mapping[op[1]], mapping[op[2]]) + tuple(op[3:]))
        for v in expires[pos]:
            if mapping[v] is not None:
                free.append(mapping[v])
    return out, [mapping[r] for r in parity_regs], fresh


def build_encode_program(k: int, m: int) -> Program:
    """Flat encode program (src/fft.rs:304-385)."""
    n = k + m
    t = 1
    while t < k:
        t <<= 1
    b = _Builder()
    base = b.alloc(t)
    b.interpolate(base, 0, k)
    parity = [0] * m
    block_base = t
    while block_base < n:
        blk = b.alloc(t)
        for j in range(min(k, t)):
            b.glue((XO, blk[j], base[j]))
        b.transform(False, blk, block_base)
        for off, reg in enumerate(blk):
            point = block_base + off
            if k <= point < n:
                parity[point - k] = reg
        block_base += t
    b.transform(False, base, 0)
    in_base = max(0, min(n, t) - k)
    parity[:in_base] = base[k:k + in_base]
    flat = _flatten(b.items)
    norm = _normalise(flat, b.next_reg, k)
    kept = _dce(norm, parity, b.next_reg)
    mults = sum(1 for op in kept if op[0] in (MX, MC))
    ops, preg, nreg = _remap(kept, parity, b.next_reg, k)
    return Program(k=k, m=m, ops=ops, parity_regs=preg, register_count=nreg, mult_count=mults)


def run_program_bytes(prog: Program, data: List[int]) -> List[int]:
    """Replays the program over one byte column (for tests)."""
    r = [0] * prog.register_count
    r[:len(data)] = data
    for op in prog.ops:
        kind, d, s = op[0], op[1], op[2]
        if kind == MX:
            r[d] ^= MUL[op[3]][r[s]]
        elif kind == XO:
            r[d] ^= r[s]
        elif kind == MC:
            r[d] = MUL[op[3]][r[s]]
        else:
            r[d] = r[s]
    return [r[p] for p in prog.parity_regs]


GENERATED_SHAPES = [(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32)]
