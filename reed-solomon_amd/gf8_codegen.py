"""Generates the bit-sliced GF(2^8) encode kernels (build step).

Bit-sliced form: a lane loads 32 bytes of a shard (two 16-byte pieces 1 KiB
apart, so each wave-wide load covers 1 KiB contiguous; bitslice.hpp) and
transposes them into 8 planes, plane b holding bit b of all 32 bytes.
Multiplying by a constant c is then the 8x8 GF(2) matrix of "x -> c*x"
(the same matrix GFNI's affine instruction takes, src/gf/gfni.rs:41-64):
output plane r is the XOR of the input planes b whose product c*2^b has bit r
set.  With c known at compile time every multiply becomes a short list of
plane XORs (3-input XORs on gfx950), and field additions are plane XORs.

Two programs per shape are emitted, both byte-identical to the generator
matrix encode:
  * fft  -- the LCH additive-FFT op list from fftgen.build_encode_program
            (the reference's generated programs, src/fft_programs.rs);
  * mat  -- the systematic generator applied row by row with constant
            coefficients (src/reedsolomon.rs:154-184).

Usage: python gf8_codegen.py OUT.hip
"""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import fftgen  # noqa: E402

MUL = fftgen.MUL

# Shapes with generated kernels: the six generated FFT shapes and (10,4).
SHAPES = [(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32), (10, 4)]
FFT_SHAPES = set(fftgen.GENERATED_SHAPES)
# Constant-matrix programs keep all k input planes live across every output;
# past ~20 inputs they exceed the register file, so only these get one.
MAT_SHAPES = {(7, 13), (10, 10), (18, 6), (10, 4)}


def bit_rows(c: int):
    """rows[r] = input planes b feeding output plane r of x -> c*x."""
    return [[b for b in range(8) if (MUL[c][1 << b] >> r) & 1] for r in range(8)]


def _gf8_generator(k: int, n: int):
    """Systematic generator V(n,k) * inv(V_top) over GF(2^8) (src/reedsolomon.rs:154-159)."""
    inv8 = [0] * 256
    for a in range(1, 256):
        for b in range(1, 256):
            if MUL[a][b] == 1:
                inv8[a] = b
                break
    V = [[0] * k for _ in range(n)]
    for r in range(n):
        p = 1
        for c in range(k):
            V[r][c] = p
            p = MUL[p][r]
    # invert top k x k
    A = [V[r][:] + [1 if i == r else 0 for i in range(k)] for r in range(k)]
    for c in range(k):
        piv = next(r for r in range(c, k) if A[r][c])
        A[c], A[piv] = A[piv], A[c]
        s = inv8[A[c][c]]
        A[c] = [MUL[s][x] for x in A[c]]
        for r in range(k):
            if r != c and A[r][c]:
                f = A[r][c]
                A[r] = [x ^ MUL[f][y] for x, y in zip(A[r], A[c])]
    inv = [row[k:] for row in A]
    G = [[0] * k for _ in range(n)]
    for r in range(n):
        for c in range(k):
            acc = 0
            for i in range(k):
                acc ^= MUL[V[r][i]][inv[i][c]]
            G[r][c] = acc
    return G


def matrix_program(k: int, m: int):
    """Row-by-row constant-coefficient program: ops plus interleaved stores."""
    G = _gf8_generator(k, k + m)
    ops = []
    acc = k
    for j in range(m):
        first = True
        for i in range(k):
            c = G[k + j][i]
            if c == 0:
                continue
            ops.append(("mc" if first else "mx", acc, i, c))
            first = False
        if first:
            ops.append(("zero", acc))
        ops.append(("st", j, acc))
    return ops, k + 1


def fft_program(k: int, m: int):
    p = fftgen.build_encode_program(k, m)
    ops = list(p.ops) + [("st", j, r) for j, r in enumerate(p.parity_regs)]
    return ops, p.register_count


def _dispatch(name: str, ins: str, len_: str, vec: str, ind: str, verify: bool = False) -> str:
    """One branch per tile on its access mode (bitslice.hpp, split_mode): the
    full and masked paths get straight-line copies of the program.  `verify`
    compares the outputs with the shards instead of storing them."""
    v = 'true' if verify else 'false'
    return "\n".join([
        f"{ind}const int mode = split_mode({vec}, tile, {len_});",
        f"{ind}uint32_t bad = 0;",
        f"{ind}if (mode == kSplitFull) {name}<{v}, true>({ins}, out, off, {len_}, kSplitFull, bad);",
        f"{ind}else if (mode == kSplitMasked) {name}<{v}, true>({ins}, out, off, {len_}, kSplitMasked, "
        "bad);",
        f"{ind}else {name}<{v}, false>({ins}, out, off, {len_}, mode, bad);",
    ])


# Code shape of the 16-byte access paths (FAST; the byte-access paths keep
# the plain form, so their byte loads do not raise the kernels' register
# counts):
#   LOAD_BLOCK  every input's loads are issued before the first transpose (a
#               scheduling barrier between), so a tile's loads are in flight
#               together; left to itself the compiler interleaves loads with
#               transposes and waits after every few (at one wave per SIMD,
#               (32,32), one input at a time);
#   OPAQUE      each op's result passes through an empty asm statement, so the
#               compiler cannot reassociate XOR terms across ops and keep the
#               shared partial sums live (232 -> 199 VGPRs for (16,16));
#   SCHED_GROUP a scheduling barrier every SCHED_GROUP ops (0 = none), so the
#               schedule follows the program's own order.
# Measured together on batched encode: 4670-4732 -> 4949-4993 GB/s.
LOAD_BLOCK = int(os.environ.get("TRS_GEN_LOAD_BLOCK", "1"))
OPAQUE = int(os.environ.get("TRS_GEN_OPAQUE", "1"))
SCHED_GROUP = int(os.environ.get("TRS_GEN_SCHED_GROUP", "1"))


# Programs whose 16-byte paths transpose with the select form of the bit swap
# (bitslice.hpp: 48 instead of 60 operations per 32 bytes, one more live
# register).  It pays only where VALU bounds the program at one wave per
# SIMD: (32,32) 2322 -> 2234 us per batched launch, while (7,13) and (16,16)
# were 3-9 % slower with it.
SELECT_SWAP = {("fft", 32, 32)}


def emit_body(k: int, ops, nreg: int, sel: bool = False) -> str:
    out = []
    fs = "FAST" if sel else "false"
    out.append(f"  uint32_t R[{nreg}][8];")
    if LOAD_BLOCK:
        out.append("  if constexpr (FAST) {")
        for s in range(k):
            out.append(f"    load_split_raw<true>(ld_const_ptr(in + {s}), off, len, mode, R[{s}]);")
        out.append("    __builtin_amdgcn_sched_barrier(0);")
        for s in range(k):
            out.append(f"    bytes_to_planes<{fs}>(R[{s}]);")
        out.append("  } else {")
        for s in range(k):
            out.append(f"    load_split<true>(ld_const_ptr(in + {s}), off, len, mode, R[{s}]);")
        out.append("  }")
    else:
        for s in range(k):
            out.append(f"  load_split<true>(ld_const_ptr(in + {s}), off, len, mode, R[{s}]);")
    for i, op in enumerate(ops):
        if SCHED_GROUP and i % SCHED_GROUP == 0:
            # verify kernels keep the order on the byte paths too: their
            # spills drop from 647 / 1733 to 4 / 23 VGPRs for (14,14) /
            # (16,16) (a verify call's time, bound by waiting for its flag,
            # did not change).  Encode kernels keep the plain byte paths,
            # which spill less there ((18,6) 8 against 35 VGPRs).
            out.append("  if constexpr (FAST || V) __builtin_amdgcn_sched_barrier(0);")
        kind = op[0]
        if kind == "st":
            _, j, r = op
            out.append(f"  if constexpr (V) verify_split<true>(ld_const_ptr(out + {j}), off, len, "
                       f"mode, R[{r}], bad);")
            out.append(f"  else store_split<true, {fs}>(ld_const_ptr(out + {j}), off, len, mode, "
                       f"R[{r}]);")
        elif kind == "zero":
            out.append(f"  for (int r = 0; r < 8; ++r) R[{op[1]}][r] = 0;")
        elif kind in ("xo", "cp"):
            _, d, s = op
            sym = "^=" if kind == "xo" else "="
            for r in range(8):
                out.append(f"  R[{d}][{r}] {sym} R[{s}][{r}];")
        else:
            _, d, s, c = op
            rows = bit_rows(c)
            for r in range(8):
                terms = " ^ ".join(f"R[{s}][{b}]" for b in rows[r])
                if kind == "mx":
                    if terms:
                        out.append(f"  R[{d}][{r}] ^= {terms};")
                else:
                    out.append(f"  R[{d}][{r}] = {terms if terms else '0u'};")
        if OPAQUE and kind in ("xo", "mx", "mc"):
            out.append(f"  if constexpr (FAST || V) opaque(R[{op[1]}]);")
    return "\n".join(out)


# Work-list programs that run as two half programs on a pair of waves.  An
# LCH encode is an IFFT over the k data registers (butterfly strides 1, 2, ...,
# k/2) followed by an FFT (strides k/2, ..., 1): only the two stride-k/2 levels
# pair a register of the lower half with one of the upper half.  Together they
# map each pair (a, b) = (R[j], R[j + k/2]) to (a ^ c t, b ^ c t) with
# t = a ^ b, the same constant c on both sides.  So one wave holds registers
# 0..k/2-1 and the other k/2..k-1, each loads its k/2 data shards, runs its
# half of the IFFT, swaps its registers with its partner through LDS, applies
# the pair map, runs its half of the FFT and stores its k/2 parity shards.
# Each wave then keeps k/2 + 1 registers live instead of k + 1: for (32,32),
# two waves per SIMD instead of one.
SPLIT = {("fft", 32, 32)} if int(os.environ.get("TRS_GEN_SPLIT", "1")) else set()
# Waves per SIMD a split program's kernels are compiled for.  (16,16) splits
# too (125-127 VGPRs, four waves per SIMD), but its work-list launch measured
# 709 -> 901 us, so it keeps the single program at two waves per SIMD.
SPLIT_WAVES = {("fft", 32, 32): 2}


def split_program(ops, nreg: int, k: int):
    """Halves of a generated encode program for a wave pair.

    Returns (pre, consts, post): pre[h] and post[h] are half h's ops before
    and after the exchange, on local registers (register j of half h is
    program register j + h*k/2); consts[j] is the pair map constant of
    (R[j], R[j + k/2]).  Raises ValueError when the program does not have
    that form.  The copies of the program (register renames between the IFFT
    and the FFT) are folded into register names first."""
    name = list(range(nreg))
    prog = []
    for op in ops:
        kind = op[0]
        if kind == "cp":
            name[op[1]] = name[op[2]]
        elif kind == "st":
            prog.append(("st", op[1], name[op[2]]))
        elif kind == "xo":
            prog.append(("xo", name[op[1]], name[op[2]]))
        elif kind in ("mx", "mc"):
            prog.append((kind, name[op[1]], name[op[2]], op[3]))
        else:
            raise ValueError(f"op {op} in a split program")
    h = k // 2
    if k % 2 or any(r >= k for op in prog for r in ((op[2],) if op[0] == "st" else op[1:3])):
        raise ValueError("registers beyond the data registers")
    side = lambda r: r // h  # noqa: E731
    cross = [i for i, op in enumerate(prog) if op[0] != "st" and side(op[1]) != side(op[2])]
    if not cross:
        raise ValueError("no cross-half ops")
    first, last = cross[0], cross[-1]
    mid = prog[first:last + 1]
    if any(op[0] == "st" or side(op[1]) == side(op[2]) for op in mid):
        raise ValueError("cross-half ops are not one block")
    consts = []
    for j in range(h):
        v = {j: (1, 0), j + h: (0, 1)}  # each value as (coefficient of a, of b)
        for op in mid:
            d, s = op[1], op[2]
            if d not in v and s not in v:
                continue
            if {d, s} != {j, j + h}:
                raise ValueError("cross-half ops do not pair j with j + k/2")
            if op[0] == "xo":
                v[d] = (v[d][0] ^ v[s][0], v[d][1] ^ v[s][1])
            elif op[0] == "mx":
                v[d] = (v[d][0] ^ MUL[op[3]][v[s][0]], v[d][1] ^ MUL[op[3]][v[s][1]])
            else:
                raise ValueError("overwriting op among the cross-half ops")
        (p, q), (r, s) = v[j], v[j + h]
        if not (q == r and p == 1 ^ q and s == 1 ^ q):
            raise ValueError(f"pair {j} map is not (a ^ c t, b ^ c t)")
        consts.append(q)

    def local(seg, hh):
        out = []
        for op in seg:
            own = op[2] if op[0] == "st" else op[1]
            if side(own) != hh:
                continue
            if op[0] == "st":
                out.append(("st", op[1], op[2] - hh * h))
            else:
                out.append((op[0], op[1] - hh * h, op[2] - hh * h) + tuple(op[3:]))
        return out

    pre = [local(prog[:first], hh) for hh in (0, 1)]
    post = [local(prog[last + 1:], hh) for hh in (0, 1)]
    return pre, consts, post


def _split_op_lines(op, sel: str) -> list:
    kind = op[0]
    if kind == "st":
        return [f"  if constexpr (V) {{ if (ALL || act) verify_split<true>(ld_const_ptr(out + {op[1]}), "
                f"off, len, mode, R[{op[2]}], bad); }}",
                f"  else if (ALL || act) store_split<true, {sel}>(ld_const_ptr(out + {op[1]}), off, len, "
                f"mode, R[{op[2]}]);"]
    lines = []
    if kind == "xo":
        for r in range(8):
            lines.append(f"  R[{op[1]}][{r}] ^= R[{op[2]}][{r}];")
    else:
        _, d, s, c = op
        rows = bit_rows(c)
        for r in range(8):
            terms = " ^ ".join(f"R[{s}][{b}]" for b in rows[r])
            if kind == "mx":
                if terms:
                    lines.append(f"  R[{d}][{r}] ^= {terms};")
            else:
                lines.append(f"  R[{d}][{r}] = {terms if terms else '0u'};")
    lines.append("  opaque(R[" + str(op[1]) + "]);")
    return lines


def emit_split(name: str, k: int, ops, nreg: int, sel: bool, waves: int = 2) -> str:
    """Device functions and the work-list kernel of a split program."""
    pre, consts, post = split_program(ops, nreg, k)
    h = k // 2
    rounds = 2  # registers cross LDS in two rounds of h/2: 32 KiB per wave pair
    per = h // rounds
    fs = "FAST" if sel else "false"
    out = []
    for hh in (0, 1):
        out.append(f"// half {hh}: data shards {hh * h}..{hh * h + h - 1}, its half of the IFFT")
        out.append(f"template <bool FAST, bool ALL> __device__ __forceinline__ void {name}_pre{hh}("
                   "const uint8_t *const *in, uint64_t off, uint64_t len, int mode, bool act, "
                   f"uint32_t (&R)[{h}][8]) {{")
        out.append("  if (ALL || act) {")
        out.append("    if constexpr (FAST) {")
        for s in range(h):
            out.append(f"      load_split_raw<true>(ld_const_ptr(in + {hh * h + s}), off, len, mode, "
                       f"R[{s}]);")
        out.append("    } else {")
        for s in range(h):  # one input at a time: byte loads are not hoisted
            out.append(f"      load_split<true>(ld_const_ptr(in + {hh * h + s}), off, len, mode, "
                       f"R[{s}]);")
            out.append("      __builtin_amdgcn_sched_barrier(0);")
        out.append("    }")
        out.append("  } else {")
        out.append(f"    for (int s = 0; s < {h}; ++s)")
        out.append("      for (int r = 0; r < 8; ++r) R[s][r] = 0;")
        out.append("  }")
        out.append("  if constexpr (FAST) {")
        out.append("    __builtin_amdgcn_sched_barrier(0);")
        for s in range(h):
            out.append(f"    bytes_to_planes<{fs}>(R[{s}]);")
        out.append("  }")
        # every path keeps the program's order: at k/2 + 1 live registers per
        # wave the budget is 256 VGPRs, which reassociated XORs overrun
        for op in pre[hh]:
            out.append("  __builtin_amdgcn_sched_barrier(0);")
            out.extend(_split_op_lines(op, fs))
        out.append("}")
        out.append(f"// half {hh}: its half of the FFT and parity shards")
        out.append(f"template <bool V, bool FAST, bool ALL> __device__ __forceinline__ void "
                   f"{name}_post{hh}(uint8_t *const *out, uint64_t off, uint64_t len, int mode, bool act, "
                   f"uint32_t (&R)[{h}][8], uint32_t &bad) {{")
        for op in post[hh]:
            out.append("  __builtin_amdgcn_sched_barrier(0);")
            out.extend(_split_op_lines(op, fs))
        out.append("}")
    out.append("// The pair map of the two stride-k/2 levels: each wave publishes its")
    out.append("// registers, reads its partner's and adds c (own ^ partner) to its own.")
    out.append("// X holds one round of both waves' registers; the barrier after a round's")
    out.append("// reads keeps the next round (or tile) from overwriting them early.")
    out.append(f"__device__ __forceinline__ void {name}_exchange(uint32_t (&R)[{h}][8], "
               f"uint4 (*X)[{per}][2][64], uint32_t half, uint32_t lane) {{")
    for rnd in range(rounds):
        for r in range(per):
            j = rnd * per + r
            out.append(f"  X[half][{r}][0][lane] = make_uint4(R[{j}][0], R[{j}][1], R[{j}][2], "
                       f"R[{j}][3]);")
            out.append(f"  X[half][{r}][1][lane] = make_uint4(R[{j}][4], R[{j}][5], R[{j}][6], "
                       f"R[{j}][7]);")
        out.append("  __syncthreads();")
        for r in range(per):
            j = rnd * per + r
            rows = bit_rows(consts[j])
            out.append("  {")
            out.append(f"    const uint4 p0 = X[half ^ 1u][{r}][0][lane], p1 = X[half ^ 1u][{r}][1][lane];")
            out.append(f"    const uint32_t t[8] = {{R[{j}][0] ^ p0.x, R[{j}][1] ^ p0.y, R[{j}][2] ^ p0.z, "
                       f"R[{j}][3] ^ p0.w, R[{j}][4] ^ p1.x, R[{j}][5] ^ p1.y, R[{j}][6] ^ p1.z, "
                       f"R[{j}][7] ^ p1.w}};")
            for row in range(8):
                if rows[row]:
                    terms = " ^ ".join(f"t[{b}]" for b in rows[row])
                    out.append(f"    R[{j}][{row}] ^= {terms};")
            out.append("  }")
            out.append("  __builtin_amdgcn_sched_barrier(0);")  # one register's reads at a time
        out.append("  __syncthreads();")
    out.append("}")
    out.append(f"template <uint32_t H, bool V, bool FAST, bool ALL> __device__ __forceinline__ void "
               f"{name}_pair(const uint8_t *const *in, uint8_t *const *out, uint64_t off, uint64_t len, "
               f"int mode, bool act, uint32_t lane, uint4 (*X)[{per}][2][64], uint32_t &bad) {{")
    out.append(f"  uint32_t R[{h}][8];")
    out.append(f"  if constexpr (H == 0) {name}_pre0<FAST, ALL>(in, off, len, mode, act, R);")
    out.append(f"  else {name}_pre1<FAST, ALL>(in, off, len, mode, act, R);")
    out.append(f"  {name}_exchange(R, X, H, lane);")
    out.append(f"  if constexpr (H == 0) {name}_post0<V, FAST, ALL>(out, off, len, mode, act, R, bad);")
    out.append(f"  else {name}_post1<V, FAST, ALL>(out, off, len, mode, act, R, bad);")
    out.append("}")
    # The work list's tiles stay the codec's 8 KiB tiles; a workgroup (one
    # wave pair) runs a quarter of one, sub-tile u = 4 t + q covering bytes
    # q*2048 .. q*2048 + 2047 of tile t.  Each wave runs its own copy of the
    # loop (its half's registers are all that is live in it); both copies
    # take the same branches, which depend only on the workgroup's sub-tile,
    # and so meet at the same barriers.  Lanes past the shard end load zeros
    # and store nothing.
    out.append(f"template <uint32_t H> __device__ __forceinline__ void {name}_run(const BatchJob *jobs, "
               f"const uint32_t *tile_job, uint64_t total, uint32_t lane, uint4 (*X)[{per}][2][64]) {{")
    out.append("  for (uint64_t u = blockIdx.x; u < total * 4; u += gridDim.x) {")
    out.append("    const uint64_t t = u >> 2;")
    out.append("    const uint32_t q = uint32_t(u & 3);")
    out.append("    const auto &J = ld_const_ref(jobs + ld_const(tile_job + t));")
    out.append("    const uint64_t len = J.len;")
    out.append("    const uint64_t tile = (t - J.tile_begin) * 8192;")
    out.append("    if (tile + uint64_t(q) * 2048 >= len) continue;")
    out.append("    const uint64_t off = tile + split_lane_offset(q * 64 + lane);")
    out.append("    const bool act = off < len;")
    out.append("    const uint8_t *const *in = J.ptrs;")
    out.append(f"    uint8_t *const *out = const_cast<uint8_t *const *>(in + {k});")
    out.append("    const int mode = split_mode(J.vec_ok != 0, tile, len);")
    out.append("    uint32_t bad = 0;")
    out.append(f"    if (mode == kSplitFull) {name}_pair<H, false, true, true>(in, out, off, len, mode, act, "
               "lane, X, bad);")
    out.append(f"    else if (mode == kSplitMasked) {name}_pair<H, false, true, false>(in, out, off, len, mode, "
               "act, lane, X, bad);")
    out.append(f"    else {name}_pair<H, false, false, false>(in, out, off, len, mode, act, lane, X, bad);")
    out.append("  }")
    out.append("}")
    out.append(f"__global__ __launch_bounds__(kPairThreads) __attribute__((amdgpu_waves_per_eu({waves}))) "
               f"void k_batch_{name}(const BatchJob *jobs,")
    out.append("    const uint32_t *tile_job, uint64_t total) {")
    out.append(f"  __shared__ uint4 X[2][{per}][2][64];")
    out.append("  const uint32_t lane = threadIdx.x & 63;")
    out.append("  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0)")
    out.append(f"    {name}_run<0>(jobs, tile_job, total, lane, X);")
    out.append("  else")
    out.append(f"    {name}_run<1>(jobs, tile_job, total, lane, X);")
    out.append("}")
    # Per-call kernels: workgroup b runs window b & 3 of tile b >> 2 (grid =
    # one workgroup per 2 KiB of shard); a window past the end returns in
    # both waves before any barrier.  The verify form compares each half's
    # parity with the shards (ReedSolomon::verify) and sets the flag.
    out.append(f"template <uint32_t H, bool V> __device__ __forceinline__ void {name}_call(const Gf8Call &a, "
               f"uint32_t lane, uint4 (*X)[{per}][2][64]) {{")
    out.append("  const uint64_t tile = uint64_t(blockIdx.x >> 2) * 8192;")
    out.append("  const uint32_t q = blockIdx.x & 3;")
    out.append("  if (tile + uint64_t(q) * 2048 >= a.len) return;")
    out.append("  const uint64_t off = tile + split_lane_offset(q * 64 + lane);")
    out.append("  const bool act = off < a.len;")
    out.append(f"  uint8_t *const *out = const_cast<uint8_t *const *>(a.ptr + {k});")
    out.append("  const int mode = split_mode(a.vec_ok != 0, tile, a.len);")
    out.append("  uint32_t bad = 0;")
    out.append(f"  if (mode == kSplitFull) {name}_pair<H, V, true, true>(a.ptr, out, off, a.len, mode, act, "
               "lane, X, bad);")
    out.append(f"  else if (mode == kSplitMasked) {name}_pair<H, V, true, false>(a.ptr, out, off, a.len, mode, "
               "act, lane, X, bad);")
    out.append(f"  else {name}_pair<H, V, false, false>(a.ptr, out, off, a.len, mode, act, lane, X, bad);")
    out.append("  if constexpr (V)")
    out.append("    if (bad) *a.flag = 1u;")
    out.append("}")
    for suffix, v in (("", "false"), ("_verify", "true")):
        out.append(f"__global__ __launch_bounds__(kPairThreads) __attribute__((amdgpu_waves_per_eu({waves}))) "
                   f"void k_{name}{suffix}(const Gf8Call a) {{")
        out.append(f"  __shared__ uint4 X[2][{per}][2][64];")
        out.append("  const uint32_t lane = threadIdx.x & 63;")
        out.append("  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0)")
        out.append(f"    {name}_call<0, {v}>(a, lane, X);")
        out.append("  else")
        out.append(f"    {name}_call<1, {v}>(a, lane, X);")
        out.append("}")
    return "\n".join(out)


def xor_cost(ops) -> int:
    """VALU estimate per 32-byte lane chunk: 3-input XORs, excluding transposes."""
    n = 0
    for op in ops:
        if op[0] in ("xo",):
            n += 8
        elif op[0] in ("mx", "mc"):
            for row in bit_rows(op[3]):
                t = len(row) + (1 if op[0] == "mx" else 0)
                n += max(0, (t - 1 + 1) // 2)
    return n


# Waves per SIMD the kernels are compiled for (the register budget is
# 512 / waves VGPRs).  Two for every program but these, which need more than
# 256 VGPRs without spilling: (32,32) fft as one program (without its wave
# pairs, SPLIT) keeps ~60 plane registers live, and
# the compiler interleaves the output rows of the wide constant-matrix
# programs.  With the full/tail branch taken once per tile the compiler
# issues loads early, which raises register pressure: unbounded, (18,6) fft
# fell to one wave.  Three waves spilled to scratch and measured slower
# ((10,4): 3.67 -> 3.17 TB/s).
ONE_WAVE = {("fft", 32, 32), ("mat", 7, 13), ("mat", 10, 10), ("mat", 18, 6)}


def _waves_attr(route: str, k: int, m: int) -> str:
    if (route, k, m) in ONE_WAVE:
        return ""
    return "__attribute__((amdgpu_waves_per_eu(2))) "


def generate() -> str:
    parts = [
        "// GENERATED by reed-solomon_amd/gf8_codegen.py -- do not edit.",
        "// Bit-sliced GF(2^8) encode programs with compile-time coefficients.",
        '#include "bitslice.hpp"',
        '#include "uniform.hpp"',
        "",
        "namespace trs {",
        "namespace {",
        "",
        "// The register's planes pass through an empty asm statement: the compiler",
        "// no longer sees how they were computed, so it cannot reassociate and share",
        "// XOR terms across program ops.",
        "__device__ __forceinline__ void opaque(uint32_t (&x)[8]) {",
        "#pragma unroll",
        "  for (int r = 0; r < 8; ++r) asm(\"\" : \"+v\"(x[r]));",
        "}",
        "",
        "// Threads of a split program's workgroup: one wave per half.",
        "[[maybe_unused]] constexpr int kPairThreads = 128;",
        "",
    ]
    table = []
    for (k, m) in SHAPES:
        routes = []
        if (k, m) in MAT_SHAPES:
            routes.append(("mat", matrix_program(k, m)))
        if (k, m) in FFT_SHAPES:
            routes.append(("fft", fft_program(k, m)))
        for route, (ops, nreg) in routes:
            name = f"bs_{route}_{k}_{m}"
            parts.append(f"// ({k},{m}) {route}: {nreg} registers, ~{xor_cost(ops)} plane XORs "
                         "per 32-byte lane chunk")
            if (route, k, m) in SPLIT:  # per-call, verify and work-list kernels on wave pairs
                parts.append(emit_split(name, k, ops, nreg, (route, k, m) in SELECT_SWAP,
                                        SPLIT_WAVES[(route, k, m)]))
                parts.append("")
                table.append((route, k, m, name))
                continue
            parts.append(f"template <bool V, bool FAST> __device__ __forceinline__ void {name}("
                         "const uint8_t *const *in, uint8_t *const *out, uint64_t off, "
                         "uint64_t len, int mode, uint32_t &bad) {")
            parts.append(emit_body(k, ops, nreg, (route, k, m) in SELECT_SWAP))
            parts.append("}")
            parts.append(f"__global__ __launch_bounds__(kBlockThreads) {_waves_attr(route, k, m)}"
                         f"void k_{name}(const Gf8Call a) {{")
            parts.append("  const uint64_t tile = uint64_t(blockIdx.x) * 8192;")
            parts.append("  const uint64_t off = tile + split_lane_offset(threadIdx.x);")
            parts.append("  if (off >= a.len) return;")
            parts.append(f"  uint8_t *const *out = const_cast<uint8_t *const *>(a.ptr + {k});")
            parts.append(_dispatch(name, "a.ptr", "a.len", "a.vec_ok != 0", "  "))
            parts.append("}")
            # verify: the same program comparing its outputs with the parity
            # shards (ReedSolomon::verify); a mismatching lane sets the flag
            parts.append(f"__global__ __launch_bounds__(kBlockThreads) {_waves_attr(route, k, m)}"
                         f"void k_{name}_verify(const Gf8Call a) {{")
            parts.append("  const uint64_t tile = uint64_t(blockIdx.x) * 8192;")
            parts.append("  const uint64_t off = tile + split_lane_offset(threadIdx.x);")
            parts.append("  if (off >= a.len) return;")
            parts.append(f"  uint8_t *const *out = const_cast<uint8_t *const *>(a.ptr + {k});")
            parts.append(_dispatch(name, "a.ptr", "a.len", "a.vec_ok != 0", "  ", verify=True))
            parts.append("  if (bad) *a.flag = 1u;")
            parts.append("}")
            parts.append("")
            table.append((route, k, m, name))
    # Persistent work-list kernels, one per generated program so that each gets
    # the register budget of its own program: tiles are (job, 8 KiB column
    # chunk), numbered job by job; tile_job[t] names the job of tile t.  The
    # full/tail choice is made once per tile, so the full path is straight-line
    # code whose loads can all be in flight together (a per-load branch would
    # put a wait after every load).
    for pid, (route, k, m, name) in enumerate(table, start=1):
        if (route, k, m) in SPLIT:  # emitted with its per-call kernels
            continue
        waves = _waves_attr(route, k, m)
        parts.append(f"__global__ __launch_bounds__(kBlockThreads) {waves}void k_batch_{name}("
                     "const BatchJob *jobs,")
        parts.append("    const uint32_t *tile_job, uint64_t total) {")
        parts.append("  for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {")
        parts.append("    // job fields are the same in every lane: scalar loads")
        parts.append("    const auto &J = ld_const_ref(jobs + ld_const(tile_job + t));")
        parts.append("    const uint64_t len = J.len;")
        parts.append("    const uint64_t tile = (t - J.tile_begin) * 8192;")
        parts.append("    const uint64_t off = tile + split_lane_offset(threadIdx.x);")
        parts.append("    if (off >= len) continue;")
        parts.append("    const uint8_t *const *in = J.ptrs;")
        parts.append(f"    uint8_t *const *out = const_cast<uint8_t *const *>(in + {k});")
        parts.append(_dispatch(name, "in", "len", "J.vec_ok != 0", "    "))
        parts.append("  }")
        parts.append("}")
        parts.append("")
    parts.append("}  // namespace")
    parts.append("")
    parts.append("// Program id of a generated (route, k, m) program for BatchJob::prog, 0 if none.")
    parts.append("uint32_t bs_program_id(int route, uint32_t k, uint32_t m) {")
    for pid, (route, k, m, name) in enumerate(table, start=1):
        rid = 1 if route == "mat" else 2
        parts.append(f"  if (route == {rid} && k == {k} && m == {m}) return {pid};")
    parts.append("  return 0;")
    parts.append("}")
    parts.append("")
    parts.append(f"uint32_t bs_program_count() {{ return {len(table)}; }}")
    parts.append("")
    parts.append("hipError_t launch_batch_bs(uint32_t prog, const BatchJob *jobs, const uint32_t *tile_job, "
                 "uint64_t total_tiles, uint32_t grid, hipStream_t stream) {")
    parts.append("  if (total_tiles == 0) return hipSuccess;")
    parts.append("  switch (prog) {")
    for pid, (route, k, m, name) in enumerate(table, start=1):
        if (route, k, m) in SPLIT:  # four wave-pair workgroups per tile
            parts.append(f"    case {pid}: hipLaunchKernelGGL(k_batch_{name}, dim3(uint32_t(std::min<uint64_t>("
                         "uint64_t(grid) * 4, total_tiles * 4))), dim3(kPairThreads), 0, stream, jobs, "
                         "tile_job, total_tiles); break;")
            continue
        parts.append(f"    case {pid}: hipLaunchKernelGGL(k_batch_{name}, dim3(grid), dim3(kBlockThreads), 0, "
                     "stream, jobs, tile_job, total_tiles); break;")
    parts.append("    default: return hipErrorNotSupported;")
    parts.append("  }")
    parts.append("  return hipGetLastError();")
    parts.append("}")
    parts.append("")
    parts.append("// route: 1 = matrix, 2 = fft.  Returns hipErrorNotSupported for shapes")
    parts.append("// without a generated program.")
    parts.append("hipError_t launch_bs_encode(int route, const Gf8Call &a, hipStream_t stream) {")
    parts.append("  if (a.len == 0) return hipSuccess;")
    parts.append("  const dim3 grid(uint32_t((a.len + 8191) / 8192));  // one 8 KiB tile per workgroup")
    for route, k, m, name in table:
        rid = 1 if route == "mat" else 2
        parts.append(f"  if (route == {rid} && a.k == {k} && a.nout == {m}) {{")
        if (route, k, m) in SPLIT:
            parts.append(f"    hipLaunchKernelGGL(k_{name}, dim3(uint32_t((a.len + 2047) / 2048)), "
                         "dim3(kPairThreads), 0, stream, a);")
        else:
            parts.append(f"    hipLaunchKernelGGL(k_{name}, grid, dim3(kBlockThreads), 0, stream, a);")
        parts.append("    return hipGetLastError();")
        parts.append("  }")
    parts.append("  return hipErrorNotSupported;")
    parts.append("}")
    parts.append("")
    parts.append("// The verify form of launch_bs_encode: a.flag is set to 1 on any mismatch.")
    parts.append("hipError_t launch_bs_verify(int route, const Gf8Call &a, hipStream_t stream) {")
    parts.append("  if (a.len == 0) return hipSuccess;")
    parts.append("  const dim3 grid(uint32_t((a.len + 8191) / 8192));")
    for route, k, m, name in table:
        rid = 1 if route == "mat" else 2
        parts.append(f"  if (route == {rid} && a.k == {k} && a.nout == {m}) {{")
        if (route, k, m) in SPLIT:
            parts.append(f"    hipLaunchKernelGGL(k_{name}_verify, dim3(uint32_t((a.len + 2047) / 2048)), "
                         "dim3(kPairThreads), 0, stream, a);")
        else:
            parts.append(f"    hipLaunchKernelGGL(k_{name}_verify, grid, dim3(kBlockThreads), 0, stream, a);")
        parts.append("    return hipGetLastError();")
        parts.append("  }")
    parts.append("  return hipErrorNotSupported;")
    parts.append("}")
    parts.append("")
    parts.append("bool bs_encode_available(int route, uint32_t k, uint32_t m) {")
    for route, k, m, name in table:
        rid = 1 if route == "mat" else 2
        parts.append(f"  if (route == {rid} && k == {k} && m == {m}) return true;")
    parts.append("  return false;")
    parts.append("}")
    parts.append("")
    parts.append("}  // namespace trs")
    return "\n".join(parts) + "\n"


if __name__ == "__main__":
    text = generate()
    with open(sys.argv[1], "w", encoding="utf-8") as f:
        f.write(text)
