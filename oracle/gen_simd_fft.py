"""Writes the straight-line GFNI / AVX-512 encoders of the six generated LCH
shapes for the CPU baseline (oracle build step; test infrastructure).

The reference runs its generated programs (src/fft_programs.rs, expanded by
the fft_run! macro, src/macros.rs:16-46) as straight-line code over byte
strips, one vector register per program register and one gf2p8affine per
multiply (src/gf/fft_x86.rs:22-165).  This script emits the same shape of
code in C from the op lists that reed-solomon_amd/fftgen.py regenerates (the
tests pin those op lists to the reference's text): one __m512i per register,
64-byte strips, a masked last strip.

Usage: python gen_simd_fft.py OUT.c
"""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "reed-solomon_amd"))
import fftgen  # noqa: E402

MUL = fftgen.MUL


def gfni_matrix(c: int) -> int:
    """src/gf/gfni.rs:47-64: A.byte[r].bit[j] = bit_(7-r)(c * 2^j)."""
    q = 0
    for r in range(8):
        byte = 0
        for j in range(8):
            if (MUL[c][1 << j] >> (7 - r)) & 1:
                byte |= 1 << j
        q |= byte << (8 * r)
    return q


def emit(k: int, m: int) -> str:
    p = fftgen.build_encode_program(k, m)
    n = p.register_count
    name = f"fft_{k}_{m}"
    out = [f"/* ({k},{m}): {n} registers, {p.mult_count} multiplies */",
           f"static SIMD_FN void {name}(const uint8_t *const *in, uint8_t *const *out, size_t len) {{",
           "  for (size_t off = 0; off < len; off += 64) {",
           "    const __mmask64 mk = tail_mask(len - off);",
           "    __m512i " + ", ".join(f"r{i}" for i in range(n)) + ";"]
    for i in range(k):
        out.append(f"    r{i} = _mm512_maskz_loadu_epi8(mk, (const void *)(in[{i}] + off));")
    for op in p.ops:
        kind = op[0]
        if kind == "xo":
            out.append(f"    r{op[1]} = _mm512_xor_si512(r{op[1]}, r{op[2]});")
        elif kind == "cp":
            out.append(f"    r{op[1]} = r{op[2]};")
        elif kind in ("mx", "mc"):
            prod = (f"_mm512_gf2p8affine_epi64_epi8(r{op[2]}, "
                    f"_mm512_set1_epi64((long long)0x{gfni_matrix(op[3]):016x}ull), 0)")
            if kind == "mx":
                out.append(f"    r{op[1]} = _mm512_xor_si512(r{op[1]}, {prod});")
            else:
                out.append(f"    r{op[1]} = {prod};")
        elif kind == "zero":
            out.append(f"    r{op[1]} = _mm512_setzero_si512();")
        else:
            raise ValueError(kind)
    for j, r in enumerate(p.parity_regs):
        out.append(f"    _mm512_mask_storeu_epi8((void *)(out[{j}] + off), mk, r{r});")
    out.append("  }")
    out.append("}")
    return "\n".join(out)


def generate() -> str:
    shapes = sorted(fftgen.GENERATED_SHAPES)
    parts = [
        "/* GENERATED by oracle/gen_simd_fft.py -- do not edit.  CPU-baseline test",
        " * infrastructure: straight-line GFNI / AVX-512 encoders of the generated",
        " * LCH shapes (src/fft_programs.rs, src/gf/fft_x86.rs:22-165). */",
        "#include <immintrin.h>",
        "#include <stddef.h>",
        "#include <stdint.h>",
        "",
        '#define SIMD_FN __attribute__((target("gfni,avx512f,avx512bw,avx512vl")))',
        "",
        "static inline SIMD_FN __mmask64 tail_mask(size_t rem) {",
        "  return rem >= 64 ? ~(__mmask64)0 : (((__mmask64)1 << rem) - 1);",
        "}",
        "",
    ]
    for (k, m) in shapes:
        parts.append(emit(k, m))
        parts.append("")
    parts.append("/* Parity of one stripe through the generated program of (k, m); -1 when")
    parts.append(" * the shape has none. */")
    parts.append("int orc_simd_fft_generated(size_t k, size_t m, const uint8_t *const *in,")
    parts.append("                           uint8_t *const *out, size_t len) {")
    for (k, m) in shapes:
        parts.append(f"  if (k == {k} && m == {m}) {{ fft_{k}_{m}(in, out, len); return 0; }}")
    parts.append("  return -1;")
    parts.append("}")
    return "\n".join(parts) + "\n"


if __name__ == "__main__":
    with open(sys.argv[1], "w", encoding="utf-8") as f:
        f.write(generate())
