"""CPU pins for the wave-pair form of the generated encode programs
(reed-solomon_amd/gf8_codegen.py: split_program).

A split program runs as two halves, one per wave: each half's IFFT ops on its
own registers, then the pair map of the two stride-k/2 levels,
a' = a ^ c (a ^ b) and b' = b ^ c (a ^ b) for the register pair
(R[j], R[j + k/2]), then each half's FFT ops and parity stores.  Replaying
exactly that, byte by byte with the oracle's GF(2^8) product, gives the CPU
oracle's generator-matrix parity (the encode of src/reedsolomon.rs:271-351).
Shapes whose programs lack the form are refused, so the generator falls back
to the one-wave kernels for them.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "reed-solomon_amd"))
import gf8_codegen  # noqa: E402


def _run_half(ops, regs):
    out = {}
    for op in ops:
        kind = op[0]
        if kind == "st":
            out[op[1]] = regs[op[2]]
        elif kind == "xo":
            regs[op[1]] ^= regs[op[2]]
        elif kind == "mx":
            regs[op[1]] ^= O.gf8_mul(op[3], regs[op[2]])
        elif kind == "mc":
            regs[op[1]] = O.gf8_mul(op[3], regs[op[2]])
        else:
            raise AssertionError(op)
    return out


@pytest.mark.parametrize("k,m", [(32, 32), (16, 16)])
def test_split_halves_give_matrix_parity(k, m):
    ops, nreg = gf8_codegen.fft_program(k, m)
    pre, consts, post = gf8_codegen.split_program(ops, nreg, k)
    h = k // 2
    assert len(consts) == h
    # each half stores its share of the parity, every parity shard once
    stores = sorted(op[1] for half in post for op in half if op[0] == "st")
    assert stores == list(range(m))
    G = O.build_matrix(8, k, k + m)
    rng = np.random.default_rng(k * 7 + m)
    data = rng.integers(0, 256, (k, 48)).astype(np.uint8)
    data[:, 0] = 0          # all-zero column: zero parity
    data[:, 1] = 0
    data[0, 1] = 1          # unit column: parity = first generator column
    want = O.rs8_encode(G, k, data)
    for col in range(data.shape[1]):
        halves = [{j: int(data[hh * h + j, col]) for j in range(h)} for hh in (0, 1)]
        for hh in (0, 1):
            _run_half(pre[hh], halves[hh])
        a, b = halves
        for j in range(h):
            u = O.gf8_mul(consts[j], a[j] ^ b[j])
            a[j] ^= u
            b[j] ^= u
        got = {}
        for hh in (0, 1):
            got.update(_run_half(post[hh], halves[hh]))
        assert [got[j] for j in range(m)] == [int(x) for x in want[:, col]], (k, m, col)


@pytest.mark.parametrize("k,m", [(7, 13), (10, 10), (14, 14), (18, 6)])
def test_split_refuses_programs_without_two_halves(k, m):
    ops, nreg = gf8_codegen.fft_program(k, m)
    with pytest.raises(ValueError):
        gf8_codegen.split_program(ops, nreg, k)


def test_split_programs_are_the_generated_wave_pair_kernels():
    src = gf8_codegen.generate()
    for route, k, m in gf8_codegen.SPLIT:
        name = f"bs_{route}_{k}_{m}"
        for kernel in (f"k_batch_{name}(", f"k_{name}(", f"k_{name}_verify("):
            head = src[:src.index(kernel)].rsplit("\n", 1)[1]
            assert "kPairThreads" in head, kernel
        assert f"void {name}(" not in src  # no one-wave body for it
